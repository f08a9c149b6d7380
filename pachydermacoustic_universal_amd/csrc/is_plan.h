// is_plan.h -- range planning of the image-source launches (host only, no HIP).
//
// pach_image_sources validates the items of one order in ranges [lo, hi) of
// whole sequences (multiples of R items).  Each launch writes its valid paths
// to a device buffer of `cap` records; a launch that finds more is redone:
// over a smaller range when the range holds more than one sequence, or, for a
// single sequence (R items), once more after the buffer has grown to hold
// every path it found.  Ranges that fill less than a quarter of the buffer
// double the next span (up to span_max).
#pragma once

#include <algorithm>
#include <cstdint>

namespace pach {

struct IsSpanStep {
    bool redo;         // launch [lo, hi) again
    int64_t span;      // items of the next launch
    uint64_t grow_to;  // > 0: grow the result buffer to this many records first
};

inline IsSpanStep is_after_launch(int64_t span, int64_t lo, int64_t hi, uint64_t n, uint64_t cap, int64_t R,
                                  int64_t span_max)
{
    IsSpanStep s{false, span, 0};
    if (n > cap) {
        s.redo = true;
        if (hi - lo > R) {
            const int64_t fit = (int64_t)((double)(hi - lo) * (double)cap / (2.0 * (double)n));
            s.span = std::max<int64_t>(R, fit / R * R);
        } else {
            s.grow_to = n;  // one sequence over every receiver: make room for all of it
        }
        return s;
    }
    if (n < cap / 4) s.span = std::min<int64_t>(span * 2, span_max);
    return s;
}

}  // namespace pach
