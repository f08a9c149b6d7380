// FP64-geometry instantiation of the tracer kernels, compiled with
// -ffp-contract=off so that each operation rounds as the FP64 oracle's does.
#include "trace_impl.h"
namespace pach {
PACH_INSTANTIATE(double)
}
