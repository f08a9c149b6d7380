// FP32-geometry instantiation of the tracer kernels (the throughput path).
#include "trace_impl.h"
namespace pach {
PACH_INSTANTIATE(float)
}
