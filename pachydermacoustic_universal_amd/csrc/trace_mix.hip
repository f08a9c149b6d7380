// Mixed-walk instantiation of the octave-ray kernel: FP64 rays, hits and
// shading (the FP64 path's results), with an FP32 screen for the voxel walk.
// Compiled with -ffp-contract=off so every FP64 operation rounds as the
// oracle's does; the FP32 screen fuses its products explicitly.
#define PACH_MIX_UNIT
#include "trace_impl.h"
