// FP32 instantiation of the step kernels: pair arithmetic in FP32 on
// cell-relative positions, FP64 rigid-body reductions and time accumulator.
#include "dmph_kernels.inc"
#include "dmph_sweep32.inc"

namespace dmph {
DMPH_INSTANTIATE(float)
}
