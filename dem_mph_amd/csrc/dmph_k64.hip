// FP64 instantiation of the step kernels (parity mode).  Built with
// -ffp-contract=off so that no multiply-add is fused and the FP64 arithmetic
// associates exactly as the reference's source does.
#include "dmph_kernels.inc"

namespace dmph {
DMPH_INSTANTIATE(double)
}
