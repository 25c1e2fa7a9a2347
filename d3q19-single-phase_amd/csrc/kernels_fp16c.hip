// FP16C-storage lattice kernels (a 16-bit format with 11 significand bits,
// FP32 arithmetic).  Contraction into fused multiply-add is allowed here, as
// for FP32 and FP16 storage: these variants are held to error budgets against
// the FP64 lattice, not to bit equality.
#include <utility>

#include "kernels_impl.h"

namespace d3q19 {
template struct Launch<3>;
}
