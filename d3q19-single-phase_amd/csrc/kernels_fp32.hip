// FP32 and FP16-storage lattice kernels (FP32 arithmetic).  Contraction into
// fused multiply-add is allowed here; these variants are held to error
// budgets against the FP64 lattice, not to bit equality.
#include <utility>

#include "kernels_impl.h"

namespace d3q19 {
template struct Launch<1>;
template struct Launch<2>;
}
