// FP32 and FP16-storage lattice kernels (FP32 arithmetic).  Contraction into
// fused multiply-add is allowed here; these variants are judged by their
// error against the FP64 lattice, not bit for bit.
#include <utility>

#include "kernels_impl.h"

namespace d3q19 {
template struct Launch<1>;
template struct Launch<2>;
}
