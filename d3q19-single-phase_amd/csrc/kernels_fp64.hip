// FP64 lattice kernels.  Built with -ffp-contract=off: no fused multiply-add
// is formed, so every operation rounds as in the reference.
#include <utility>

#include "kernels_impl.h"

namespace d3q19 {
template struct Launch<0>;
}
