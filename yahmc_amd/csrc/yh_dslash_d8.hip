// Instantiation of the hopping-term and clover kernels for D = 8 (SU(3)
// adjoint).  The multi-RHS hop is not instantiated for D = 8.
#include "yh_dslash_impl.hpp"

namespace yh {
template int hop_dispatch_d<8>(Ctx*, const HopArgs&, int, int, int);
template int clover_apply_d<8>(Ctx*, int, int, double, Field*, const Field*, int);
}  // namespace yh
