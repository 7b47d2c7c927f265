// Instantiation of the hopping-term and clover kernels for D = 3.
#include "yh_dslash_impl.hpp"

namespace yh {
template int hop_dispatch_d<3>(Ctx*, const HopArgs&, int, int, int);
template int clover_apply_d<3>(Ctx*, int, int, double, Field*, const Field*, int);
template int hop_mr_dispatch_d<3>(Ctx*, const HopArgsM&, int, int, int);
}  // namespace yh
