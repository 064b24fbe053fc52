// kernels_f64.hip - FP64 instantiation of the step kernels. Built with -ffp-contract=off so that
// every FP64 expression is evaluated as written (see kExactOrder in node_math.h).
#include "kernels_impl.h"

namespace mrlbm {
MRLBM_INSTANTIATE(double, double)
}  // namespace mrlbm
