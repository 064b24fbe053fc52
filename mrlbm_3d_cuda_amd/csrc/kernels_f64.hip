// kernels_f64.hip - FP64 instantiation of the step kernels, and the force-record launch. Built
// with -ffp-contract=off so that every FP64 expression is evaluated as written (see kExactOrder in
// node_math.h).
#include "kernels_impl.h"

namespace mrlbm {
MRLBM_INSTANTIATE(double, double)

void launch_record(const RecordArgs &r, hipStream_t stream) {
    hipLaunchKernelGGL(record_kernel, dim3(1), dim3(64), 0, stream, r);
}
}  // namespace mrlbm
