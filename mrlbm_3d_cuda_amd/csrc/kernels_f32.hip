// kernels_f32.hip - FP32-arithmetic instantiations: FP32 storage and the 16-bit storage forms.
#include "kernels_impl.h"

namespace mrlbm {
MRLBM_INSTANTIATE(float, float)
MRLBM_INSTANTIATE(float, __hip_bfloat16)
MRLBM_INSTANTIATE(float, _Float16)
}  // namespace mrlbm
