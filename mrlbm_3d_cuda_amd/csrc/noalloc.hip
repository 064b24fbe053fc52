// noalloc.hip - mrlbm_get_moments_range: a moment read-back that allocates nothing.
//
// mrlbm_get_moments decodes all ten fields of every node into a device staging array it allocates
// (hipMalloc / hipFree on every call). Callers that may not allocate device memory -- code that
// runs under HIP graph capture, tight monitoring loops, harnesses that forbid hipMalloc -- read a
// range of nodes of one field through a fixed staging array instead: `scratch` below, one copy per
// GPU, held until the process ends. The caller passes the host output; ranges longer than
// MRLBM_RANGE_MAX_VALUES are refused.
//
// Every call on one GPU shares that GPU's array and leaves its last contents there; each call
// overwrites exactly the values it then copies, so nothing needs clearing. Calls that overlap in
// time on one GPU (several host threads, or several simulations stepped from different threads)
// race on the array and must be serialised by the caller.
#include <mutex>

#include "sim.h"

using mrlbm::set_error;

namespace mrlbm {

namespace {

// Staging array of the range read-back (8 B per value).
static __device__ __align__(256) double scratch[MRLBM_RANGE_MAX_VALUES];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on each GPU, looked up once per device with that device current.
// In host code the array's name is only a stand-in: kernels and copies get this address instead.
int scratch_address(int dev, double **out) {
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    if (dev < 0 || dev >= kMaxDevices) return set_error(MRLBM_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        int prev = 0;
        HIP_TRY(hipGetDevice(&prev));
        if (prev != dev) HIP_TRY(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != dev) HIP_TRY(hipSetDevice(prev));
        if (e != hipSuccess)
            return set_error(MRLBM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[dev] = static_cast<double *>(p);
    }
    *out = addr[dev];
    return MRLBM_OK;
}

// Field f of nodes first .. first + count - 1 as double (field 0 = rho - rho0, the stored
// deviation), with the arithmetic of the full read-back's decode.
template <class R, class S>
__global__ __launch_bounds__(256) void decode_range_kernel(FieldSet<const S> in, int f, uint32_t first,
                                                           uint32_t count, double *__restrict__ out,
                                                           StoreScale<R> sc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t node = first + i;
    if constexpr (!Conv<S>::scaled) {
        out[i] = double(Conv<S>::to(ld_field(in, f, node)));
    } else if (f == 0) {
        out[i] = double(R(Conv<S>::to(ld_field(in, 0, node))) * sc.dec[0]);
    } else {
        R M[10];
        load_node<R, S>(in, node, sc, R(0), M);
        out[i] = double(M[f]);
    }
}

template <class R, class S>
int read_range(SimBase &base, int f, uint32_t first, uint32_t count, double *staging, double *host) {
    Sim<R, S> &s = static_cast<Sim<R, S> &>(base);
    if (int rc = s.wait_batch()) return rc;
    hipLaunchKernelGGL((decode_range_kernel<R, S>), dim3((count + 255) / 256), dim3(256), 0, s.stream(), s.state(),
                       f, first, count, staging, s.store_scale());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host, staging, size_t(count) * sizeof(double), hipMemcpyDeviceToHost, s.stream()));
    HIP_TRY(hipStreamSynchronize(s.stream()));
    return MRLBM_OK;
}

}  // namespace

}  // namespace mrlbm

extern "C" {

int mrlbm_get_moments_range(mrlbm_sim *s, int32_t field, int64_t first, int64_t count, double *host) {
    if (!s || !s->impl) return set_error(MRLBM_ERR_INVALID, "null simulation");
    mrlbm::SimBase &sim = *s->impl;
    if (!sim.device_ready) return set_error(MRLBM_ERR_INVALID, "geometry-only simulation");
    if (field < 0 || field >= MRLBM_NUM_FIELDS) return set_error(MRLBM_ERR_INVALID, "bad field");
    if (first < 0 || count < 0 || first > sim.g.n || count > sim.g.n - first)
        return set_error(MRLBM_ERR_INVALID, "node range outside the grid");
    if (count > MRLBM_RANGE_MAX_VALUES)
        return set_error(MRLBM_ERR_INVALID, "range of " + std::to_string(count) + " nodes exceeds the staging array (" +
                                                std::to_string(MRLBM_RANGE_MAX_VALUES) +
                                                " values): read it in several ranges");
    if (count == 0) return MRLBM_OK;
    if (!host) return set_error(MRLBM_ERR_INVALID, "null output");
    double *staging = nullptr;
    if (int rc = mrlbm::scratch_address(sim.device, &staging)) return rc;
    const uint32_t f0 = uint32_t(first), n = uint32_t(count);
    switch (sim.d.precision) {
    case MRLBM_FP64: return mrlbm::read_range<double, double>(sim, field, f0, n, staging, host);
    case MRLBM_FP32: return mrlbm::read_range<float, float>(sim, field, f0, n, staging, host);
    case MRLBM_FP32_BF16: return mrlbm::read_range<float, __hip_bfloat16>(sim, field, f0, n, staging, host);
    case MRLBM_FP32_FP16: return mrlbm::read_range<float, _Float16>(sim, field, f0, n, staging, host);
    default: return set_error(MRLBM_ERR_INVALID, "unknown precision");
    }
}

}  // extern "C"
