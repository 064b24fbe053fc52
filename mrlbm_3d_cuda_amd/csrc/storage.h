// storage.h - how the 10 moments of a node are kept in HBM.
//
// FP64/FP32 store the reference's fields as they are (rho - rho0, u, m; mlbm.cu:154-163).
// The 16-bit forms store scaled deviations: (rho - rho0) * s_rho, u * s_u and the
// non-equilibrium second moments (m - u u) * s_m, decoded in FP32.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mrlbm {

template <class R>
struct StoreScale {
    R enc[3];  // rho, u, m multipliers (16-bit storage only)
    R dec[3];  // their inverses
};

template <class S>
struct Conv;

template <>
struct Conv<double> {
    static constexpr bool scaled = false;
    __device__ __host__ static double to(double v) { return v; }
    template <class R>
    __device__ __host__ static double from(R v) { return double(v); }
};

template <>
struct Conv<float> {
    static constexpr bool scaled = false;
    __device__ __host__ static float to(float v) { return v; }
    template <class R>
    __device__ __host__ static float from(R v) { return float(v); }
};

// 16-bit forms round stochastically when given random bits: a uniform integer is added below
// the kept mantissa bits before truncation, so the stored value is unbiased. Round-to-nearest
// would drop every per-step increment smaller than half an ulp and stall slowly evolving flow.
template <>
struct Conv<__hip_bfloat16> {
    static constexpr bool scaled = true;
    __device__ static float to(__hip_bfloat16 v) { return __bfloat162float(v); }
    __device__ static __hip_bfloat16 from(float v) { return __float2bfloat16(v); }
    __device__ static __hip_bfloat16 from_sr(float v, uint32_t r) {
        const uint32_t b = __float_as_uint(v) + (r & 0xFFFFu);
        return __ushort_as_bfloat16(static_cast<unsigned short>(b >> 16));
    }
};

template <>
struct Conv<_Float16> {
    static constexpr bool scaled = true;
    __device__ static float to(_Float16 v) { return float(v); }
    __device__ static _Float16 from(float v) { return _Float16(v); }
    __device__ static _Float16 from_sr(float v, uint32_t r) {
        // 13 float mantissa bits fall below the half mantissa (normal range)
        const float t = __uint_as_float(__float_as_uint(v) + (r & 0x1FFFu));
        return _Float16(__float2half_rz(t));
    }
};

// Random bits for stochastic rounding: a 32-bit integer hash of (node, field pair, step seed).
__device__ __forceinline__ uint32_t sr_hash(uint32_t idx, uint32_t k, uint32_t seed) {
    uint32_t h = idx * 0x9E3779B1u ^ (seed + k * 0x85EBCA77u);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// Field scale class: 0 for rho, 1 for u, 2 for the second moments.
__device__ __host__ constexpr int field_class(int f) { return f == 0 ? 0 : (f < 4 ? 1 : 2); }

// Decode the moments of node `idx` (M[0] absolute density).
template <class R, class S>
__device__ __forceinline__ void load_node(const S *__restrict__ base, size_t stride, uint32_t idx,
                                          const StoreScale<R> &sc, R rho0, R *M) {
    if constexpr (!Conv<S>::scaled) {
#pragma unroll
        for (int f = 0; f < 10; ++f) M[f] = R(Conv<S>::to(base[f * stride + idx]));
        M[0] += rho0;
    } else {
#pragma unroll
        for (int f = 0; f < 10; ++f) M[f] = R(Conv<S>::to(base[f * stride + idx])) * sc.dec[field_class(f)];
        M[0] += rho0;
        M[4] += M[1] * M[1];
        M[5] += M[2] * M[2];
        M[6] += M[3] * M[3];
        M[7] += M[1] * M[2];
        M[8] += M[1] * M[3];
        M[9] += M[2] * M[3];
    }
}

// Store the moments of node `idx`. For the 16-bit forms `seed` selects stochastic rounding
// (seed != 0) or round-to-nearest (seed == 0).
template <class R, class S>
__device__ __forceinline__ void store_node(S *__restrict__ base, size_t stride, uint32_t idx, const StoreScale<R> &sc,
                                           R rho0, const R *M, uint32_t seed = 0) {
    if constexpr (!Conv<S>::scaled) {
        base[idx] = Conv<S>::from(M[0] - rho0);
#pragma unroll
        for (int f = 1; f < 10; ++f) base[f * stride + idx] = Conv<S>::from(M[f]);
    } else {
        R v[10];
        v[0] = M[0] - rho0;
        v[1] = M[1];
        v[2] = M[2];
        v[3] = M[3];
        v[4] = M[4] - M[1] * M[1];
        v[5] = M[5] - M[2] * M[2];
        v[6] = M[6] - M[3] * M[3];
        v[7] = M[7] - M[1] * M[2];
        v[8] = M[8] - M[1] * M[3];
        v[9] = M[9] - M[2] * M[3];
        if (seed == 0) {
#pragma unroll
            for (int f = 0; f < 10; ++f) base[f * stride + idx] = Conv<S>::from(v[f] * sc.enc[field_class(f)]);
        } else {
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const uint32_t h = sr_hash(idx, uint32_t(k), seed);
                base[(2 * k) * stride + idx] = Conv<S>::from_sr(v[2 * k] * sc.enc[field_class(2 * k)], h);
                base[(2 * k + 1) * stride + idx] =
                    Conv<S>::from_sr(v[2 * k + 1] * sc.enc[field_class(2 * k + 1)], h >> 16 | h << 16);
            }
        }
    }
}

}  // namespace mrlbm
