// storage.h - how the 10 moments of a node are kept in HBM.
//
// FP64/FP32 store the reference's fields as they are (rho - rho0, u, m; mlbm.cu:154-163).
// The 16-bit forms store deviations: rho - rho0, u_x - u_ref (u_ref: the case's mean inflow
// velocity), u_y, u_z and the non-equilibrium second moments m - u u, decoded in FP32. FP16 scales
// them by s_rho, s_u, s_m into its normal range; the BF16 form keeps rho - rho0 and m - u u in BF16
// unscaled (float's exponent range; power-of-two scales would not change a single rounding) and
// the velocity in FP16 bits times 16.
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
    R uref;    // reference velocity (x): the 16-bit forms store u_x - uref
};

template <class S>
struct Conv;

template <>
struct Conv<double> {
    static constexpr bool scaled = false;
    static constexpr bool unit_scale = true;
    __device__ __host__ static double to(double v) { return v; }
    template <class R>
    __device__ __host__ static double from(R v) { return double(v); }
};

template <>
struct Conv<float> {
    static constexpr bool scaled = false;
    static constexpr bool unit_scale = true;
    __device__ __host__ static float to(float v) { return v; }
    template <class R>
    __device__ __host__ static float from(R v) { return float(v); }
};

// Two half values, the operand form of the half conversion with stochastic rounding.
typedef __attribute__((__vector_size__(2 * sizeof(_Float16)))) _Float16 f16x2_t;

// 16-bit forms round stochastically when given random bits: a uniform integer is added below
// the kept mantissa bits before truncation, so the stored value is unbiased. Round-to-nearest
// would drop every per-step increment smaller than half an ulp and stall slowly evolving flow.
// BF16 takes 16 random bits (the low half of `r`, the rest zero), all the bits it drops.
// The BF16 form keeps the velocity in FP16 bits (u_x - u_ref, u_y, u_z, times kVelocityScale):
// BF16's 8-bit significand made the velocity's rounding noise the dominant storage error (drag
// error on 16 seeded 64-192 x 64-96 x 32-96 channels, mean / max over the cases after 10,000
// steps: 0.92% / 1.73% with a BF16 velocity, 0.12% / 0.28% with an FP16 one); density and
// non-equilibrium second moments stay BF16.
template <>
struct Conv<__hip_bfloat16> {
    static constexpr bool scaled = true;
    static constexpr bool unit_scale = true;
    static constexpr bool half_velocity = true;
    __device__ static float to(__hip_bfloat16 v) { return __bfloat162float(v); }
    __device__ static __hip_bfloat16 from(float v) { return __float2bfloat16(v); }
    __device__ static __hip_bfloat16 from_sr(float v, uint32_t r16) {
        const uint32_t b = __float_as_uint(v) + r16;
        return __ushort_as_bfloat16(static_cast<unsigned short>(b >> 16));
    }
};

template <>
struct Conv<_Float16> {
    static constexpr bool scaled = true;
    static constexpr bool unit_scale = false;
    static constexpr bool half_velocity = false;
    __device__ static float to(_Float16 v) { return float(v); }
    __device__ static _Float16 from(float v) { return _Float16(v); }
    // v_cvt_sr_f16_f32 adds the high 13 bits of its random operand below the kept mantissa bits
    // and truncates (13 float mantissa bits fall below the half mantissa in the normal range):
    // `r` is a whole random word
    __device__ static _Float16 from_sr(float v, uint32_t r) {
        f16x2_t w = {};
        w = __builtin_amdgcn_cvt_sr_f16_f32(w, v, r, false);
        return w[0];
    }
};

// Random bits for stochastic rounding: one 32-bit hash of (node, step seed) per node, from which
// every field takes a window (overlapping windows inside the word: the rounding of each field is
// unbiased, correlations between the fields of one node do not matter).
__device__ __forceinline__ uint32_t sr_hash(uint32_t idx, uint32_t seed) {
    uint32_t h = (idx ^ 0x61C88647u) * 0x9E3779B1u + seed;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}


// Field scale class: 0 for rho, 1 for u, 2 for the second moments.
__device__ __host__ constexpr int field_class(int f) { return f == 0 ? 0 : (f < 4 ? 1 : 2); }

// One moment buffer, node-chunked: the nodes are cut into chunks of kChunk<S> consecutive indices
// (32, or 16 for FP64: one tile row of the default 32- and 16-wide tiles; a row of the 16-wide
// FP32 tiles of cache-resident grids is half a chunk), and the fields of a node are grouped so
// that a node moves in few wide accesses. FP32 and 16-bit storage group fields 0-3, 4-7 and 8-9:
// chunk c holds group g of its nodes at element c * 10 kChunk + g * 4 kChunk + width * lane, with
// width 4, 4, 2 (3 accesses per node: 16 + 16 + 8 bytes in FP32, 8 + 8 + 4 at 16-bit storage).
// FP64 pairs the fields (2j, 2j+1) at element c * 10 kChunk + j * 2 kChunk + 2 lane (5 accesses of
// 16 bytes). A chunk is one contiguous run (1280 bytes in FP32 and FP64, 640 bytes at 16-bit
// storage), so the loads (stores) of a tile row fall into consecutive lines.
template <class S>
constexpr uint32_t kChunk = sizeof(S) == 8 ? 16u : 32u;

// Pair word of two stored values: the bits of each value and the word moved as one access.
template <size_t B>
struct PairBits;
template <>
struct PairBits<2> {
    using U = uint16_t;
    using W = uint32_t;
};
template <>
struct PairBits<4> {
    using U = uint32_t;
    using W = uint64_t;
};
template <>
struct PairBits<8> {
    using U = uint64_t;
    typedef uint64_t W __attribute__((ext_vector_type(2)));
};
template <class S>
using pair_t = typename PairBits<sizeof(S)>::W;

template <class S>
__device__ __host__ inline void split_pair(pair_t<S> w, S &lo, S &hi) {
    using U = typename PairBits<sizeof(S)>::U;
    if constexpr (sizeof(S) == 8) {
        lo = __builtin_bit_cast(S, U(w.x));
        hi = __builtin_bit_cast(S, U(w.y));
    } else {
        lo = __builtin_bit_cast(S, static_cast<U>(w));
        hi = __builtin_bit_cast(S, static_cast<U>(w >> (8 * sizeof(S))));
    }
}

template <class S>
__device__ __host__ inline pair_t<S> join_pair(S lo, S hi) {
    using U = typename PairBits<sizeof(S)>::U;
    if constexpr (sizeof(S) == 8) {
        pair_t<S> w;
        w.x = __builtin_bit_cast(U, lo);
        w.y = __builtin_bit_cast(U, hi);
        return w;
    } else {
        return pair_t<S>(__builtin_bit_cast(U, lo)) | (pair_t<S>(__builtin_bit_cast(U, hi)) << (8 * sizeof(S)));
    }
}

// Quad word of four stored values (FP32 and 16-bit storage).
template <size_t B>
struct QuadBits;
template <>
struct QuadBits<2> {
    using U = uint16_t;
    using W = uint64_t;
};
template <>
struct QuadBits<4> {
    using U = uint32_t;
    typedef uint32_t W __attribute__((ext_vector_type(4)));
};
template <class S>
using quad_t = typename QuadBits<sizeof(S)>::W;

template <class S>
__device__ __host__ inline void split_quad(quad_t<S> w, S *v) {
    using U = typename QuadBits<sizeof(S)>::U;
    if constexpr (sizeof(S) == 4) {
        v[0] = __builtin_bit_cast(S, U(w.x));
        v[1] = __builtin_bit_cast(S, U(w.y));
        v[2] = __builtin_bit_cast(S, U(w.z));
        v[3] = __builtin_bit_cast(S, U(w.w));
    } else {  // two 32-bit halves, two values each
        const uint32_t lo = uint32_t(w), hi = uint32_t(w >> 32);
        v[0] = __builtin_bit_cast(S, static_cast<U>(lo));
        v[1] = __builtin_bit_cast(S, static_cast<U>(lo >> 16));
        v[2] = __builtin_bit_cast(S, static_cast<U>(hi));
        v[3] = __builtin_bit_cast(S, static_cast<U>(hi >> 16));
    }
}

template <class S>
__device__ __host__ inline quad_t<S> join_quad(const S *v) {
    using U = typename QuadBits<sizeof(S)>::U;
    if constexpr (sizeof(S) == 4) {
        quad_t<S> w;
        w.x = __builtin_bit_cast(U, v[0]);
        w.y = __builtin_bit_cast(U, v[1]);
        w.z = __builtin_bit_cast(U, v[2]);
        w.w = __builtin_bit_cast(U, v[3]);
        return w;
    } else {  // two 32-bit halves, two values each
        const uint32_t lo = uint32_t(__builtin_bit_cast(U, v[0])) | (uint32_t(__builtin_bit_cast(U, v[1])) << 16);
        const uint32_t hi = uint32_t(__builtin_bit_cast(U, v[2])) | (uint32_t(__builtin_bit_cast(U, v[3])) << 16);
        return quad_t<S>(lo) | (quad_t<S>(hi) << 32);
    }
}

template <class S>
struct FieldSet {
    S *base;
};

// Element offset of the chunk that holds node idx.
template <class S>
__host__ __device__ inline size_t chunk_base(uint32_t idx) {
    return size_t(idx / kChunk<S>) * (10 * kChunk<S>);
}

// Four fields per access (FP32 and 16-bit storage) or two (FP64).
template <class S>
constexpr bool kQuads = sizeof(S) <= 4;

// Element offset of field f of node idx.
template <class S>
__host__ __device__ inline size_t field_elem(uint32_t idx, int f) {
    const uint32_t l = idx % kChunk<S>;
    if constexpr (kQuads<S>) {
        const uint32_t g = uint32_t(f) >> 2;
        return chunk_base<S>(idx) + g * (4 * kChunk<S>) + (g < 2 ? 4u : 2u) * l + uint32_t(f & 3);
    } else {
        return chunk_base<S>(idx) + size_t(f >> 1) * (2 * kChunk<S>) + 2 * l + uint32_t(f & 1);
    }
}

// Pointer to pair word 0 of node idx (paired layout); pair word j is at [j kChunk].
template <class S>
__host__ __device__ inline pair_t<S> *pair_ptr(S *base, uint32_t idx) {
    return reinterpret_cast<pair_t<S> *>(base + chunk_base<S>(idx)) + idx % kChunk<S>;
}
template <class S>
__host__ __device__ inline const pair_t<S> *pair_ptr(const S *base, uint32_t idx) {
    return reinterpret_cast<const pair_t<S> *>(base + chunk_base<S>(idx)) + idx % kChunk<S>;
}

// Elements of a buffer of n nodes plus the padding slot at index n.
template <class S>
inline size_t field_buffer_elems(int64_t n) {
    return size_t((n + 1 + kChunk<S> - 1) / kChunk<S>) * (10 * kChunk<S>);
}

template <class S>
__host__ __device__ inline FieldSet<S> make_fields(S *base) {
    return FieldSet<S>{base};
}

template <class S>
__host__ __device__ inline FieldSet<const S> as_const(const FieldSet<S> &fs) {
    return FieldSet<const S>{fs.base};
}

template <class S>
__device__ __forceinline__ S ld_field(const FieldSet<const S> &fs, int f, uint32_t idx) {
    return fs.base[field_elem<S>(idx, f)];
}

template <class S>
__device__ __forceinline__ void st_field(const FieldSet<S> &fs, int f, uint32_t idx, S v) {
    fs.base[field_elem<S>(idx, f)] = v;
}

// The 10 stored values of node `idx`, untouched: decoding is left to the consumer so that a
// prefetch does not wait for its data.
template <class S>
__device__ __forceinline__ void load_raw(const FieldSet<const S> &fs, uint32_t idx, S *raw) {
    if constexpr (kQuads<S>) {
        const S *cb = fs.base + chunk_base<S>(idx);
        const uint32_t l = idx % kChunk<S>;
        const quad_t<S> *__restrict__ q = reinterpret_cast<const quad_t<S> *>(cb) + l;
        const pair_t<S> *__restrict__ p = reinterpret_cast<const pair_t<S> *>(cb + 8 * kChunk<S>) + l;
        split_quad<S>(q[0], raw);
        split_quad<S>(q[kChunk<S>], raw + 4);
        split_pair<S>(p[0], raw[8], raw[9]);
    } else {
        const pair_t<S> *__restrict__ p = pair_ptr(fs.base, idx);
#pragma unroll
        for (int j = 0; j < 5; ++j) split_pair<S>(p[j * kChunk<S>], raw[2 * j], raw[2 * j + 1]);
    }
}

// Multiplier of the velocity kept in FP16 bits by the BF16 form (a power of two: exact).
constexpr float kVelocityScale = 16.0f;

// Whether the 16-bit form S keeps its velocity in FP16 bits.
template <class S>
__host__ __device__ constexpr bool has_half_velocity() {
    if constexpr (Conv<S>::scaled) {
        return Conv<S>::half_velocity;
    } else {
        return false;
    }
}

// Random operand of field f's stochastic rounding from the node's hash h: a whole word for the
// FP16 conversions (its high 13 bits are used; the word rotated by 3 f, or 7 f for the BF16 form's
// velocity), 16 bits for a BF16 field (the halves of h and of h rotated by 5 and 10 bits, so a
// field costs one rotation per two fields and an add).
template <class S>
__device__ __forceinline__ uint32_t sr_rand(uint32_t h, int f) {
    if constexpr (has_half_velocity<S>()) {
        if (f >= 1 && f <= 3) return __builtin_rotateright32(h, 7 * f);
        const int k = f == 0 ? 0 : f - 3;  // BF16 fields 0, 4..9 -> 0..6
        const uint32_t w = k < 2 ? h : __builtin_rotateright32(h, 5 * (k / 2));
        return (k & 1) ? (w >> 16) : (w & 0xFFFFu);
    } else {
        return __builtin_rotateright32(h, 3 * f);
    }
}

// Value of stored field f before the deviation is undone (scaled back for FP16).
template <class R, class S>
__device__ __forceinline__ R stored_value(S raw, int f, const StoreScale<R> &sc) {
    if constexpr (has_half_velocity<S>()) {
        if (f >= 1 && f <= 3) {
            const _Float16 h = __builtin_bit_cast(_Float16, __builtin_bit_cast(uint16_t, raw));
            return R(float(h)) * R(1.0f / kVelocityScale);
        }
    }
    R v = R(Conv<S>::to(raw));
    if constexpr (!Conv<S>::unit_scale) v *= sc.dec[field_class(f)];
    return v;
}

// Stored bits of deviation v of field f (v already scaled for FP16 storage); r: random operand of
// the stochastic rounding (sr, from sr_rand), unused with round-to-nearest.
template <class R, class S>
__device__ __forceinline__ S stored_bits(R v, int f, uint32_t r, bool sr) {
    if constexpr (has_half_velocity<S>()) {
        if (f >= 1 && f <= 3) {
            const float hv = float(v) * kVelocityScale;
            const _Float16 h = sr ? Conv<_Float16>::from_sr(hv, r) : _Float16(hv);
            return __builtin_bit_cast(S, __builtin_bit_cast(uint16_t, h));
        }
    }
    return sr ? Conv<S>::from_sr(v, r) : Conv<S>::from(v);
}

// The stored words of one node as loaded -- three accesses (FP32 and 16-bit storage: fields 0-3,
// 4-7, 8-9) or five pairs (FP64) -- kept packed until the node is decoded. A 16-bit node then
// occupies five 32-bit registers instead of ten, which the marches' prefetched planes need.
template <class S, bool Q = kQuads<S>>
struct RawNode;
template <class S>
struct RawNode<S, true> {
    quad_t<S> q0, q1;
    pair_t<S> p;
};
template <class S>
struct RawNode<S, false> {
    pair_t<S> p[5];
};

template <class S>
__device__ __forceinline__ void load_raw(const FieldSet<const S> &fs, uint32_t idx, RawNode<S> &r) {
    if constexpr (kQuads<S>) {
        const S *cb = fs.base + chunk_base<S>(idx);
        const uint32_t l = idx % kChunk<S>;
        const quad_t<S> *__restrict__ q = reinterpret_cast<const quad_t<S> *>(cb) + l;
        const pair_t<S> *__restrict__ p = reinterpret_cast<const pair_t<S> *>(cb + 8 * kChunk<S>) + l;
        r.q0 = q[0];
        r.q1 = q[kChunk<S>];
        r.p = p[0];
    } else {
        const pair_t<S> *__restrict__ p = pair_ptr(fs.base, idx);
#pragma unroll
        for (int j = 0; j < 5; ++j) r.p[j] = p[j * kChunk<S>];
    }
}

template <class S>
__device__ __forceinline__ void unpack_raw(const RawNode<S> &r, S *v) {
    if constexpr (kQuads<S>) {
        split_quad<S>(r.q0, v);
        split_quad<S>(r.q1, v + 4);
        split_pair<S>(r.p, v[8], v[9]);
    } else {
#pragma unroll
        for (int j = 0; j < 5; ++j) split_pair<S>(r.p[j], v[2 * j], v[2 * j + 1]);
    }
}

// Moments from stored values (M[0] absolute density).
template <class R, class S>
__device__ __forceinline__ void decode_node(const S *raw, const StoreScale<R> &sc, R rho0, R *M) {
    if constexpr (!Conv<S>::scaled) {
#pragma unroll
        for (int f = 0; f < 10; ++f) M[f] = R(Conv<S>::to(raw[f]));
        M[0] += rho0;
    } else {
#pragma unroll
        for (int f = 0; f < 10; ++f) {
            M[f] = stored_value<R, S>(raw[f], f, sc);
        }
        M[0] += rho0;
        M[1] += sc.uref;
        M[4] += M[1] * M[1];
        M[5] += M[2] * M[2];
        M[6] += M[3] * M[3];
        M[7] += M[1] * M[2];
        M[8] += M[1] * M[3];
        M[9] += M[2] * M[3];
    }
}

template <class R, class S>
__device__ __forceinline__ void decode_node(const RawNode<S> &r, const StoreScale<R> &sc, R rho0, R *M) {
    S raw[10];
    unpack_raw<S>(r, raw);
    decode_node<R, S>(raw, sc, rho0, M);
}

// Decode the moments of node `idx` (M[0] absolute density).
template <class R, class S>
__device__ __forceinline__ void load_node(const FieldSet<const S> &fs, uint32_t idx, const StoreScale<R> &sc, R rho0,
                                          R *M) {
    S raw[10];
    load_raw(fs, idx, raw);
    decode_node<R, S>(raw, sc, rho0, M);
}

// The 10 stored values of the moments M of node `idx`. For the 16-bit forms `seed` selects
// stochastic rounding (seed != 0) or round-to-nearest (seed == 0), and NEQ says that M[4..9]
// already hold the non-equilibrium second moments m - u u.
template <class R, class S, bool NEQ = false>
__device__ __forceinline__ void encode_node(uint32_t idx, const StoreScale<R> &sc, R rho0, const R *M, uint32_t seed,
                                            S *out) {
    if constexpr (!Conv<S>::scaled) {
        out[0] = Conv<S>::from(M[0] - rho0);
#pragma unroll
        for (int f = 1; f < 10; ++f) out[f] = Conv<S>::from(M[f]);
    } else {
        R v[10];
        v[0] = M[0] - rho0;
        v[1] = M[1] - sc.uref;
        v[2] = M[2];
        v[3] = M[3];
        v[4] = NEQ ? M[4] : M[4] - M[1] * M[1];
        v[5] = NEQ ? M[5] : M[5] - M[2] * M[2];
        v[6] = NEQ ? M[6] : M[6] - M[3] * M[3];
        v[7] = NEQ ? M[7] : M[7] - M[1] * M[2];
        v[8] = NEQ ? M[8] : M[8] - M[1] * M[3];
        v[9] = NEQ ? M[9] : M[9] - M[2] * M[3];
        if constexpr (!Conv<S>::unit_scale) {
#pragma unroll
            for (int f = 0; f < 10; ++f) v[f] *= sc.enc[field_class(f)];
        }
        if (seed == 0) {
#pragma unroll
            for (int f = 0; f < 10; ++f) out[f] = stored_bits<R, S>(v[f], f, 0u, false);
        } else {
            const uint32_t h = sr_hash(idx, seed);
#pragma unroll
            for (int f = 0; f < 10; ++f) out[f] = stored_bits<R, S>(v[f], f, sr_rand<S>(h, f), true);
        }
    }
}

// Store 10 stored values of node `idx` (the layout of load_raw).
template <class S>
__device__ __forceinline__ void store_raw(const FieldSet<S> &fs, uint32_t idx, const S *v) {
    if constexpr (kQuads<S>) {
        S *cb = fs.base + chunk_base<S>(idx);
        const uint32_t l = idx % kChunk<S>;
        quad_t<S> *__restrict__ q = reinterpret_cast<quad_t<S> *>(cb) + l;
        pair_t<S> *__restrict__ p = reinterpret_cast<pair_t<S> *>(cb + 8 * kChunk<S>) + l;
#ifdef MRLBM_NT_STORES
        __builtin_nontemporal_store(join_quad<S>(v), q);
        __builtin_nontemporal_store(join_quad<S>(v + 4), q + kChunk<S>);
        __builtin_nontemporal_store(join_pair<S>(v[8], v[9]), p);
#else
        q[0] = join_quad<S>(v);
        q[kChunk<S>] = join_quad<S>(v + 4);
        p[0] = join_pair<S>(v[8], v[9]);
#endif
    } else {
        pair_t<S> *__restrict__ p = pair_ptr(fs.base, idx);
#pragma unroll
        for (int j = 0; j < 5; ++j) p[j * kChunk<S>] = join_pair<S>(v[2 * j], v[2 * j + 1]);
    }
}

template <class S>
__device__ __forceinline__ void store_raw(const FieldSet<S> &fs, uint32_t idx, const RawNode<S> &r) {
    S v[10];
    unpack_raw<S>(r, v);
    store_raw<S>(fs, idx, v);
}

// Store the moments of node `idx`.
template <class R, class S>
__device__ __forceinline__ void store_node(const FieldSet<S> &fs, uint32_t idx, const StoreScale<R> &sc, R rho0,
                                           const R *M, uint32_t seed = 0) {
    S v[10];
    encode_node<R, S>(idx, sc, rho0, M, seed, v);
    store_raw<S>(fs, idx, v);
}

// Planes that hold a whole number of chunks (nx ny a multiple of kChunk) put node `off` of every
// plane at the same byte offsets from the plane's first element, so a march forms them once:
// .a for the quads (fields 0-3, and 4-7 at .a + 4 kChunk values) or FP64's first pair (pair j at
// .a + 2 j kChunk values), .b for the 8-9 pair of the quad layout.
struct LaneOffsets {
    uint32_t a, b;
};

template <class S>
__host__ __device__ inline LaneOffsets lane_offsets(uint32_t off) {
    const uint32_t cb = (off / kChunk<S>) * (10u * kChunk<S>), l = off % kChunk<S>;
    if constexpr (kQuads<S>)
        return {uint32_t((cb + 4u * l) * sizeof(S)), uint32_t((cb + 8u * kChunk<S> + 2u * l) * sizeof(S))};
    else
        return {uint32_t((cb + 2u * l) * sizeof(S)), 0u};
}

// Addresses of a node's stored words: .a the first quad (FP32 and 16-bit storage; the second is
// 4 kChunk values on) or the first pair (FP64; pair j is 2 j kChunk values on), .b the 8-9 pair of
// the quad layout. From the node index, or from a plane of whole chunks and the node's offsets in it.
struct NodePtrs {
    const char *a, *b;
};

template <class S>
__device__ __forceinline__ NodePtrs node_ptrs(const S *base, uint32_t idx) {
    const char *cb = reinterpret_cast<const char *>(base + chunk_base<S>(idx));
    const uint32_t l = idx % kChunk<S>;
    if constexpr (kQuads<S>)
        return {cb + size_t(4 * l) * sizeof(S), cb + size_t(8 * kChunk<S> + 2 * l) * sizeof(S)};
    else
        return {cb + size_t(2 * l) * sizeof(S), nullptr};
}

template <class S>
__device__ __forceinline__ NodePtrs node_ptrs_at(const S *plane, LaneOffsets o) {
    const char *b = reinterpret_cast<const char *>(plane);
    return {b + o.a, b + o.b};
}

template <class S>
__device__ __forceinline__ void load_raw_p(NodePtrs p, RawNode<S> &r) {
    if constexpr (kQuads<S>) {
        r.q0 = *reinterpret_cast<const quad_t<S> *>(p.a);
        r.q1 = *reinterpret_cast<const quad_t<S> *>(p.a + 4 * kChunk<S> * sizeof(S));
        r.p = *reinterpret_cast<const pair_t<S> *>(p.b);
    } else {
#pragma unroll
        for (int j = 0; j < 5; ++j) r.p[j] = *reinterpret_cast<const pair_t<S> *>(p.a + size_t(j) * 2 * kChunk<S> * sizeof(S));
    }
}

template <class S>
__device__ __forceinline__ void store_raw_p(NodePtrs p, const S *v) {
    char *a = const_cast<char *>(p.a);
    if constexpr (kQuads<S>) {
        *reinterpret_cast<quad_t<S> *>(a) = join_quad<S>(v);
        *reinterpret_cast<quad_t<S> *>(a + 4 * kChunk<S> * sizeof(S)) = join_quad<S>(v + 4);
        *reinterpret_cast<pair_t<S> *>(const_cast<char *>(p.b)) = join_pair<S>(v[8], v[9]);
    } else {
#pragma unroll
        for (int j = 0; j < 5; ++j)
            *reinterpret_cast<pair_t<S> *>(a + size_t(j) * 2 * kChunk<S> * sizeof(S)) = join_pair<S>(v[2 * j], v[2 * j + 1]);
    }
}

template <class S>
__device__ __forceinline__ void store_raw_p(NodePtrs p, const RawNode<S> &r) {
    S v[10];
    unpack_raw<S>(r, v);
    store_raw_p<S>(p, v);
}

// Node-major buffers (the 10 values of node i at base[10 i .. 10 i + 9]): the donor snapshot,
// read at scattered stencil corners, where one node's values then share a cache line. A node
// starts on a pair-word boundary (10 i values) and moves as 5 pair words.
template <class S>
__device__ __forceinline__ void store_raw_nm(S *__restrict__ base, uint32_t idx, const S *v) {
    pair_t<S> *__restrict__ p = reinterpret_cast<pair_t<S> *>(base + size_t(idx) * 10);
#pragma unroll
    for (int j = 0; j < 5; ++j) p[j] = join_pair<S>(v[2 * j], v[2 * j + 1]);
}

template <class R, class S>
__device__ __forceinline__ void load_node_nm(const S *__restrict__ base, uint32_t idx, const StoreScale<R> &sc, R rho0,
                                             R *M) {
    S raw[10];
    const pair_t<S> *__restrict__ p = reinterpret_cast<const pair_t<S> *>(base + size_t(idx) * 10);
#pragma unroll
    for (int j = 0; j < 5; ++j) split_pair<S>(p[j], raw[2 * j], raw[2 * j + 1]);
    decode_node<R, S>(raw, sc, rho0, M);
}

}  // namespace mrlbm
