// node_math.h - per-node D3Q27 moment-space arithmetic for the HIP kernels.
//
// Moment vector M[10] = (rho, ux, uy, uz, mxx, myy, mzz, mxy, mxz, myz) with rho absolute and
// m_ab = Pi_ab / rho - cs2 delta_ab (common/globalStructs.cuh:6-19,
// core/D3Q27/eval_moments.cuh:16-21). R is the arithmetic type (double or float).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lattice.h"

namespace mrlbm {

// FP64 arithmetic follows one fixed expression order per quantity (the order of the plain
// per-direction loops the reference writes), and the FP64 kernels are compiled without FMA
// contraction, so FP64 results are reproducible bit for bit by any IEEE implementation of the
// same loops. FP32 and the 16-bit storage forms use the faster pairwise forms below.
template <class R>
constexpr bool kExactOrder = std::is_same<R, double>::value;

// Population q of the regularised distribution of M, per-direction loop form
// (core/D3Q27/second_order/reconstruction.cuh:5-60 as a Hermite sum).
template <class R>
__device__ __forceinline__ R population(const R *M, int q) {
    const R cx = R(LATTICE.cx[q]), cy = R(LATTICE.cy[q]), cz = R(LATTICE.cz[q]);
    const R third = R(1) / R(3);
    const R cu = cx * M[1] + cy * M[2] + cz * M[3];
    const R hd = M[4] * (cx * cx - third) + M[5] * (cy * cy - third) + M[6] * (cz * cz - third);
    const R ho = M[7] * cx * cy + M[8] * cx * cz + M[9] * cy * cz;
    return R(weight(q)) * M[0] * (R(1) + R(3) * cu + R(4.5) * hd + R(9) * ho);
}

// 1/x: the hardware reciprocal (about 1 ulp) in FP32, the IEEE quotient in FP64.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }

// ---------------------------------------------------------------- mask helpers

// Bit set of directions satisfying a predicate on (cx, cy, cz).
template <class Pred>
constexpr uint32_t dir_mask(Pred p) {
    uint32_t m = 0;
    for (int q = 0; q < Q; ++q)
        if (p(LATTICE.cx[q], LATTICE.cy[q], LATTICE.cz[q])) m |= 1u << q;
    return m;
}

// Sum of f over the set bits of M, in increasing q.
template <uint32_t M, class R>
__device__ __forceinline__ R msum(const R *f) {
    R s = R(0);
    bool first = true;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if (M & (1u << q)) {
            s = first ? f[q] : s + f[q];
            first = false;
        }
    }
    return s;
}

// Sequential signed sum in increasing q: + over K & P, - over K & N (first term keeps its sign).
template <uint32_t K, uint32_t P, uint32_t N, class R>
__device__ __forceinline__ R mseq(const R *f) {
    R s = R(0);
    bool first = true;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const bool pos = (K & P) & (1u << q), neg = (K & N) & (1u << q);
        if (!pos && !neg) continue;
        s = first ? (pos ? f[q] : -f[q]) : (pos ? s + f[q] : s - f[q]);
        first = false;
    }
    return s;
}

// ---------------------------------------------------------------- regularised populations

// All 27 populations of the second-order regularised distribution of M
// (core/D3Q27/second_order/reconstruction.cuh:5-60), evaluated pairwise: for each pair of
// opposite directions f = w (S +- V) with S the even (rho, m) part and V = 3 rho c.u.
template <class R>
__device__ __forceinline__ void reconstruct(const R *M, R *f) {
    if constexpr (kExactOrder<R>) {
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = population(M, q);
        return;
    }
    const R rho = M[0];
    const R P = rho * (R(1) - R(1.5) * (M[4] + M[5] + M[6]));
    const R r3 = R(3) * rho, r45 = R(4.5) * rho, r9 = R(9) * rho;
    const R Ux = r3 * M[1], Uy = r3 * M[2], Uz = r3 * M[3];
    const R Dx = r45 * M[4], Dy = r45 * M[5], Dz = r45 * M[6];
    const R Oxy = r9 * M[7], Oxz = r9 * M[8], Oyz = r9 * M[9];
    const R w0 = R(W_REST), w1 = R(W_AXIS), w2 = R(W_FACE), w3 = R(W_CORNER);
    f[0] = w0 * P;
    auto pair = [&](int q, R w, R S, R V) {
        f[q] = w * (S + V);
        f[q + 1] = w * (S - V);
    };
    pair(1, w1, P + Dx, Ux);
    pair(3, w1, P + Dy, Uy);
    pair(5, w1, P + Dz, Uz);
    const R Pxy = P + Dx + Dy, Pxz = P + Dx + Dz, Pyz = P + Dy + Dz, Pxyz = Pxy + Dz;
    pair(7, w2, Pxy + Oxy, Ux + Uy);
    pair(9, w2, Pxz + Oxz, Ux + Uz);
    pair(11, w2, Pyz + Oyz, Uy + Uz);
    pair(13, w2, Pxy - Oxy, Ux - Uy);
    pair(15, w2, Pxz - Oxz, Ux - Uz);
    pair(17, w2, Pyz - Oyz, Uy - Uz);
    pair(19, w3, Pxyz + Oxy + Oxz + Oyz, Ux + Uy + Uz);
    pair(21, w3, Pxyz + Oxy - Oxz - Oyz, Ux + Uy - Uz);
    pair(23, w3, Pxyz - Oxy + Oxz - Oyz, Ux - Uy + Uz);
    pair(25, w3, Pxyz - Oxy - Oxz + Oyz, Uy + Uz - Ux);
}

// Moments of 27 populations (core/D3Q27/eval_moments.cuh:5-22), from the sums s and
// differences d of opposite pairs.
template <class R>
__device__ __forceinline__ void moments(const R *f, R *M) {
    if constexpr (kExactOrder<R>) {
        R rho = 0, jx = 0, jy = 0, jz = 0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int cx = LATTICE.cx[q], cy = LATTICE.cy[q], cz = LATTICE.cz[q];
            rho += f[q];
            jx += f[q] * R(cx);
            jy += f[q] * R(cy);
            jz += f[q] * R(cz);
            pxx += f[q] * R(cx * cx);
            pyy += f[q] * R(cy * cy);
            pzz += f[q] * R(cz * cz);
            pxy += f[q] * R(cx * cy);
            pxz += f[q] * R(cx * cz);
            pyz += f[q] * R(cy * cz);
        }
        const R inv = R(1) / rho;
        const R cs2 = R(1) / R(3);
        M[0] = rho;
        M[1] = jx * inv;
        M[2] = jy * inv;
        M[3] = jz * inv;
        M[4] = pxx * inv - cs2;
        M[5] = pyy * inv - cs2;
        M[6] = pzz * inv - cs2;
        M[7] = pxy * inv;
        M[8] = pxz * inv;
        M[9] = pyz * inv;
        return;
    }
    R s[13], d[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        s[k] = f[2 * k + 1] + f[2 * k + 2];
        d[k] = f[2 * k + 1] - f[2 * k + 2];
    }
    // pair k: 0 x, 1 y, 2 z, 3 (1,1,0), 4 (1,0,1), 5 (0,1,1), 6 (1,-1,0), 7 (1,0,-1),
    // 8 (0,1,-1), 9 (1,1,1), 10 (1,1,-1), 11 (1,-1,1), 12 (-1,1,1)
    const R corners = (s[9] + s[10]) + (s[11] + s[12]);
    const R rho = f[0] + ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])) + (s[8] + corners);
    const R jx = (d[0] + d[3]) + (d[4] + d[6]) + d[7] + (d[9] + d[10]) + (d[11] - d[12]);
    const R jy = (d[1] + d[3]) + (d[5] - d[6]) + d[8] + (d[9] + d[10]) + (d[12] - d[11]);
    const R jz = (d[2] + d[4]) + (d[5] - d[7]) - d[8] + (d[9] - d[10]) + (d[11] + d[12]);
    const R pxx = (s[0] + s[3]) + (s[4] + s[6]) + s[7] + corners;
    const R pyy = (s[1] + s[3]) + (s[5] + s[6]) + s[8] + corners;
    const R pzz = (s[2] + s[4]) + (s[5] + s[7]) + s[8] + corners;
    const R pxy = (s[3] - s[6]) + (s[9] + s[10]) - (s[11] + s[12]);
    const R pxz = (s[4] - s[7]) + (s[9] + s[11]) - (s[10] + s[12]);
    const R pyz = (s[5] - s[8]) + (s[9] + s[12]) - (s[10] + s[11]);
    const R inv = fast_rcp(rho);
    const R cs2 = R(1) / R(3);
    M[0] = rho;
    M[1] = jx * inv;
    M[2] = jy * inv;
    M[3] = jz * inv;
    M[4] = pxx * inv - cs2;
    M[5] = pyy * inv - cs2;
    M[6] = pzz * inv - cs2;
    M[7] = pxy * inv;
    M[8] = pxz * inv;
    M[9] = pyz * inv;
}

// ---------------------------------------------------------------- populations as held in LDS
// The fused kernel streams populations through LDS. In FP64 it keeps the weighted populations
// f_q, evaluated exactly as reconstruct() does. In FP32 it keeps g_q = f_q / w_q: the weight
// products are dropped from the reconstruction and applied once per lattice class (rest, axis,
// face diagonal, corner) when the moments are summed.

// Direction q of velocity (cx, cy, cz), indexed [cx + 1][cy + 1][cz + 1].
struct QIndex {
    int q[3][3][3];
};
constexpr QIndex make_qindex() {
    QIndex t{};
    for (int q = 0; q < Q; ++q) t.q[LATTICE.cx[q] + 1][LATTICE.cy[q] + 1][LATTICE.cz[q] + 1] = q;
    return t;
}
constexpr QIndex QOF = make_qindex();

// Regularised populations in the LDS form, g = f / w (core/D3Q27/second_order/reconstruction.cuh:
// 5-60 as a Hermite sum). FP32: for fixed (cy, cz), g is a quadratic in cx, g = A + B cx + Dx cx^2,
// and A (a quadratic in cy and cz) and B (linear in cy and cz) are built the same way, so the 27
// values take 3 operations per x line after 22 for the coefficients.
template <class R>
__device__ __forceinline__ void reconstruct_lds(const R *M, R *g) {
    if constexpr (kExactOrder<R>) {
        reconstruct(M, g);
    } else {
        const R rho = M[0];
        const R P = rho * (R(1) - R(1.5) * (M[4] + M[5] + M[6]));
        const R r3 = R(3) * rho, r45 = R(4.5) * rho, r9 = R(9) * rho;
        const R Ux = r3 * M[1], Uy = r3 * M[2], Uz = r3 * M[3];
        const R Dx = r45 * M[4], Dy = r45 * M[5], Dz = r45 * M[6];
        const R Oxy = r9 * M[7], Oxz = r9 * M[8], Oyz = r9 * M[9];
        // A(cy, cz) = A0(cz) + cy Byz(cz) + cy^2 Dy; B(cy, cz) = Bx(cz) + cy Oxy
        const R Pz = P + Dz;
        const R A0[3] = {Pz - Uz, P, Pz + Uz};
        const R Byz[3] = {Uy - Oyz, Uy, Uy + Oyz};
        const R Bx[3] = {Ux - Oxz, Ux, Ux + Oxz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const R Ay = A0[c] + Dy;
            const R A[3] = {Ay - Byz[c], A0[c], Ay + Byz[c]};
            const R B[3] = {Bx[c] - Oxy, Bx[c], Bx[c] + Oxy};
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const R t = A[b] + Dx;
                g[QOF.q[0][b][c]] = t - B[b];
                g[QOF.q[1][b][c]] = A[b];
                g[QOF.q[2][b][c]] = t + B[b];
            }
        }
    }
}

// The nine populations in the LDS form (FP32 arithmetic) with one in-plane velocity component
// fixed, c_a = s (a = y when `along_y`, else x; s = +-1), the ones a ring position of a tile hands
// to the tile. g[3 k + j] has c_b = j - 1 (b the other in-plane axis) and c_z = k - 1. With c_a
// fixed the Hermite sum of reconstruct_lds is a quadratic in c_b and c_z:
// g = K + Lz cz + Dz cz^2 + (Lb + Obz cz) cb + Db cb^2, K = P + Da + s Ua, Lb = Ub + s Oab,
// Lz = Uz + s Oaz.
template <class R>
__device__ __forceinline__ void reconstruct_ring(const R *M, bool along_y, R s, R *g) {
    const R rho = M[0];
    const R P = rho * (R(1) - R(1.5) * (M[4] + M[5] + M[6]));
    const R r3 = R(3) * rho, r45 = R(4.5) * rho, r9 = R(9) * rho;
    const R sr3 = s * r3, sr9 = s * r9;
    const R ua = along_y ? M[2] : M[1], ub = along_y ? M[1] : M[2];
    const R maa = along_y ? M[5] : M[4], mbb = along_y ? M[4] : M[5];
    const R maz = along_y ? M[9] : M[8], mbz = along_y ? M[8] : M[9];
    const R K = P + r45 * maa + sr3 * ua;
    const R Lb = r3 * ub + sr9 * M[7];
    const R Lz = r3 * M[3] + sr9 * maz;
    const R Dz = r45 * M[6], Db = r45 * mbb, Obz = r9 * mbz;
    const R Kz = K + Dz;
    const R A[3] = {Kz - Lz, K, Kz + Lz};
    const R B[3] = {Lb - Obz, Lb, Lb + Obz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R t = A[k] + Db;
        g[3 * k + 0] = t - B[k];
        g[3 * k + 1] = A[k];
        g[3 * k + 2] = t + B[k];
    }
}

// Moments of the populations held in LDS (eval_moments.cuh:5-22). FP32: the D3Q27 weights are
// products of the 1-D weights w1(0) = 2/3, w1(+-1) = 1/6, so the weighted sums factor into three
// passes of 1-D sums (x, then y, then z), each of which keeps, per line, the zeroth, first and
// second moment times 6: (g- + g+) + 4 g0, g+ - g-, g+ + g-. Pruned to the ten moments needed
// (order <= 2 in every pair of axes) that is 66 additions; the common factor 6^3 = 216 cancels in
// u and m and is divided out of rho.
template <class R>
__device__ __forceinline__ void moments_lds(const R *g, R *M) {
    if constexpr (kExactOrder<R>) {
        moments(g, M);
    } else {
        // x pass: per (cy, cz) line
        R X0[3][3], X1[3][3], X2[3][3];
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const R gm = g[QOF.q[0][b][c]], g0 = g[QOF.q[1][b][c]], gp = g[QOF.q[2][b][c]];
                X2[b][c] = gp + gm;
                X1[b][c] = gp - gm;
                X0[b][c] = R(4) * g0 + X2[b][c];
            }
        // y pass: per cz, the (x, y) orders 00, 01, 02, 10, 11, 20
        R Y00[3], Y01[3], Y02[3], Y10[3], Y11[3], Y20[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Y02[c] = X0[2][c] + X0[0][c];
            Y00[c] = R(4) * X0[1][c] + Y02[c];
            Y01[c] = X0[2][c] - X0[0][c];
            Y10[c] = R(4) * X1[1][c] + (X1[2][c] + X1[0][c]);
            Y11[c] = X1[2][c] - X1[0][c];
            Y20[c] = R(4) * X2[1][c] + (X2[2][c] + X2[0][c]);
        }
        // z pass: the ten moments times 216
        const R pzz = Y00[2] + Y00[0];
        const R rho = R(4) * Y00[1] + pzz;
        const R jz = Y00[2] - Y00[0];
        const R jy = R(4) * Y01[1] + (Y01[2] + Y01[0]);
        const R pyz = Y01[2] - Y01[0];
        const R pyy = R(4) * Y02[1] + (Y02[2] + Y02[0]);
        const R jx = R(4) * Y10[1] + (Y10[2] + Y10[0]);
        const R pxz = Y10[2] - Y10[0];
        const R pxy = R(4) * Y11[1] + (Y11[2] + Y11[0]);
        const R pxx = R(4) * Y20[1] + (Y20[2] + Y20[0]);
        const R inv = fast_rcp(rho);
        const R cs2 = R(1) / R(3);
        M[0] = rho * R(1.0 / 216.0);
        M[1] = jx * inv;
        M[2] = jy * inv;
        M[3] = jz * inv;
        M[4] = pxx * inv - cs2;
        M[5] = pyy * inv - cs2;
        M[6] = pzz * inv - cs2;
        M[7] = pxy * inv;
        M[8] = pxz * inv;
        M[9] = pyz * inv;
    }
}

// LDS form -> weighted populations f_q, in place (no-op in FP64).
template <class R>
__device__ __forceinline__ void lds_to_pop(R *g) {
    if constexpr (!kExactOrder<R>) {
#pragma unroll
        for (int q = 0; q < Q; ++q) g[q] *= R(weight(q));
    }
}

// BGK relaxation of the second moments towards u u (moment_collision, solver/mlbm.cu:3-16).
template <class R>
__device__ __forceinline__ void collide(R omega, R *M) {
    const R om1 = R(1) - omega;
    M[4] = om1 * M[4] + omega * M[1] * M[1];
    M[5] = om1 * M[5] + omega * M[2] * M[2];
    M[6] = om1 * M[6] + omega * M[3] * M[3];
    M[7] = om1 * M[7] + omega * M[1] * M[2];
    M[8] = om1 * M[8] + omega * M[1] * M[3];
    M[9] = om1 * M[9] + omega * M[2] * M[3];
}

// ---------------------------------------------------------------- domain-face closed forms
// cases/cylinder/boundaries.cuh:201-843. The known ("incoming") populations of a face are those
// whose velocity does not point out of the domain through that face; every moment below is a
// sum over them. `old`/`old_xm1` are the node's and its x-1 neighbour's moments at the start of
// the step.

namespace face {
constexpr uint32_t K_W = dir_mask([](int cx, int, int) { return cx <= 0; });
constexpr uint32_t K_E = dir_mask([](int cx, int, int) { return cx >= 0; });
constexpr uint32_t K_N = dir_mask([](int, int cy, int) { return cy >= 0; });
constexpr uint32_t K_S = dir_mask([](int, int cy, int) { return cy <= 0; });
constexpr uint32_t K_F = dir_mask([](int, int, int cz) { return cz >= 0; });
constexpr uint32_t K_B = dir_mask([](int, int, int cz) { return cz <= 0; });
constexpr uint32_t XY_P = dir_mask([](int cx, int cy, int) { return cx * cy > 0; });
constexpr uint32_t XY_N = dir_mask([](int cx, int cy, int) { return cx * cy < 0; });
constexpr uint32_t XZ_P = dir_mask([](int cx, int, int cz) { return cx * cz > 0; });
constexpr uint32_t XZ_N = dir_mask([](int cx, int, int cz) { return cx * cz < 0; });
constexpr uint32_t YZ_P = dir_mask([](int, int cy, int cz) { return cy * cz > 0; });
constexpr uint32_t YZ_N = dir_mask([](int, int cy, int cz) { return cy * cz < 0; });
constexpr uint32_t YY = dir_mask([](int, int cy, int) { return cy != 0; });
constexpr uint32_t ZZ = dir_mask([](int, int, int cz) { return cz != 0; });
}  // namespace face

// sum over known & P minus sum over known & N
template <uint32_t K, uint32_t P, uint32_t N, class R>
__device__ __forceinline__ R mdiff(const R *f) {
    return msum<K & P>(f) - msum<K & N>(f);
}

// Convective outlet blend (update_neumaan_density_velocity, boundaries.cuh:163-199 with
// CONVECTIVE_OUTLET): Uc = |u_x(x-1)|, phi = (1 - Uc) phi_b + Uc phi_{x-1}.
template <class R>
__device__ __forceinline__ void outlet_blend(const R *old, const R *old_xm1, R *M) {
    const R Uc = sqrt(old_xm1[1] * old_xm1[1]);
    const R a = R(1) - Uc;
#pragma unroll
    for (int k = 0; k < 4; ++k) M[k] = a * old[k] + Uc * old_xm1[k];
}

template <class R>
__device__ __forceinline__ void bc_west(const R *f, const R *old, R u_in, R *M) {
    using namespace face;
    const R rI = msum<K_W>(f);
    const R inv = R(1) / rI;
    const R mxyI = mdiff<K_W, XY_P, XY_N>(f) * inv;
    const R mxzI = mdiff<K_W, XZ_P, XZ_N>(f) * inv;
    const R uo = old[1];  // previous-step u_x of the node (boundaries.cuh:216)
    const R rho = -R(6) * rI / (-R(5) + R(3) * uo + R(3) * uo * uo);
    M[0] = rho;
    M[1] = u_in;
    M[2] = R(0);
    M[3] = R(0);
    M[4] = u_in * u_in;
    M[5] = R(0);
    M[6] = R(0);
    M[7] = R(2) * mxyI * rI / rho;
    M[8] = R(2) * mxzI * rI / rho;
    M[9] = R(0);
}

template <class R>
__device__ __forceinline__ void bc_east(const R *f, const R *old, const R *old_xm1, R *M) {
    using namespace face;
    const R rI = msum<K_E>(f);
    const R inv = R(1) / rI;
    const R cs2 = R(1) / R(3);
    const R myyI = msum<K_E & YY>(f) * inv - cs2;
    const R mzzI = msum<K_E & ZZ>(f) * inv - cs2;
    const R mxyI = mdiff<K_E, XY_P, XY_N>(f) * inv;
    const R mxzI = mdiff<K_E, XZ_P, XZ_N>(f) * inv;
    const R myzI = mseq<K_E, YZ_P, YZ_N>(f) * inv;  // boundaries.cuh:241 order
    outlet_blend(old, old_xm1, M);
    const R rho = M[0], uy = M[2], uz = M[3];
    M[4] = M[1] * M[1];
    M[5] = (R(6) * myyI * rI - R(6) * mzzI * rI + R(5) * uy * uy * rho + R(5) * uz * uz * rho) / (R(10) * rho);
    M[6] = (-R(6) * myyI * rI + R(6) * mzzI * rI + R(5) * uy * uy * rho + R(5) * uz * uz * rho) / (R(10) * rho);
    M[7] = (R(6) * mxyI * rI - uy * rho) / (R(3) * rho);
    M[8] = (R(6) * mxzI * rI - uz * rho) / (R(3) * rho);
    M[9] = R(6) * myzI * rI / (R(5) * rho);
}

// no-slip y walls (NORTH/SOUTH, boundaries.cuh:272-315)
template <uint32_t K, class R>
__device__ __forceinline__ void bc_ywall(const R *f, R *M) {
    using namespace face;
    const R rI = msum<K>(f);
    const R inv = R(1) / rI;
    const R mxyI = mdiff<K, XY_P, XY_N>(f) * inv;
    const R myzI = mdiff<K, YZ_P, YZ_N>(f) * inv;
    const R rho = R(6) * rI / R(5);
    M[0] = rho;
    M[1] = M[2] = M[3] = R(0);
    M[4] = M[5] = M[6] = R(0);
    M[7] = R(2) * mxyI * rI / rho;
    M[8] = R(0);
    M[9] = R(2) * myzI * rI / rho;
}

// no-slip z walls (BACK/FRONT, boundaries.cuh:316-359), without z periodicity
template <uint32_t K, class R>
__device__ __forceinline__ void bc_zwall(const R *f, R *M) {
    using namespace face;
    const R rI = msum<K>(f);
    const R inv = R(1) / rI;
    const R mxzI = mdiff<K, XZ_P, XZ_N>(f) * inv;
    const R myzI = mdiff<K, YZ_P, YZ_N>(f) * inv;
    const R rho = R(6) * rI / R(5);
    M[0] = rho;
    M[1] = M[2] = M[3] = R(0);
    M[4] = M[5] = M[6] = R(0);
    M[7] = R(0);
    M[8] = R(2) * mxzI * rI / rho;
    M[9] = R(2) * myzI * rI / rho;
}

// Inlet edges. With the y walls (NORTH_WEST sgn = -1, SOUTH_WEST sgn = +1,
// boundaries.cuh:360-403) the edge shear is m_xy (SHEAR = 7); with the z walls (WEST_FRONT
// sgn = -1, WEST_BACK sgn = +1, :404-447) it is m_xz (SHEAR = 8). sgn is the sign of the
// product of the two face-normal velocity components of the known diagonal pair.
template <uint32_t K, int SGN, int SHEAR = 7, class R>
__device__ __forceinline__ void bc_west_edge(const R *f, R omega, R *M) {
    using namespace face;
    constexpr uint32_t P = SHEAR == 7 ? XY_P : XZ_P, N = SHEAR == 7 ? XY_N : XZ_N;
    const R rI = msum<K>(f);
    const R inv = R(1) / rI;
    const R msI = (SGN > 0 ? msum<K & P>(f) : -msum<K & N>(f)) * inv;
    const R myzI = mseq<K, YZ_P, YZ_N>(f) * inv;
    R rho;
    if (SGN > 0)
        rho = R(36) * rI * (R(1) - msI + msI * omega) / (R(24) + omega);
    else
        rho = R(36) * rI * (R(1) + msI - msI * omega) / (R(24) + omega);
    M[0] = rho;
    M[1] = M[2] = M[3] = R(0);
    M[4] = M[5] = M[6] = R(0);
    M[7] = M[8] = R(0);
    M[SHEAR] = (R(36) * msI * rI - R(SGN) * rho) / (R(9) * rho);
    M[9] = R(12) * myzI * rI / (R(5) * rho);
}

// Outlet edges: with the y walls (NORTH_EAST sgn = +1, SOUTH_EAST sgn = -1,
// boundaries.cuh:448-511, SHEAR = 7) and the z walls (EAST_FRONT sgn = +1, EAST_BACK sgn = -1,
// :512-577, SHEAR = 8); rho and u from the outlet blend.
template <uint32_t K, int SGN, int SHEAR = 7, class R>
__device__ __forceinline__ void bc_east_edge(const R *f, const R *old, const R *old_xm1, R *M) {
    using namespace face;
    constexpr uint32_t P = SHEAR == 7 ? XY_P : XZ_P, N = SHEAR == 7 ? XY_N : XZ_N;
    const R rI = msum<K>(f);
    const R inv = R(1) / rI;
    const R msI = (SGN > 0 ? msum<K & P>(f) : -msum<K & N>(f)) * inv;
    const R myzI = mseq<K, YZ_P, YZ_N>(f) * inv;
    outlet_blend(old, old_xm1, M);
    const R rho = M[0];
    M[4] = M[5] = M[6] = R(0);
    M[7] = M[8] = R(0);
    M[SHEAR] = (R(36) * msI * rI - R(SGN) * rho) / (R(9) * rho);
    M[9] = (R(12) * myzI * rI) / (R(5) * rho);
}

// y-wall / z-wall edges (NORTH_FRONT and SOUTH_BACK sgn = +1, NORTH_BACK and SOUTH_FRONT
// sgn = -1, boundaries.cuh:578-661): only m_yz survives.
template <uint32_t K, int SGN, class R>
__device__ __forceinline__ void bc_yz_edge(const R *f, R omega, R *M) {
    using namespace face;
    const R rI = msum<K>(f);
    const R myzI = (SGN > 0 ? msum<K & YZ_P>(f) : -msum<K & YZ_N>(f)) * (R(1) / rI);
    R rho;
    if (SGN > 0)
        rho = R(36) * rI * (R(1) - myzI + myzI * omega) / (R(24) + omega);
    else
        rho = R(36) * rI * (R(1) + myzI - myzI * omega) / (R(24) + omega);
    M[0] = rho;
    M[1] = M[2] = M[3] = R(0);
    M[4] = M[5] = M[6] = R(0);
    M[7] = M[8] = R(0);
    M[9] = (R(36) * myzI * rI - R(SGN) * rho) / (R(9) * rho);
}

// Inlet corners (boundaries.cuh:662-733): rho = 216 rho_I / 125 at rest; corner 0..3 = NWF,
// NWB, SWF, SWB.
template <class R>
__device__ __forceinline__ void bc_west_corner(uint32_t corner, const R *f, R *M) {
    using namespace face;
    R rI;
    if (corner == 0)
        rI = msum<K_W & K_N & K_F>(f);
    else if (corner == 1)
        rI = msum<K_W & K_N & K_B>(f);
    else if (corner == 2)
        rI = msum<K_W & K_S & K_F>(f);
    else
        rI = msum<K_W & K_S & K_B>(f);
    M[0] = R(216) * rI / R(125);
#pragma unroll
    for (int k = 1; k < 10; ++k) M[k] = R(0);
}

// Outlet corners (boundaries.cuh:734-841): outlet blend, second moments zero.
template <class R>
__device__ __forceinline__ void bc_east_corner(const R *old, const R *old_xm1, R *M) {
    outlet_blend(old, old_xm1, M);
#pragma unroll
    for (int k = 4; k < 10; ++k) M[k] = R(0);
}

template <int... Qs>
constexpr uint32_t qbits() {
    return ((1u << Qs) | ...);
}

// BCSOLID end-wall nodes next to a curved wall (bcsolid_boundary_condition,
// cases/cylinder/boundaries.cuh:1039-1214): known set K per tag, m_xz and m_yz from signed
// sums over K, then one of four closed-form families (tags 7/224, 11/208, 13/176, 14/112).
template <uint32_t K, int FAM, class R>
__device__ __forceinline__ void bc_solid_tag(R omega, const R *f, R *M) {
    using namespace face;
    const R rI = msum<K>(f);
    const R inv = R(1.0) / rI;
    const R mxzI = mseq<K, XZ_P, XZ_N>(f) * inv;
    const R myzI = mseq<K, YZ_P, YZ_N>(f) * inv;
    const R om = omega;
    R rho, mxz, myz;
    if constexpr (FAM == 0) {
        rho = (R(108) * (R(10) + mxzI * (R(-1) + om) + myzI * (R(-1) + om)) * rI) / (R(874) + om);
        mxz = (R(198) * mxzI * rI + R(18) * myzI * rI - rho) / (R(90.) * rho);
        myz = (R(18) * mxzI * rI + R(198) * myzI * rI - rho) / (R(90.) * rho);
    } else if constexpr (FAM == 1) {
        rho = (R(108) * (R(10) + mxzI + myzI * (R(-1) + om) - mxzI * om) * rI) / (R(874) + om);
        mxz = (R(198) * mxzI * rI - R(18) * myzI * rI + rho) / (R(90.) * rho);
        myz = -(R(18) * mxzI * rI - R(198) * myzI * rI + rho) / (R(90.) * rho);
    } else if constexpr (FAM == 2) {
        rho = (R(108) * (R(10) + myzI + mxzI * (R(-1) + om) - myzI * om) * rI) / (R(874) + om);
        mxz = -(R(-198) * mxzI * rI + R(18) * myzI * rI + rho) / (R(90.) * rho);
        myz = (R(-18) * mxzI * rI + R(198) * myzI * rI + rho) / (R(90.) * rho);
    } else {
        rho = (R(-108) * (R(-10) + mxzI * (R(-1) + om) + myzI * (R(-1) + om)) * rI) / (R(874) + om);
        mxz = (R(198) * mxzI * rI + R(18) * myzI * rI + rho) / (R(90.) * rho);
        myz = (R(18) * mxzI * rI + R(198) * myzI * rI + rho) / (R(90.) * rho);
    }
    M[0] = rho;
#pragma unroll
    for (int k = 1; k < 8; ++k) M[k] = R(0);
    M[8] = mxz;
    M[9] = myz;
}

template <class R>
__device__ __forceinline__ void bc_solid(uint32_t tag, R omega, const R *f, R *M) {
    switch (tag) {
    case 7: bc_solid_tag<qbits<0, 1, 2, 3, 4, 5, 7, 9, 11, 13, 14, 16, 18, 19, 23, 25>(), 0>(omega, f, M); break;
    case 11: bc_solid_tag<qbits<0, 1, 2, 3, 4, 5, 7, 8, 9, 11, 14, 16, 18, 19, 22, 25>(), 1>(omega, f, M); break;
    case 13: bc_solid_tag<qbits<0, 1, 2, 3, 4, 5, 7, 8, 9, 11, 13, 16, 18, 19, 22, 23>(), 2>(omega, f, M); break;
    case 14: bc_solid_tag<qbits<0, 1, 2, 3, 4, 5, 8, 9, 11, 13, 14, 16, 18, 22, 23, 25>(), 3>(omega, f, M); break;
    case 112: bc_solid_tag<qbits<0, 1, 2, 3, 4, 6, 7, 10, 12, 13, 14, 15, 17, 21, 24, 26>(), 3>(omega, f, M); break;
    case 176: bc_solid_tag<qbits<0, 1, 2, 3, 4, 6, 7, 8, 10, 12, 14, 15, 17, 20, 21, 24>(), 2>(omega, f, M); break;
    case 208: bc_solid_tag<qbits<0, 1, 2, 3, 4, 6, 7, 8, 10, 12, 13, 15, 17, 20, 21, 26>(), 1>(omega, f, M); break;
    default: bc_solid_tag<qbits<0, 1, 2, 3, 4, 6, 8, 10, 12, 13, 14, 15, 17, 20, 24, 26>(), 0>(omega, f, M); break;
    }
}

// ---------------------------------------------------------------- masked lattice-frame sums

// Zeroth, first and second lattice-frame moments of the populations selected by `mask`
// (runtime bit set): S0 = sum f, J = sum f c, P = sum f c c (xx, yy, zz, xy, xz, yz).
template <class R>
struct MaskedSums {
    R s0, jx, jy, jz, pxx, pyy, pzz, pxy, pxz, pyz;
};

template <class R>
__device__ __forceinline__ MaskedSums<R> masked_sums(uint32_t mask, const R *f) {
    MaskedSums<R> m{R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const R v = ((mask >> q) & 1u) ? f[q] : R(0);
        const int cx = LATTICE.cx[q], cy = LATTICE.cy[q], cz = LATTICE.cz[q];
        m.s0 += v;
        if (cx) m.jx += cx > 0 ? v : -v;
        if (cy) m.jy += cy > 0 ? v : -v;
        if (cz) m.jz += cz > 0 ? v : -v;
        if (cx) m.pxx += v;
        if (cy) m.pyy += v;
        if (cz) m.pzz += v;
        if (cx && cy) m.pxy += cx * cy > 0 ? v : -v;
        if (cx && cz) m.pxz += cx * cz > 0 ? v : -v;
        if (cy && cz) m.pyz += cy * cz > 0 ? v : -v;
    }
    return m;
}

// ---------------------------------------------------------------- BCFLUID closed forms
// fluid_boundary_condition, cases/cylinder/boundaries.cuh:845-1037 (Equilibrium mass variant).
// The four tags are x/y mirror images: with (sx, sy) = (+,+) for 17, (-,+) for 34, (+,-) for 68
// and (-,-) for 136, each tag's formulas are tag 17's applied to (sx ux, sy uy, sx sy mxy,
// sx mxz, sy myz), with the outputs mapped back the same way.
template <class R>
__device__ __forceinline__ void bc_fluid(uint32_t tag, uint32_t in_mask, R omega, bool strong, const R *f, R *M) {
    const R third = R(1) / R(3);
    const R sx = (tag == 34u || tag == 136u) ? R(-1) : R(1);
    const R sy = (tag == 68u || tag == 136u) ? R(-1) : R(1);
    R rI, ux, uy, uz, mxx, myy, mzz, mxy, mxz, myz;
    if constexpr (kExactOrder<R>) {
        // per-direction accumulation of the incoming moments (boundaries.cuh:861-888)
        R a[10] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
        for (int q = 0; q < Q; ++q) {
            if (!(in_mask & (1u << q))) continue;
            const int cx = LATTICE.cx[q], cy = LATTICE.cy[q], cz = LATTICE.cz[q];
            const R fq = f[q];
            a[0] += fq;
            a[1] += fq * R(cx);
            a[2] += fq * R(cy);
            a[3] += fq * R(cz);
            a[4] += fq * (R(cx * cx) - third);
            a[5] += fq * (R(cy * cy) - third);
            a[6] += fq * (R(cz * cz) - third);
            a[7] += fq * R(cx * cy);
            a[8] += fq * R(cx * cz);
            a[9] += fq * R(cy * cz);
        }
        rI = a[0];
        const R inv = R(1) / rI;
        ux = sx * (a[1] * inv);
        uy = sy * (a[2] * inv);
        uz = a[3] * inv;
        mxx = a[4] * inv;
        myy = a[5] * inv;
        mzz = a[6] * inv;
        mxy = (sx * sy) * (a[7] * inv);
        mxz = sx * (a[8] * inv);
        myz = sy * (a[9] * inv);
    } else {
        const MaskedSums<R> S = masked_sums(in_mask, f);
        rI = S.s0;
        const R inv = R(1) / rI;
        ux = sx * (S.jx * inv);
        uy = sy * (S.jy * inv);
        uz = S.jz * inv;
        mxx = S.pxx * inv - third;
        myy = S.pyy * inv - third;
        mzz = S.pzz * inv - third;
        mxy = (sx * sy) * (S.pxy * inv);
        mxz = sx * (S.pxz * inv);
        myz = sy * (S.pyz * inv);
    }
    // density: root of a rho^2 + b rho + c = 0 closest to rho_I, Equilibrium form
    // (boundaries.cuh:905-915) or Strong form (:894-904; written times -1 for tags 34 and 68,
    // which swaps the labels of the roots, so a, b, c carry that sign)
    const R poly = R(45) * (mxx * mxx) + R(405) * (mxy * mxy) + R(45) * (myy * myy) + R(345) * myy * ux +
                   R(589) * (ux * ux) + R(345) * myy * uy + R(1467) * ux * uy + R(589) * (uy * uy) +
                   R(45) * mxy * (R(6) * myy + R(23) * (ux + uy)) +
                   R(15) * mxx * (R(18) * mxy + R(6) * myy + R(23) * (ux + uy));
    R a, b, c;
    if (strong) {
        const R sg = (tag == 34u || tag == 68u) ? R(-1) : R(1);
        a = sg * (R(-9996) - R(206) * omega);
        b = sg * (R(-6) * rI *
                  (R(12) * mxx * (R(-17) + R(20) * omega) -
                   R(17) * (R(102) + R(36) * mxy + R(12) * myy - R(5) * ux - R(5) * uy) +
                   omega * (R(720) * mxy + R(240) * myy + R(53) * (ux + uy))));
        c = sg * (R(3) * omega * (rI * rI) * poly);
    } else {
        a = R(-9998);
        b = R(6) * rI * (R(1734) + R(66) * mxx + R(198) * mxy + R(66) * myy + R(253) * ux + R(253) * uy);
        c = R(3) * (rI * rI) * poly;
    }
    const R disc = fmax(b * b - R(4) * a * c, R(0));
    const R sq = sqrt(disc);
    const R rho1 = (-b + sq) / (R(2) * a);
    const R rho2 = (-b - sq) / (R(2) * a);
    const R rho = (rho1 > R(0) && fabs(rho1 - rI) < fabs(rho2 - rI)) ? rho1 : rho2;
    // closed forms, boundaries.cuh:916-924 (FP64 divides as written, FP32 by reciprocals)
    const R d17 = R(17) * rho, d29 = R(29) * rho, d51 = R(51) * rho;
    auto div = [](R num, R den) {
        if constexpr (kExactOrder<R>)
            return num / den;
        else
            return num * (R(1) / den);
    };
    const R uxo = div(R(3) * mxx * rI + R(9) * mxy * rI + R(3) * myy * rI + R(20) * rI * ux + R(3) * rI * uy + rho, d17);
    const R uyo = div(R(3) * mxx * rI + R(9) * mxy * rI + R(3) * myy * rI + R(3) * rI * ux + R(20) * rI * uy + rho, d17);
    const R uzo = div(R(3) * rI * (mxz + myz + R(10) * uz), d29);
    const R mxxo = div(R(57) * mxx * rI + R(18) * mxy * rI + R(6) * myy * rI + R(6) * rI * ux + R(6) * rI * uy + R(2) * rho, d51);
    const R myyo = div(R(6) * mxx * rI + R(18) * mxy * rI + R(57) * myy * rI + R(6) * rI * ux + R(6) * rI * uy + R(2) * rho, d51);
    const R mzzo = (R(36) * mzz * rI) / (R(35) * rho);
    const R mxyo = div(R(3) * mxx * rI + R(26) * mxy * rI + R(3) * myy * rI + R(3) * rI * ux + R(3) * rI * uy + rho, d17);
    const R mxzo = div(rI * (R(32) * mxz + R(3) * myz + uz), d29);
    const R myzo = div(rI * (R(3) * mxz + R(32) * myz + uz), d29);
    M[0] = rho;
    M[1] = sx * uxo;
    M[2] = sy * uyo;
    M[3] = uzo;
    M[4] = mxxo;
    M[5] = myyo;
    M[6] = mzzo;
    M[7] = (sx * sy) * mxyo;
    M[8] = sx * mxzo;
    M[9] = sy * myzo;
}

// ---------------------------------------------------------------- curved walls

// Incoming density and rotated incoming second moments at a wall node, x' along the normal
// (nx, ny) and y' = (-ny, nx) (evaluate_incoming_moments_rotated, boundary/curvedLBM.cu:5-59).
// The rotated sums are the rotation of the lattice-frame masked second-moment tensor.
template <class R>
__device__ __forceinline__ void incoming_rotated(R nx, R ny, uint32_t in_mask, const R *f, R *M) {
    if constexpr (kExactOrder<R>) {
        const R third = R(1) / R(3);
        R r = 0, axx = 0, ayy = 0, azz = 0, axy = 0, axz = 0, ayz = 0;
        for (int q = 0; q < Q; ++q) {
            if (!(in_mask & (1u << q))) continue;
            const R cx = R(LATTICE.cx[q]), cy = R(LATTICE.cy[q]), cz = R(LATTICE.cz[q]);
            const R xp = cx * nx + cy * ny, yp = -cx * ny + cy * nx, zp = cz;
            const R fq = f[q];
            r += fq;
            axx += fq * (xp * xp - third);
            ayy += fq * (yp * yp - third);
            azz += fq * (zp * zp - third);
            axy += fq * (xp * yp);
            axz += fq * (xp * zp);
            ayz += fq * (yp * zp);
        }
        const R inv = R(1) / r;
        M[0] = r;
        M[4] = axx * inv;
        M[5] = ayy * inv;
        M[6] = azz * inv;
        M[7] = axy * inv;
        M[8] = axz * inv;
        M[9] = ayz * inv;
        return;
    }
    const MaskedSums<R> S = masked_sums(in_mask, f);
    const R inv = R(1) / S.s0;
    const R third = R(1) / R(3);
    const R nn = nx * ny, cc = nx * nx, ss = ny * ny;
    const R txx = cc * S.pxx + R(2) * nn * S.pxy + ss * S.pyy;
    const R tyy = ss * S.pxx - R(2) * nn * S.pxy + cc * S.pyy;
    const R txy = nn * (S.pyy - S.pxx) + (cc - ss) * S.pxy;
    const R txz = nx * S.pxz + ny * S.pyz;
    const R tyz = nx * S.pyz - ny * S.pxz;
    M[0] = S.s0;
    M[4] = txx * inv - third;
    M[5] = tyy * inv - third;
    M[6] = S.pzz * inv - third;
    M[7] = txy * inv;
    M[8] = txz * inv;
    M[9] = tyz * inv;
}

// Geometry constants of the wall-frame closure of one wall node: masked sums over its incoming
// directions of the weights times wall-frame Hermite tensors (c' rotated by the normal). With
// them, numerical_solution_rhoeq_rotated (boundary/numerical_solution.cu:4-122) is
//   E'   = G0 + 3 u'.G1 + 4.5 u'u' : G2
//   D'ab = G2ab + 3 u'.Dab      (ab = xy, xz)
//   F'ab,cd = 4.5 sum w H'cd H'ab
// and the 2x2 solve for m'xy, m'xz. Computed on the host once per wall node.
constexpr int WALL_CONSTS = 28;
// layout: [0] G0, [1..3] G1, [4..9] G2 (xx yy zz xy xz yz), [10..12] Dxy, [13..15] Dxz,
//         [16..21] Fxy (xx yy zz xy xz yz), [22..27] Fxz
inline void wall_constants(double nx, double ny, uint32_t in_mask, double *k) {
    for (int i = 0; i < WALL_CONSTS; ++i) k[i] = 0.0;
    const double third = 1.0 / 3.0;
    for (int q = 0; q < Q; ++q) {
        if (!(in_mask & (1u << q))) continue;
        const double cx = LATTICE.cx[q], cy = LATTICE.cy[q], cz = LATTICE.cz[q];
        const double xp = cx * nx + cy * ny, yp = -cx * ny + cy * nx, zp = cz;
        const double H[6] = {xp * xp - third, yp * yp - third, zp * zp - third, xp * yp, xp * zp, yp * zp};
        const double w = weight(q);
        const double cp[3] = {xp, yp, zp};
        k[0] += w;
        for (int a = 0; a < 3; ++a) {
            k[1 + a] += w * cp[a];
            k[10 + a] += w * cp[a] * H[3];
            k[13 + a] += w * cp[a] * H[4];
        }
        for (int b = 0; b < 6; ++b) {
            k[4 + b] += w * H[b];
            k[16 + b] += 4.5 * w * H[b] * H[3];
            k[22 + b] += 4.5 * w * H[b] * H[4];
        }
    }
}

// Wall-frame closure with precomputed constants k (see wall_constants): p holds the extrapolated
// u'x, u'y, u'z, m'xx, m'yy, m'zz, m'yz; rhoI, mxyI, mxzI the node's incoming values. Solves for
// m'xy, m'xz, rotates back to the lattice frame and closes rho = rho_I / E'.
template <class R>
__device__ __forceinline__ void wall_closure(R nx, R ny, const R *k, const R *p, R rhoI, R mxyI, R mxzI, R *M) {
    const R c = nx, s = ny;
    const R s2 = R(2) * s * c, c2 = c * c - s * s;
    const R ux = p[0], uy = p[1], uz = p[2], mxx = p[3], myy = p[4], mzz = p[5], myz = p[6];
    const R E = k[0] + R(3) * (ux * k[1] + uy * k[2] + uz * k[3]) +
                R(4.5) * (ux * ux * k[4] + uy * uy * k[5] + uz * uz * k[6] +
                          R(2) * (ux * uy * k[7] + ux * uz * k[8] + uy * uz * k[9]));
    const R Dxy = k[7] + R(3) * (ux * k[10] + uy * k[11] + uz * k[12]);
    const R Dxz = k[8] + R(3) * (ux * k[13] + uy * k[14] + uz * k[15]);
    const R a1 = R(2) * k[19], b1 = R(2) * k[20];
    const R c1 = (E * mxyI - Dxy) - mxx * k[16] - myy * k[17] - mzz * k[18] - R(2) * myz * k[21];
    const R a2 = R(2) * k[25], b2 = R(2) * k[26];
    const R cc = (E * mxzI - Dxz) - mxx * k[22] - myy * k[23] - mzz * k[24] - R(2) * myz * k[27];
    const R inv = R(1) / (a2 * b1 - a1 * b2);
    const R mxyp = (b1 * cc - b2 * c1) * inv;
    const R mxzp = (a2 * c1 - a1 * cc) * inv;
    M[1] = ux * c - uy * s;
    M[2] = ux * s + uy * c;
    M[3] = uz;
    M[4] = mxx * c * c + myy * s * s - mxyp * s2;
    M[5] = mxx * s * s + myy * c * c + mxyp * s2;
    M[6] = mzz;
    M[7] = (mxx - myy) * R(0.5) * s2 + mxyp * c2;
    M[8] = mxzp * c - myz * s;
    M[9] = mxzp * s + myz * c;
    M[0] = rhoI / E;
}

// Per-direction loop form of the same closure (numerical_solution.cu:37-121), used in FP64.
template <class R>
__device__ __forceinline__ void wall_closure_loop(R nx, R ny, uint32_t in_mask, const R *p, R rhoI, R mxyI, R mxzI,
                                                  R *M) {
    const R c = nx, s = ny, third = R(1) / R(3);
    const R s2 = R(2) * s * c, c2 = c * c - s * s;
    const R ux = p[0], uy = p[1], uz = p[2], mxx = p[3], myy = p[4], mzz = p[5], myz = p[6];
    R E = 0, Dxy = 0, Dxz = 0;
    R F11a = 0, F22a = 0, F33a = 0, F12a = 0, F13a = 0, F23a = 0;
    R F11b = 0, F22b = 0, F33b = 0, F12b = 0, F13b = 0, F23b = 0;
    for (int q = 0; q < Q; ++q) {
        if (!(in_mask & (1u << q))) continue;
        const R cx = R(LATTICE.cx[q]), cy = R(LATTICE.cy[q]), cz = R(LATTICE.cz[q]);
        const R xp = cx * c + cy * s, yp = -cx * s + cy * c, zp = cz;
        const R Hxx = xp * xp - third, Hyy = yp * yp - third, Hzz = zp * zp - third;
        const R Hxy = xp * yp, Hxz = xp * zp, Hyz = yp * zp;
        const R wq = R(weight(q));
        const R cf = R(4.5) * wq;
        const R A = wq * (R(1) + R(3) * (ux * xp + uy * yp + uz * zp));
        E += A + cf * (ux * ux * Hxx + uy * uy * Hyy + uz * uz * Hzz + R(2) * ux * uy * Hxy + R(2) * ux * uz * Hxz +
                       R(2) * uy * uz * Hyz);
        Dxy += A * Hxy;
        Dxz += A * Hxz;
        F11a += (cf * Hxx) * Hxy;
        F22a += (cf * Hyy) * Hxy;
        F33a += (cf * Hzz) * Hxy;
        F12a += (cf * Hxy) * Hxy;
        F13a += (cf * Hxz) * Hxy;
        F23a += (cf * Hyz) * Hxy;
        F11b += (cf * Hxx) * Hxz;
        F22b += (cf * Hyy) * Hxz;
        F33b += (cf * Hzz) * Hxz;
        F12b += (cf * Hxy) * Hxz;
        F13b += (cf * Hxz) * Hxz;
        F23b += (cf * Hyz) * Hxz;
    }
    const R a1 = R(2) * F12a, b1 = R(2) * F13a;
    const R c1 = (E * mxyI - Dxy) - mxx * F11a - myy * F22a - mzz * F33a - R(2) * myz * F23a;
    const R a2 = R(2) * F12b, b2 = R(2) * F13b;
    const R cc = (E * mxzI - Dxz) - mxx * F11b - myy * F22b - mzz * F33b - R(2) * myz * F23b;
    const R inv = R(1) / (a2 * b1 - a1 * b2);
    const R mxyp = (b1 * cc - b2 * c1) * inv;
    const R mxzp = (a2 * c1 - a1 * cc) * inv;
    M[1] = ux * c - uy * s;
    M[2] = ux * s + uy * c;
    M[3] = uz;
    M[4] = mxx * c * c + myy * s * s - mxyp * s2;
    M[5] = mxx * s * s + myy * c * c + mxyp * s2;
    M[6] = mzz;
    M[7] = (mxx - myy) * R(0.5) * s2 + mxyp * c2;
    M[8] = mxzp * c - myz * s;
    M[9] = mxzp * s + myz * c;
    M[0] = rhoI / E;
}

// MassBC::Strong form of the closure (numerical_solution_strong_rotated,
// boundary/numerical_solution.cu:124-279): A', E' and B' over the outgoing directions, D' and F'
// over the incoming ones, F' corrected by (1 - omega) m'_I B', G' = (1 - omega) A' + omega E',
// and rho = rho_I / (G' + (1 - omega) m' : B'), so that the collided populations leaving along
// the outgoing directions carry rho_I. Per-direction loop in the FP64 evaluation order.
template <class R>
__device__ __forceinline__ void wall_closure_strong(R nx, R ny, uint32_t in_mask, uint32_t out_mask, R omega,
                                                    const R *p, R rhoI, R mxyI, R mxzI, R *M) {
    const R c = nx, s = ny, third = R(1) / R(3);
    const R s2 = R(2) * s * c, c2 = c * c - s * s;
    const R ux = p[0], uy = p[1], uz = p[2], mxx = p[3], myy = p[4], mzz = p[5], myz = p[6];
    R A = 0, E = 0, B11 = 0, B22 = 0, B33 = 0, B12 = 0, B13 = 0, B23 = 0, Dxy = 0, Dxz = 0;
    R F11a = 0, F22a = 0, F33a = 0, F12a = 0, F13a = 0, F23a = 0;
    R F11b = 0, F22b = 0, F33b = 0, F12b = 0, F13b = 0, F23b = 0;
    for (int q = 0; q < Q; ++q) {
        const bool o = out_mask & (1u << q), in = in_mask & (1u << q);
        if (!o && !in) continue;
        const R cx = R(LATTICE.cx[q]), cy = R(LATTICE.cy[q]), cz = R(LATTICE.cz[q]);
        const R xp = cx * c + cy * s, yp = -cx * s + cy * c, zp = cz;
        const R Hxx = xp * xp - third, Hyy = yp * yp - third, Hzz = zp * zp - third;
        const R Hxy = xp * yp, Hxz = xp * zp, Hyz = yp * zp;
        const R wq = R(weight(q));
        const R cf = R(4.5) * wq;
        const R Ai = wq * (R(1) + R(3) * (ux * xp + uy * yp + uz * zp));
        if (o) {
            A += Ai;
            E += Ai + cf * (ux * ux * Hxx + uy * uy * Hyy + uz * uz * Hzz + R(2) * ux * uy * Hxy +
                            R(2) * ux * uz * Hxz + R(2) * uy * uz * Hyz);
            B11 += cf * Hxx;
            B22 += cf * Hyy;
            B33 += cf * Hzz;
            B12 += cf * Hxy;
            B13 += cf * Hxz;
            B23 += cf * Hyz;
        }
        if (in) {
            Dxy += Ai * Hxy;
            Dxz += Ai * Hxz;
            F11a += (cf * Hxx) * Hxy;
            F22a += (cf * Hyy) * Hxy;
            F33a += (cf * Hzz) * Hxy;
            F12a += (cf * Hxy) * Hxy;
            F13a += (cf * Hxz) * Hxy;
            F23a += (cf * Hyz) * Hxy;
            F11b += (cf * Hxx) * Hxz;
            F22b += (cf * Hyy) * Hxz;
            F33b += (cf * Hzz) * Hxz;
            F12b += (cf * Hxy) * Hxz;
            F13b += (cf * Hxz) * Hxz;
            F23b += (cf * Hyz) * Hxz;
        }
    }
    const R om1 = R(1) - omega;
    const R G = om1 * A + omega * E;
    const R oxy = om1 * mxyI, oxz = om1 * mxzI;
    F11a -= oxy * B11;
    F22a -= oxy * B22;
    F33a -= oxy * B33;
    F12a -= oxy * B12;
    F13a -= oxy * B13;
    F23a -= oxy * B23;
    F11b -= oxz * B11;
    F22b -= oxz * B22;
    F33b -= oxz * B33;
    F12b -= oxz * B12;
    F13b -= oxz * B13;
    F23b -= oxz * B23;
    const R a1 = R(2) * F12a, b1 = R(2) * F13a;
    const R c1 = (G * mxyI - Dxy) - mxx * F11a - myy * F22a - mzz * F33a - R(2) * myz * F23a;
    const R a2 = R(2) * F12b, b2 = R(2) * F13b;
    const R cc = (G * mxzI - Dxz) - mxx * F11b - myy * F22b - mzz * F33b - R(2) * myz * F23b;
    const R inv = R(1) / (a2 * b1 - a1 * b2);
    const R mxyp = (b1 * cc - b2 * c1) * inv;
    const R mxzp = (a2 * c1 - a1 * cc) * inv;
    M[1] = ux * c - uy * s;
    M[2] = ux * s + uy * c;
    M[3] = uz;
    M[4] = mxx * c * c + myy * s * s - mxyp * s2;
    M[5] = mxx * s * s + myy * c * c + mxyp * s2;
    M[6] = mzz;
    M[7] = (mxx - myy) * R(0.5) * s2 + mxyp * c2;
    M[8] = mxzp * c - myz * s;
    M[9] = mxzp * s + myz * c;
    const R den = G + om1 * (mxx * B11 + myy * B22 + mzz * B33 + R(2) * mxyp * B12 + R(2) * mxzp * B13 +
                             R(2) * myz * B23);
    M[0] = rhoI / den;
}

// Momentum and mass carried by the populations of M selected by mask (compute_forces_mass,
// statistics/forces.cu:3-23), accumulated in FP64 with the given sign.
template <class R>
__device__ __forceinline__ void masked_force(const R *M, uint32_t mask, double sign, double *F) {
    R f[Q];
    reconstruct(M, f);
    const MaskedSums<R> S = masked_sums(mask, f);
    F[0] += sign * double(S.jx);
    F[1] += sign * double(S.jy);
    F[2] += sign * double(S.jz);
    F[3] += sign * double(S.s0);
}

}  // namespace mrlbm
