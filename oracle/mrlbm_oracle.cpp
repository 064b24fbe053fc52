// mrlbm_oracle.cpp - CPU oracle: a plain restatement of the reference's D3Q27 MR-LBM step.
//
// TEST INFRASTRUCTURE ONLY (see mrlbm_oracle.h). It follows the reference's two-kernel order
// literally, on x-fastest host arrays, with a naive global periodic pull instead of the
// reference's tile/ghost-plane machinery (the two are equal because ghost pops and in-tile pops
// are reconstructed from the same stored moments: core/streaming.cuh:6-35,
// core/D3Q27/halo_interface.cuh:5-201, solver/mlbm.cu:186-242).
//
// Paths below are relative to the reference repository root, MRLBM_3D/src/ omitted.
#include "mrlbm_oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

namespace {

constexpr int Q = 27;

// Velocity ordering of core/D3Q27/latticeProperties.cuh:17-21 (0 rest, 1-6 axes, 7-18 face
// diagonals, 19-26 corners). The ordering defines the wall masks and BCFLUID tags, so it has
// to be the reference's.
const int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
const int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
const int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

int opposite(int q) {
    for (int p = 0; p < Q; ++p)
        if (CX[p] == -CX[q] && CY[p] == -CY[q] && CZ[p] == -CZ[q]) return p;
    return -1;
}

// Weights from the 1-D factorisation w = w1(cx) w1(cy) w1(cz), w1(0) = 2/3, w1(+-1) = 1/6
// (equals W0..W3 of latticeProperties.cuh:7-10).
double weight(int q) {
    auto w1 = [](int c) { return c == 0 ? 2.0 / 3.0 : 1.0 / 6.0; };
    return w1(CX[q]) * w1(CY[q]) * w1(CZ[q]);
}

// Reference node types, common/nodeTypeMap.cuh:38-80.
enum : uint32_t {
    T_BULK = 0, T_SOLID, T_INNER, T_OUTER, T_BCF_INNER, T_BCF_OUTER, T_BCS_INNER, T_BCS_OUTER,
    T_NORTH, T_SOUTH, T_EAST, T_WEST, T_FRONT, T_BACK,
    T_NORTH_EAST, T_NORTH_WEST, T_NORTH_FRONT, T_NORTH_BACK, T_SOUTH_EAST, T_SOUTH_WEST,
    T_SOUTH_FRONT, T_SOUTH_BACK, T_EAST_FRONT, T_EAST_BACK, T_WEST_FRONT, T_WEST_BACK,
    T_NWF, T_NWB, T_SWF, T_SWB, T_NEF, T_NEB, T_SEF, T_SEB
};

constexpr uint32_t TAG_BCF_1 = 17, TAG_BCF_2 = 34, TAG_BCF_3 = 68, TAG_BCF_4 = 136;

// Octant direction sets used for wall/BCFLUID classification and masks
// (cases/cylinder/nodeClass.cuh:67-75, cylinder_helpers.cuh:143-151).
const int OCTANT_DIRS[8][7] = {{2, 4, 6, 8, 10, 12, 20}, {1, 4, 6, 12, 13, 15, 26},
                               {2, 3, 6, 10, 14, 17, 24}, {1, 3, 6, 7, 15, 17, 21},
                               {2, 4, 5, 8, 16, 18, 22}, {1, 4, 5, 9, 13, 18, 23},
                               {2, 3, 5, 11, 14, 16, 25}, {1, 3, 5, 7, 9, 11, 19}};

thread_local std::string g_err;
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------- per-node math (templated)

// Moment vector convention: m[0] = rho (absolute), m[1..3] = u, m[4..6] = mxx, myy, mzz,
// m[7..9] = mxy, mxz, myz.

// Regularised population q (core/D3Q27/second_order/reconstruction.cuh:5-60), written as the
// second-order Hermite expansion w rho [1 + 3 c.u + 9/2 sum_ab m_ab H_ab].
template <class R>
R pop_q(const R *m, int q) {
    const R cx = R(CX[q]), cy = R(CY[q]), cz = R(CZ[q]);
    const R third = R(1) / R(3);
    const R cu = cx * m[1] + cy * m[2] + cz * m[3];
    const R hd = m[4] * (cx * cx - third) + m[5] * (cy * cy - third) + m[6] * (cz * cz - third);
    const R ho = m[7] * cx * cy + m[8] * cx * cz + m[9] * cy * cz;
    return R(weight(q)) * m[0] * (R(1) + R(3) * cu + R(4.5) * hd + R(9) * ho);
}

// core/D3Q27/eval_moments.cuh:5-22.
template <class R>
void moments(const R *f, R *m) {
    R rho = 0, jx = 0, jy = 0, jz = 0, pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
    for (int q = 0; q < Q; ++q) {
        rho += f[q];
        jx += f[q] * R(CX[q]);
        jy += f[q] * R(CY[q]);
        jz += f[q] * R(CZ[q]);
        pxx += f[q] * R(CX[q] * CX[q]);
        pyy += f[q] * R(CY[q] * CY[q]);
        pzz += f[q] * R(CZ[q] * CZ[q]);
        pxy += f[q] * R(CX[q] * CY[q]);
        pxz += f[q] * R(CX[q] * CZ[q]);
        pyz += f[q] * R(CY[q] * CZ[q]);
    }
    const R inv = R(1) / rho;
    const R cs2 = R(1) / R(3);
    m[0] = rho;
    m[1] = jx * inv;
    m[2] = jy * inv;
    m[3] = jz * inv;
    m[4] = pxx * inv - cs2;
    m[5] = pyy * inv - cs2;
    m[6] = pzz * inv - cs2;
    m[7] = pxy * inv;
    m[8] = pxz * inv;
    m[9] = pyz * inv;
}

// moment_collision, solver/mlbm.cu:3-16.
template <class R>
void collide(R omega, R *m) {
    const R om1 = R(1) - omega;
    m[4] = om1 * m[4] + omega * m[1] * m[1];
    m[5] = om1 * m[5] + omega * m[2] * m[2];
    m[6] = om1 * m[6] + omega * m[3] * m[3];
    m[7] = om1 * m[7] + omega * m[1] * m[2];
    m[8] = om1 * m[8] + omega * m[1] * m[3];
    m[9] = om1 * m[9] + omega * m[2] * m[3];
}

template <class R>
R sum_pops(const R *f, std::initializer_list<int> qs) {
    R s = 0;
    for (int q : qs) s += f[q];
    return s;
}

// Domain-face closed forms, cases/cylinder/boundaries.cuh:201-843 (faces and the x-y edges
// used with z periodicity). `old` and `old_xm1` are the node's and its x-1 neighbour's moments
// at the start of the step (update_neumaan_density_velocity, boundaries.cuh:163-199).
// Returns false for types without a closed form (SOLID and the non-matching cases), which
// leave the node unchanged.
template <class R>
bool face_bc(uint32_t type, R omega, R u_in, const R *f, const R *old, const R *old_xm1, R *m) {
    auto neumann = [&](R &rho, R &ux, R &uy, R &uz) {
        // CONVECTIVE_OUTLET = 1 (cases/cylinder/constants.cuh:12): blend of boundary and
        // interior old values with Uc = |u_x(x-1)| (boundaries.cuh:174-192).
        const R rhoi = old_xm1[0], uxi = old_xm1[1], uyi = old_xm1[2], uzi = old_xm1[3];
        const R rhob = old[0], uxb = old[1], uyb = old[2], uzb = old[3];
        const R Uc = std::sqrt(uxi * uxi);
        rho = (R(1) - Uc) * rhob + Uc * rhoi;
        ux = (R(1) - Uc) * uxb + Uc * uxi;
        uy = (R(1) - Uc) * uyb + Uc * uyi;
        uz = (R(1) - Uc) * uzb + Uc * uzi;
    };
    R &rho = m[0], &ux = m[1], &uy = m[2], &uz = m[3];
    R &mxx = m[4], &myy = m[5], &mzz = m[6], &mxy = m[7], &mxz = m[8], &myz = m[9];
    switch (type) {
    case T_WEST: {  // boundaries.cuh:208-229, with the prescribed inlet velocity u_in
        const R rI = sum_pops(f, {0, 2, 3, 4, 5, 6, 8, 10, 11, 12, 14, 16, 17, 18, 20, 22, 24, 25});
        const R inv = R(1) / rI;
        const R mxyI = ((f[8] + f[20] + f[22]) - (f[14] + f[24] + f[25])) * inv;
        const R mxzI = ((f[10] + f[20] + f[24]) - (f[16] + f[22] + f[25])) * inv;
        const R uxo = old[1];  // the node's previous u_x (D11)
        rho = -R(6) * rI / (-R(5) + R(3) * uxo + R(3) * uxo * uxo);
        ux = u_in;
        uy = 0;
        uz = 0;
        mxx = ux * ux;
        myy = 0;
        mzz = 0;
        mxy = R(2) * mxyI * rI / rho;
        mxz = R(2) * mxzI * rI / rho;
        myz = 0;
        return true;
    }
    case T_EAST: {  // boundaries.cuh:230-271
        const R rI = sum_pops(f, {0, 1, 3, 4, 5, 6, 7, 9, 11, 12, 13, 15, 17, 18, 19, 21, 23, 26});
        const R inv = R(1) / rI;
        const R myyI = sum_pops(f, {3, 4, 7, 11, 12, 13, 17, 18, 19, 21, 23, 26}) * inv - R(1) / R(3);
        const R mzzI = sum_pops(f, {5, 6, 9, 11, 12, 15, 17, 18, 19, 21, 23, 26}) * inv - R(1) / R(3);
        const R mxyI = ((f[7] + f[19] + f[21]) - (f[13] + f[23] + f[26])) * inv;
        const R mxzI = ((f[9] + f[19] + f[23]) - (f[15] + f[21] + f[26])) * inv;
        const R myzI = (f[11] + f[12] - f[17] - f[18] + f[19] - f[21] - f[23] + f[26]) * inv;
        neumann(rho, ux, uy, uz);
        mxx = ux * ux;
        myy = (R(6) * myyI * rI - R(6) * mzzI * rI + R(5) * uy * uy * rho + R(5) * uz * uz * rho) / (R(10) * rho);
        mzz = (-R(6) * myyI * rI + R(6) * mzzI * rI + R(5) * uy * uy * rho + R(5) * uz * uz * rho) / (R(10) * rho);
        mxy = (R(6) * mxyI * rI - uy * rho) / (R(3) * rho);
        mxz = (R(6) * mxzI * rI - uz * rho) / (R(3) * rho);
        myz = R(6) * myzI * rI / (R(5) * rho);
        return true;
    }
    case T_NORTH: {  // boundaries.cuh:272-293
        const R rI = sum_pops(f, {0, 1, 2, 3, 5, 6, 7, 9, 10, 11, 14, 15, 16, 17, 19, 21, 24, 25});
        const R inv = R(1) / rI;
        const R mxyI = ((f[7] + f[19] + f[21]) - (f[14] + f[24] + f[25])) * inv;
        const R myzI = ((f[11] + f[19] + f[25]) - (f[17] + f[21] + f[24])) * inv;
        rho = R(6) * rI / R(5);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = R(2) * mxyI * rI / rho;
        mxz = 0;
        myz = R(2) * myzI * rI / rho;
        return true;
    }
    case T_SOUTH: {  // boundaries.cuh:294-315
        const R rI = sum_pops(f, {0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 15, 16, 18, 20, 22, 23, 26});
        const R inv = R(1) / rI;
        const R mxyI = ((f[8] + f[20] + f[22]) - (f[13] + f[23] + f[26])) * inv;
        const R myzI = ((f[12] + f[20] + f[26]) - (f[18] + f[22] + f[23])) * inv;
        rho = R(6) * rI / R(5);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = R(2) * mxyI * rI / rho;
        mxz = 0;
        myz = R(2) * myzI * rI / rho;
        return true;
    }
    case T_NORTH_WEST: {  // boundaries.cuh:360-381
        const R rI = sum_pops(f, {0, 2, 3, 5, 6, 10, 11, 14, 16, 17, 24, 25});
        const R inv = R(1) / rI;
        const R mxyI = -(f[14] + f[24] + f[25]) * inv;
        const R myzI = (f[11] - f[17] - f[24] + f[25]) * inv;
        rho = R(36) * rI * (R(1) + mxyI - mxyI * omega) / (R(24) + omega);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = (R(36) * mxyI * rI + rho) / (R(9) * rho);
        mxz = 0;
        myz = R(12) * myzI * rI / (R(5) * rho);
        return true;
    }
    case T_SOUTH_WEST: {  // boundaries.cuh:382-403
        const R rI = sum_pops(f, {0, 2, 4, 5, 6, 8, 10, 12, 16, 18, 20, 22});
        const R inv = R(1) / rI;
        const R mxyI = (f[8] + f[20] + f[22]) * inv;
        const R myzI = (f[12] - f[18] + f[20] - f[22]) * inv;
        rho = R(36) * rI * (R(1) - mxyI + mxyI * omega) / (R(24) + omega);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = (R(36) * mxyI * rI - rho) / (R(9) * rho);
        mxz = 0;
        myz = R(12) * myzI * rI / (R(5) * rho);
        return true;
    }
    case T_NORTH_EAST: {  // boundaries.cuh:448-480
        const R rI = sum_pops(f, {0, 1, 3, 5, 6, 7, 9, 11, 15, 17, 19, 21});
        const R inv = R(1) / rI;
        const R mxyI = (f[7] + f[19] + f[21]) * inv;
        const R myzI = (f[11] - f[17] + f[19] - f[21]) * inv;
        neumann(rho, ux, uy, uz);
        mxx = myy = mzz = 0;
        mxy = (R(36) * mxyI * rI - rho) / (R(9) * rho);
        mxz = 0;
        myz = (R(12) * myzI * rI) / (R(5) * rho);
        return true;
    }
    case T_SOUTH_EAST: {  // boundaries.cuh:481-513
        const R rI = sum_pops(f, {0, 1, 4, 5, 6, 9, 12, 13, 15, 18, 23, 26});
        const R inv = R(1) / rI;
        const R mxyI = -(f[13] + f[23] + f[26]) * inv;
        const R myzI = (f[12] - f[18] - f[23] + f[26]) * inv;
        neumann(rho, ux, uy, uz);
        mxx = myy = mzz = 0;
        mxy = (R(36) * mxyI * rI + rho) / (R(9) * rho);
        mxz = 0;
        myz = (R(12) * myzI * rI) / (R(5) * rho);
        return true;
    }
    case T_BACK: {  // boundaries.cuh:316-337, the z = 0 no-slip wall
        const R rI = sum_pops(f, {0, 1, 2, 3, 4, 6, 7, 8, 10, 12, 13, 14, 15, 17, 20, 21, 24, 26});
        const R inv = R(1) / rI;
        const R mxzI = ((f[10] + f[20] + f[24]) - (f[15] + f[21] + f[26])) * inv;
        const R myzI = ((f[12] + f[20] + f[26]) - (f[17] + f[21] + f[24])) * inv;
        rho = R(6) * rI / R(5);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = 0;
        mxz = R(2) * mxzI * rI / rho;
        myz = R(2) * myzI * rI / rho;
        return true;
    }
    case T_FRONT: {  // boundaries.cuh:338-359, the z = NZ-1 no-slip wall
        const R rI = sum_pops(f, {0, 1, 2, 3, 4, 5, 7, 8, 9, 11, 13, 14, 16, 18, 19, 22, 23, 25});
        const R inv = R(1) / rI;
        const R mxzI = ((f[9] + f[19] + f[23]) - (f[16] + f[22] + f[25])) * inv;
        const R myzI = ((f[11] + f[19] + f[25]) - (f[18] + f[22] + f[23])) * inv;
        rho = R(6) * rI / R(5);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = 0;
        mxz = R(2) * mxzI * rI / rho;
        myz = R(2) * myzI * rI / rho;
        return true;
    }
    case T_WEST_FRONT: {  // boundaries.cuh:404-425
        const R rI = sum_pops(f, {0, 2, 3, 4, 5, 8, 11, 14, 16, 18, 22, 25});
        const R inv = R(1) / rI;
        const R mxzI = -(f[16] + f[22] + f[25]) * inv;
        const R myzI = (f[11] - f[18] - f[22] + f[25]) * inv;
        rho = R(36) * rI * (R(1) + mxzI - mxzI * omega) / (R(24) + omega);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = 0;
        mxz = (R(36) * mxzI * rI + rho) / (R(9) * rho);
        myz = R(12) * myzI * rI / (R(5) * rho);
        return true;
    }
    case T_WEST_BACK: {  // boundaries.cuh:426-447
        const R rI = sum_pops(f, {0, 2, 3, 4, 6, 8, 10, 12, 14, 17, 20, 24});
        const R inv = R(1) / rI;
        const R mxzI = (f[10] + f[20] + f[24]) * inv;
        const R myzI = (f[12] - f[17] + f[20] - f[24]) * inv;
        rho = R(36) * rI * (R(1) - mxzI + mxzI * omega) / (R(24) + omega);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = 0;
        mxz = (R(36) * mxzI * rI - rho) / (R(9) * rho);
        myz = R(12) * myzI * rI / (R(5) * rho);
        return true;
    }
    case T_EAST_FRONT: {  // boundaries.cuh:512-544 (outlet blend for rho and u)
        const R rI = sum_pops(f, {0, 1, 3, 4, 5, 7, 9, 11, 13, 18, 19, 23});
        const R inv = R(1) / rI;
        const R mxzI = (f[9] + f[19] + f[23]) * inv;
        const R myzI = (f[11] - f[18] + f[19] - f[23]) * inv;
        neumann(rho, ux, uy, uz);
        mxx = myy = mzz = 0;
        mxy = 0;
        mxz = (R(36) * mxzI * rI - rho) / (R(9) * rho);
        myz = (R(12) * myzI * rI) / (R(5) * rho);
        return true;
    }
    case T_EAST_BACK: {  // boundaries.cuh:545-577
        const R rI = sum_pops(f, {0, 1, 3, 4, 6, 7, 12, 13, 15, 17, 21, 26});
        const R inv = R(1) / rI;
        const R mxzI = -(f[15] + f[21] + f[26]) * inv;
        const R myzI = (f[12] - f[17] - f[21] + f[26]) * inv;
        neumann(rho, ux, uy, uz);
        mxx = myy = mzz = 0;
        mxy = 0;
        mxz = (R(36) * mxzI * rI + rho) / (R(9) * rho);
        myz = (R(12) * myzI * rI) / (R(5) * rho);
        return true;
    }
    case T_NORTH_FRONT:    // boundaries.cuh:578-598
    case T_NORTH_BACK:     // :599-619
    case T_SOUTH_FRONT:    // :620-640
    case T_SOUTH_BACK: {   // :641-661
        R rI, myzI;
        if (type == T_NORTH_FRONT) {
            rI = sum_pops(f, {0, 1, 2, 3, 5, 7, 9, 11, 14, 16, 19, 25});
            myzI = (f[11] + f[19] + f[25]) * (R(1) / rI);
        } else if (type == T_NORTH_BACK) {
            rI = sum_pops(f, {0, 1, 2, 3, 6, 7, 10, 14, 15, 17, 21, 24});
            myzI = -(f[17] + f[21] + f[24]) * (R(1) / rI);
        } else if (type == T_SOUTH_FRONT) {
            rI = sum_pops(f, {0, 1, 2, 4, 5, 8, 9, 13, 16, 18, 22, 23});
            myzI = -(f[18] + f[22] + f[23]) * (R(1) / rI);
        } else {
            rI = sum_pops(f, {0, 1, 2, 4, 6, 8, 10, 12, 13, 15, 20, 26});
            myzI = (f[12] + f[20] + f[26]) * (R(1) / rI);
        }
        // the known diagonal pair has c_y c_z > 0 (NORTH_FRONT, SOUTH_BACK) or < 0
        const bool pos = type == T_NORTH_FRONT || type == T_SOUTH_BACK;
        if (pos)
            rho = R(36) * rI * (R(1) - myzI + myzI * omega) / (R(24) + omega);
        else
            rho = R(36) * rI * (R(1) + myzI - myzI * omega) / (R(24) + omega);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = 0;
        mxz = 0;
        myz = pos ? (R(36) * myzI * rI - rho) / (R(9) * rho) : (R(36) * myzI * rI + rho) / (R(9) * rho);
        return true;
    }
    case T_NWF: case T_NWB: case T_SWF: case T_SWB: {  // boundaries.cuh:662-733: resting corners
        R rI;
        if (type == T_NWF) rI = sum_pops(f, {0, 2, 3, 5, 11, 14, 16, 25});
        else if (type == T_NWB) rI = sum_pops(f, {0, 2, 3, 6, 10, 14, 17, 24});
        else if (type == T_SWF) rI = sum_pops(f, {0, 2, 4, 5, 8, 16, 18, 22});
        else rI = sum_pops(f, {0, 2, 4, 6, 8, 10, 12, 20});
        rho = R(216) * rI / R(125);
        ux = uy = uz = 0;
        mxx = myy = mzz = 0;
        mxy = mxz = myz = 0;
        return true;
    }
    case T_NEF: case T_NEB: case T_SEF: case T_SEB: {  // boundaries.cuh:734-841: outlet corners
        neumann(rho, ux, uy, uz);
        mxx = myy = mzz = 0;
        mxy = mxz = myz = 0;
        return true;
    }
    default:
        return false;
    }
}

// bcsolid_boundary_condition, cases/cylinder/boundaries.cuh:1039-1214 (identical in
// cases/annulus/boundaries.cuh:290-465): end-wall nodes next to a curved wall, tags
// TAG_BCS_1..8 (common/nodeTypeMap.cuh:29-36). Four closed-form families: tags 7 and 224,
// 11 and 208, 13 and 176, 14 and 112. Other tags leave the node unchanged (no default).
template <class R>
bool bcsolid_bc(uint32_t tag, R omega, const R *f, R *m) {
    R rI, mxzI, myzI;
    int fam;
    switch (tag) {
    case 7:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 5, 7, 9, 11, 13, 14, 16, 18, 19, 23, 25});
        mxzI = (f[9] - f[16] + f[19] + f[23] - f[25]) * (R(1.0) / rI);
        myzI = (f[11] - f[18] + f[19] - f[23] + f[25]) * (R(1.0) / rI);
        fam = 0;
        break;
    case 11:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 5, 7, 8, 9, 11, 14, 16, 18, 19, 22, 25});
        mxzI = (f[9] - f[16] + f[19] - f[22] - f[25]) * (R(1.0) / rI);
        myzI = (f[11] - f[18] + f[19] - f[22] + f[25]) * (R(1.0) / rI);
        fam = 1;
        break;
    case 13:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 5, 7, 8, 9, 11, 13, 16, 18, 19, 22, 23});
        mxzI = (f[9] - f[16] + f[19] - f[22] + f[23]) * (R(1.0) / rI);
        myzI = (f[11] - f[18] + f[19] - f[22] - f[23]) * (R(1.0) / rI);
        fam = 2;
        break;
    case 14:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 5, 8, 9, 11, 13, 14, 16, 18, 22, 23, 25});
        mxzI = (f[9] - f[16] - f[22] + f[23] - f[25]) * (R(1.0) / rI);
        myzI = (f[11] - f[18] - f[22] - f[23] + f[25]) * (R(1.0) / rI);
        fam = 3;
        break;
    case 112:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 6, 7, 10, 12, 13, 14, 15, 17, 21, 24, 26});
        mxzI = (f[10] - f[15] - f[21] + f[24] - f[26]) * (R(1.0) / rI);
        myzI = (f[12] - f[17] - f[21] - f[24] + f[26]) * (R(1.0) / rI);
        fam = 3;
        break;
    case 176:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 6, 7, 8, 10, 12, 14, 15, 17, 20, 21, 24});
        mxzI = (f[10] - f[15] + f[20] - f[21] + f[24]) * (R(1.0) / rI);
        myzI = (f[12] - f[17] + f[20] - f[21] - f[24]) * (R(1.0) / rI);
        fam = 2;
        break;
    case 208:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 6, 7, 8, 10, 12, 13, 15, 17, 20, 21, 26});
        mxzI = (f[10] - f[15] + f[20] - f[21] - f[26]) * (R(1.0) / rI);
        myzI = (f[12] - f[17] + f[20] - f[21] + f[26]) * (R(1.0) / rI);
        fam = 1;
        break;
    case 224:
        rI = sum_pops(f, {0, 1, 2, 3, 4, 6, 8, 10, 12, 13, 14, 15, 17, 20, 24, 26});
        mxzI = (f[10] - f[15] + f[20] + f[24] - f[26]) * (R(1.0) / rI);
        myzI = (f[12] - f[17] + f[20] - f[24] + f[26]) * (R(1.0) / rI);
        fam = 0;
        break;
    default:
        return false;
    }
    const R om = omega;
    R rho, mxz, myz;
    if (fam == 0) {
        rho = (R(108) * (R(10) + mxzI * (R(-1) + om) + myzI * (R(-1) + om)) * rI) / (R(874) + om);
        mxz = (R(198) * mxzI * rI + R(18) * myzI * rI - rho) / (R(90.) * rho);
        myz = (R(18) * mxzI * rI + R(198) * myzI * rI - rho) / (R(90.) * rho);
    } else if (fam == 1) {
        rho = (R(108) * (R(10) + mxzI + myzI * (R(-1) + om) - mxzI * om) * rI) / (R(874) + om);
        mxz = (R(198) * mxzI * rI - R(18) * myzI * rI + rho) / (R(90.) * rho);
        myz = -(R(18) * mxzI * rI - R(198) * myzI * rI + rho) / (R(90.) * rho);
    } else if (fam == 2) {
        rho = (R(108) * (R(10) + myzI + mxzI * (R(-1) + om) - myzI * om) * rI) / (R(874) + om);
        mxz = -(R(-198) * mxzI * rI + R(18) * myzI * rI + rho) / (R(90.) * rho);
        myz = (R(-18) * mxzI * rI + R(198) * myzI * rI + rho) / (R(90.) * rho);
    } else {
        rho = (R(-108) * (R(-10) + mxzI * (R(-1) + om) + myzI * (R(-1) + om)) * rI) / (R(874) + om);
        mxz = (R(198) * mxzI * rI + R(18) * myzI * rI + rho) / (R(90.) * rho);
        myz = (R(18) * mxzI * rI + R(198) * myzI * rI + rho) / (R(90.) * rho);
    }
    m[0] = rho;
    m[1] = m[2] = m[3] = 0;
    m[4] = m[5] = m[6] = 0;
    m[7] = 0;
    m[8] = mxz;
    m[9] = myz;
    return true;
}

// fluid_boundary_condition, cases/cylinder/boundaries.cuh:845-1037 (Equilibrium variant,
// BCF_MASS_CONSERV at cases/cylinder/constants.cuh:18). Tags other than 17/34/68/136 leave the
// node unchanged (the switch has no default).
template <class R>
bool bcfluid_bc(uint32_t tag, uint32_t in_mask, R omega, bool strong, const R *f, R *m) {
    if (tag != TAG_BCF_1 && tag != TAG_BCF_2 && tag != TAG_BCF_3 && tag != TAG_BCF_4) return false;
    const R third = R(1) / R(3);
    R rhoI = 0, uxI = 0, uyI = 0, uzI = 0, mxxI = 0, myyI = 0, mzzI = 0, mxyI = 0, mxzI = 0, myzI = 0;
    for (int q = 0; q < Q; ++q) {
        if (!(in_mask & (1u << q))) continue;
        const R fq = f[q];
        rhoI += fq;
        uxI += fq * R(CX[q]);
        uyI += fq * R(CY[q]);
        uzI += fq * R(CZ[q]);
        mxxI += fq * (R(CX[q] * CX[q]) - third);
        myyI += fq * (R(CY[q] * CY[q]) - third);
        mzzI += fq * (R(CZ[q] * CZ[q]) - third);
        mxyI += fq * R(CX[q] * CY[q]);
        mxzI += fq * R(CX[q] * CZ[q]);
        myzI += fq * R(CY[q] * CZ[q]);
    }
    const R inv = R(1) / rhoI;
    uxI *= inv; uyI *= inv; uzI *= inv;
    mxxI *= inv; myyI *= inv; mzzI *= inv; mxyI *= inv; mxzI *= inv; myzI *= inv;

    // The four tags are mirror images of each other: sx, sy are the signs applied to x and y.
    R sx = 1, sy = 1;
    if (tag == TAG_BCF_2) { sx = -1; sy = 1; }
    if (tag == TAG_BCF_3) { sx = 1; sy = -1; }
    if (tag == TAG_BCF_4) { sx = -1; sy = -1; }
    R rho;
    {
        // Literal coefficients of the Equilibrium quadratic, boundaries.cuh:905-915 (TAG_BCF_1),
        // :942-951 (2), :978-987 (3), :1014-1023 (4); tags 2-4 equal tag 1 with
        // u_x -> sx u_x, u_y -> sy u_y, m_xy -> sx sy m_xy. The Strong quadratic
        // (BCF_MASS_CONSERV == MassBC::Strong, boundaries.cuh:894-904, 930-940, 966-976,
        // 1002-1012) mirrors the same way; for tags 34 and 68 the reference writes it multiplied
        // by -1, which swaps the labels of the two roots, so a, b, c carry that sign here too.
        const R ux = sx * uxI, uy = sy * uyI, mxy = sx * sy * mxyI;
        const R poly = R(45) * (mxxI * mxxI) + R(405) * (mxy * mxy) + R(45) * (myyI * myyI) + R(345) * myyI * ux +
                       R(589) * (ux * ux) + R(345) * myyI * uy + R(1467) * ux * uy + R(589) * (uy * uy) +
                       R(45) * mxy * (R(6) * myyI + R(23) * (ux + uy)) +
                       R(15) * mxxI * (R(18) * mxy + R(6) * myyI + R(23) * (ux + uy));
        R a, b, c;
        if (strong) {
            const R sg = (tag == TAG_BCF_2 || tag == TAG_BCF_3) ? R(-1) : R(1);
            a = sg * (R(-9996) - R(206) * omega);
            b = sg * (R(-6) * rhoI *
                      (R(12) * mxxI * (R(-17) + R(20) * omega) -
                       R(17) * (R(102) + R(36) * mxy + R(12) * myyI - R(5) * ux - R(5) * uy) +
                       omega * (R(720) * mxy + R(240) * myyI + R(53) * (ux + uy))));
            c = sg * (R(3) * omega * (rhoI * rhoI) * poly);
        } else {
            a = R(-9998);
            b = R(6) * rhoI * (R(1734) + R(66) * mxxI + R(198) * mxy + R(66) * myyI + R(253) * ux + R(253) * uy);
            c = R(3) * (rhoI * rhoI) * poly;
        }
        const R disc = std::max(b * b - R(4) * a * c, R(0));
        const R rho1 = (-b + std::sqrt(disc)) / (R(2) * a);
        const R rho2 = (-b - std::sqrt(disc)) / (R(2) * a);
        rho = (rho1 > R(0) && std::fabs(rho1 - rhoI) < std::fabs(rho2 - rhoI)) ? rho1 : rho2;
    }
    // Rational closed forms for u and m, boundaries.cuh:916-924 and mirrors.
    const R ux = sx * uxI, uy = sy * uyI, mxy = sx * sy * mxyI, mxz = sx * mxzI, myz = sy * myzI;
    const R uxo = (R(3) * mxxI * rhoI + R(9) * mxy * rhoI + R(3) * myyI * rhoI + R(20) * rhoI * ux + R(3) * rhoI * uy + rho) / (R(17) * rho);
    const R uyo = (R(3) * mxxI * rhoI + R(9) * mxy * rhoI + R(3) * myyI * rhoI + R(3) * rhoI * ux + R(20) * rhoI * uy + rho) / (R(17) * rho);
    const R uzo = (R(3) * rhoI * (mxz + myz + R(10) * uzI)) / (R(29) * rho);
    const R mxxo = (R(57) * mxxI * rhoI + R(18) * mxy * rhoI + R(6) * myyI * rhoI + R(6) * rhoI * ux + R(6) * rhoI * uy + R(2) * rho) / (R(51) * rho);
    const R myyo = (R(6) * mxxI * rhoI + R(18) * mxy * rhoI + R(57) * myyI * rhoI + R(6) * rhoI * ux + R(6) * rhoI * uy + R(2) * rho) / (R(51) * rho);
    const R mzzo = (R(36) * mzzI * rhoI) / (R(35) * rho);
    const R mxyo = (R(3) * mxxI * rhoI + R(26) * mxy * rhoI + R(3) * myyI * rhoI + R(3) * rhoI * ux + R(3) * rhoI * uy + rho) / (R(17) * rho);
    const R mxzo = (rhoI * (R(32) * mxz + R(3) * myz + uzI)) / (R(29) * rho);
    const R myzo = (rhoI * (R(3) * mxz + R(32) * myz + uzI)) / (R(29) * rho);
    m[0] = rho;
    m[1] = sx * uxo;
    m[2] = sy * uyo;
    m[3] = uzo;
    m[4] = mxxo;
    m[5] = myyo;
    m[6] = mzzo;
    m[7] = sx * sy * mxyo;
    m[8] = sx * mxzo;
    m[9] = sy * myzo;
    return true;
}

// evaluate_incoming_moments_rotated, boundary/curvedLBM.cu:5-59. Writes rho and the rotated
// incoming second moments m' (6 values); u is left to the caller (unchanged at wall nodes).
template <class R>
void incoming_rotated(R nx, R ny, uint32_t in_mask, const R *f, R &rho, R *m6) {
    const R c = nx, s = ny, third = R(1) / R(3);
    R r = 0, axx = 0, ayy = 0, azz = 0, axy = 0, axz = 0, ayz = 0;
    for (int q = 0; q < Q; ++q) {
        if (!(in_mask & (1u << q))) continue;
        const R cx = R(CX[q]), cy = R(CY[q]), cz = R(CZ[q]);
        const R xp = cx * c + cy * s, yp = -cx * s + cy * c, zp = cz;
        r += f[q];
        axx += f[q] * (xp * xp - third);
        ayy += f[q] * (yp * yp - third);
        azz += f[q] * (zp * zp - third);
        axy += f[q] * (xp * yp);
        axz += f[q] * (xp * zp);
        ayz += f[q] * (yp * zp);
    }
    const R inv = R(1) / r;
    rho = r;
    m6[0] = axx * inv;
    m6[1] = ayy * inv;
    m6[2] = azz * inv;
    m6[3] = axy * inv;
    m6[4] = axz * inv;
    m6[5] = ayz * inv;
}

// extrapolation_quadratic, boundary/extrapolation_utils.cuh:49-61 (coefficients kept literal;
// their delta^2 terms differ in sign from the Lagrange extrapolation, see D6).
template <class R>
void extrap_coeffs(R delta, R *c3) {
    const R d = delta / R(1), d2 = d * d;  // delx = 1
    c3[0] = R(1) - R(0.5) * d2 + R(1.5) * d;
    c3[1] = d2 - R(2) * d;
    c3[2] = -R(0.5) * (d2 - d);
}

// numerical_solution_rhoeq_rotated, boundary/numerical_solution.cu:4-122. prime7 holds the
// extrapolated wall-frame u'x, u'y, u'z, m'xx, m'yy, m'zz, m'yz; out gets rho, u, m in the
// lattice frame.
template <class R>
void curved_solution(R nx, R ny, uint32_t in_mask, const R *p7, R rhoI, R mxyI, R mxzI, R *out) {
    const R c = nx, s = ny, third = R(1) / R(3);
    const R s2 = R(2) * s * c, c2 = c * c - s * s;
    const R uxp = p7[0], uyp = p7[1], uzp = p7[2], mxxp = p7[3], myyp = p7[4], mzzp = p7[5], myzp = p7[6];
    R E = 0, Dxy = 0, Dxz = 0;
    R F11xy = 0, F22xy = 0, F33xy = 0, F12xy = 0, F13xy = 0, F23xy = 0;
    R F11xz = 0, F22xz = 0, F33xz = 0, F12xz = 0, F13xz = 0, F23xz = 0;
    const R base = R(0.5) * R(3) * R(3);
    for (int q = 0; q < Q; ++q) {
        if (!(in_mask & (1u << q))) continue;
        const R cx = R(CX[q]), cy = R(CY[q]), cz = R(CZ[q]);
        const R xp = cx * c + cy * s, yp = -cx * s + cy * c, zp = cz;
        const R Hxx = xp * xp - third, Hyy = yp * yp - third, Hzz = zp * zp - third;
        const R Hxy = xp * yp, Hxz = xp * zp, Hyz = yp * zp;
        const R wq = R(weight(q));
        const R cf = base * wq;
        const R A = wq * (R(1) + R(3) * (uxp * xp + uyp * yp + uzp * zp));
        E += A + cf * (uxp * uxp * Hxx + uyp * uyp * Hyy + uzp * uzp * Hzz + R(2) * uxp * uyp * Hxy +
                       R(2) * uxp * uzp * Hxz + R(2) * uyp * uzp * Hyz);
        Dxy += A * Hxy;
        Dxz += A * Hxz;
        F11xy += (cf * Hxx) * Hxy; F22xy += (cf * Hyy) * Hxy; F33xy += (cf * Hzz) * Hxy;
        F12xy += (cf * Hxy) * Hxy; F13xy += (cf * Hxz) * Hxy; F23xy += (cf * Hyz) * Hxy;
        F11xz += (cf * Hxx) * Hxz; F22xz += (cf * Hyy) * Hxz; F33xz += (cf * Hzz) * Hxz;
        F12xz += (cf * Hxy) * Hxz; F13xz += (cf * Hxz) * Hxz; F23xz += (cf * Hyz) * Hxz;
    }
    const R Rxy = E * mxyI - Dxy, Rxz = E * mxzI - Dxz;
    const R a1 = R(2) * F12xy, b1 = R(2) * F13xy;
    const R c1 = Rxy - mxxp * F11xy - myyp * F22xy - mzzp * F33xy - R(2) * myzp * F23xy;
    const R a2 = R(2) * F12xz, b2 = R(2) * F13xz;
    const R cc2 = Rxz - mxxp * F11xz - myyp * F22xz - mzzp * F33xz - R(2) * myzp * F23xz;
    const R inv = R(1) / (a2 * b1 - a1 * b2);
    const R mxyp = (b1 * cc2 - b2 * c1) * inv;
    const R mxzp = (a2 * c1 - a1 * cc2) * inv;
    out[1] = uxp * c - uyp * s;
    out[2] = uxp * s + uyp * c;
    out[3] = uzp;
    out[4] = mxxp * c * c + myyp * s * s - mxyp * s2;
    out[5] = mxxp * s * s + myyp * c * c + mxyp * s2;
    out[6] = mzzp;
    out[7] = (mxxp - myyp) * R(0.5) * s2 + mxyp * c2;
    out[8] = mxzp * c - myzp * s;
    out[9] = mxzp * s + myzp * c;
    out[0] = rhoI / E;
}

// numerical_solution_strong_rotated, boundary/numerical_solution.cu:124-279 (MassBC::Strong):
// A', E' and the B' tensor summed over the outgoing directions, D' and F' over the incoming ones;
// F' is corrected by (1 - omega) m'_I B', the 2x2 solve uses G' = (1 - omega) A' + omega E', and
// rho = rho_I / (G' + (1 - omega) m' : B').
template <class R>
void curved_solution_strong(R nx, R ny, uint32_t in_mask, uint32_t out_mask, R omega, const R *p7, R rhoI, R mxyI,
                            R mxzI, R *out) {
    const R c = nx, s = ny, third = R(1) / R(3);
    const R s2 = R(2) * s * c, c2 = c * c - s * s;
    const R uxp = p7[0], uyp = p7[1], uzp = p7[2], mxxp = p7[3], myyp = p7[4], mzzp = p7[5], myzp = p7[6];
    R A = 0, E = 0, B11 = 0, B22 = 0, B33 = 0, B12 = 0, B13 = 0, B23 = 0, Dxy = 0, Dxz = 0;
    R F11xy = 0, F22xy = 0, F33xy = 0, F12xy = 0, F13xy = 0, F23xy = 0;
    R F11xz = 0, F22xz = 0, F33xz = 0, F12xz = 0, F13xz = 0, F23xz = 0;
    const R base = R(0.5) * R(3) * R(3);
    for (int q = 0; q < Q; ++q) {
        const R cx = R(CX[q]), cy = R(CY[q]), cz = R(CZ[q]);
        const R xp = cx * c + cy * s, yp = -cx * s + cy * c, zp = cz;
        const R Hxx = xp * xp - third, Hyy = yp * yp - third, Hzz = zp * zp - third;
        const R Hxy = xp * yp, Hxz = xp * zp, Hyz = yp * zp;
        const R wq = R(weight(q));
        const R cf = base * wq;
        const R Ai = wq * (R(1) + R(3) * (uxp * xp + uyp * yp + uzp * zp));
        if (out_mask & (1u << q)) {
            A += Ai;
            E += Ai + cf * (uxp * uxp * Hxx + uyp * uyp * Hyy + uzp * uzp * Hzz + R(2) * uxp * uyp * Hxy +
                            R(2) * uxp * uzp * Hxz + R(2) * uyp * uzp * Hyz);
            B11 += cf * Hxx; B22 += cf * Hyy; B33 += cf * Hzz;
            B12 += cf * Hxy; B13 += cf * Hxz; B23 += cf * Hyz;
        }
        if (in_mask & (1u << q)) {
            Dxy += Ai * Hxy;
            Dxz += Ai * Hxz;
            F11xy += (cf * Hxx) * Hxy; F22xy += (cf * Hyy) * Hxy; F33xy += (cf * Hzz) * Hxy;
            F12xy += (cf * Hxy) * Hxy; F13xy += (cf * Hxz) * Hxy; F23xy += (cf * Hyz) * Hxy;
            F11xz += (cf * Hxx) * Hxz; F22xz += (cf * Hyy) * Hxz; F33xz += (cf * Hzz) * Hxz;
            F12xz += (cf * Hxy) * Hxz; F13xz += (cf * Hxz) * Hxz; F23xz += (cf * Hyz) * Hxz;
        }
    }
    const R om1 = R(1) - omega;
    const R G = om1 * A + omega * E;
    const R oxy = om1 * mxyI, oxz = om1 * mxzI;
    F11xy -= oxy * B11; F22xy -= oxy * B22; F33xy -= oxy * B33;
    F12xy -= oxy * B12; F13xy -= oxy * B13; F23xy -= oxy * B23;
    F11xz -= oxz * B11; F22xz -= oxz * B22; F33xz -= oxz * B33;
    F12xz -= oxz * B12; F13xz -= oxz * B13; F23xz -= oxz * B23;
    const R Rxy = G * mxyI - Dxy, Rxz = G * mxzI - Dxz;
    const R a1 = R(2) * F12xy, b1 = R(2) * F13xy;
    const R c1 = Rxy - mxxp * F11xy - myyp * F22xy - mzzp * F33xy - R(2) * myzp * F23xy;
    const R a2 = R(2) * F12xz, b2 = R(2) * F13xz;
    const R cc2 = Rxz - mxxp * F11xz - myyp * F22xz - mzzp * F33xz - R(2) * myzp * F23xz;
    const R inv = R(1) / (a2 * b1 - a1 * b2);
    const R mxyp = (b1 * cc2 - b2 * c1) * inv;
    const R mxzp = (a2 * c1 - a1 * cc2) * inv;
    out[1] = uxp * c - uyp * s;
    out[2] = uxp * s + uyp * c;
    out[3] = uzp;
    out[4] = mxxp * c * c + myyp * s * s - mxyp * s2;
    out[5] = mxxp * s * s + myyp * c * c + mxyp * s2;
    out[6] = mzzp;
    out[7] = (mxxp - myyp) * R(0.5) * s2 + mxyp * c2;
    out[8] = mxzp * c - myzp * s;
    out[9] = mxzp * s + myzp * c;
    const R den = G + om1 * (mxxp * B11 + myyp * B22 + mzzp * B33 + R(2) * mxyp * B12 + R(2) * mxzp * B13 +
                             R(2) * myzp * B23);
    out[0] = rhoI / den;
}

// ---------------------------------------------------------------- initial-state hash

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double uniform01(uint64_t seed, uint64_t counter) {
    return double(splitmix64(seed ^ splitmix64(counter)) >> 11) * 0x1.0p-53;
}

// surface_pressure_extrapolation (boundary/extrapolation_utils.cuh:109-154): pressures
// p_k = rho_k cs2 at three points along the wall normal, a quadratic in the distance from the
// cylinder centre through them, evaluated at the wall radius.
double surface_pressure_extrapolation(double xc, double yc, double xw, double yw, double x1, double y1, double x2,
                                      double y2, double x3, double y3, double rho1, double rho2, double rho3) {
    const double cs2 = 1.0 / 3.0;
    const double xw_diff = xw - xc, yw_diff = yw - yc;
    const double x1_diff = x1 - xc, y1_diff = y1 - yc;
    const double x2_diff = x2 - xc, y2_diff = y2 - yc;
    const double x3_diff = x3 - xc, y3_diff = y3 - yc;
    const double rw2 = xw_diff * xw_diff + yw_diff * yw_diff;
    const double r12 = x1_diff * x1_diff + y1_diff * y1_diff;
    const double r22 = x2_diff * x2_diff + y2_diff * y2_diff;
    const double r32 = x3_diff * x3_diff + y3_diff * y3_diff;
    const double rw = std::sqrt(rw2), r1 = std::sqrt(r12), r2 = std::sqrt(r22), r3 = std::sqrt(r32);
    const double denom = (r1 - r2) * (r1 - r3) * (r2 - r3);
    const double p1 = rho1 * cs2, p2 = rho2 * cs2, p3 = rho3 * cs2;
    const double a0 =
        (r1 * r3 * p2 * (r3 - r1) + (r2 * r2) * (r3 * p1 - r1 * p3) + r2 * ((r1 * r1) * p3 - (r3 * r3) * p1)) / denom;
    const double a1 = ((r3 * r3) * (p1 - p2) + (r1 * r1) * (p2 - p3) + (r2 * r2) * (p3 - p1)) / denom;
    const double a2 = (r3 * (p2 - p1) + r2 * (p1 - p3) + r1 * (p3 - p2)) / denom;
    return a0 + a1 * rw + a2 * (rw * rw);
}

// ---------------------------------------------------------------- simulation

struct WallGroup {
    std::vector<int64_t> node;
    std::vector<double> nx, ny, delta;
    std::vector<uint32_t> in, out;
    std::vector<int64_t> bcf;
    std::vector<uint32_t> bcf_tag;
    double radius = 0;
    double uw[3] = {0, 0, 0};
    double fluid_dir = 1;  // -1 for the outer wall (curvedLBM.cu:72-76)
};

struct Base {
    virtual ~Base() = default;
    mrlbm_case_desc d{};
    int NX = 0, NY = 0, NZ = 0;
    int64_t N = 0;
    std::vector<uint32_t> type, tag;  // reference node type, and tag (BCFLUID) or list index
    std::vector<uint8_t> active;      // node belongs to an active reference tile
    WallGroup grp[2];
    int ngroups = 0;
    uint32_t bcf_in[256] = {0}, bcf_out[256] = {0};
    std::vector<double> inlet;        // prescribed inlet u_x per (y, z)
    std::vector<std::array<double, 4>> forces;
    std::vector<std::array<double, 2>> planes;  // inlet mean density, outlet-plane mean u_x
    int64_t steps = 0;
    int active_tiles = 0, total_tiles = 0;
    int64_t idx(int x, int y, int z) const { return x + int64_t(NX) * (y + int64_t(NY) * z); }

    virtual int init() = 0;
    virtual int step(int n, uint32_t flags) = 0;
    virtual void get(int f, double *h) const = 0;
    virtual void set(int f, const double *h) = 0;
    virtual int surface_pressure(double *out) const = 0;
    int setup();
};

// boundary_definitions, cases/cylinder/boundaries.cuh:8-94.
uint32_t boundary_definition(const mrlbm_case_desc &d, int x, int y, int z) {
    bool W = x == 0, E = x == d.nx - 1, S = y == 0, Nn = y == d.ny - 1, B = z == 0, F = z == d.nz - 1;
    if (d.periodic_x) W = E = false;
    if (d.periodic_y) S = Nn = false;
    if (d.periodic_z) B = F = false;
    if (Nn && W && F) return T_NWF;
    if (Nn && W && B) return T_NWB;
    if (Nn && E && F) return T_NEF;
    if (Nn && E && B) return T_NEB;
    if (S && W && F) return T_SWF;
    if (S && W && B) return T_SWB;
    if (S && E && F) return T_SEF;
    if (S && E && B) return T_SEB;
    if (Nn && W) return T_NORTH_WEST;
    if (Nn && E) return T_NORTH_EAST;
    if (Nn && F) return T_NORTH_FRONT;
    if (Nn && B) return T_NORTH_BACK;
    if (S && W) return T_SOUTH_WEST;
    if (S && E) return T_SOUTH_EAST;
    if (S && F) return T_SOUTH_FRONT;
    if (S && B) return T_SOUTH_BACK;
    if (W && F) return T_WEST_FRONT;
    if (W && B) return T_WEST_BACK;
    if (E && F) return T_EAST_FRONT;
    if (E && B) return T_EAST_BACK;
    if (Nn) return T_NORTH;
    if (S) return T_SOUTH;
    if (W) return T_WEST;
    if (E) return T_EAST;
    if (F) return T_FRONT;
    if (B) return T_BACK;
    return T_BULK;
}

int Base::setup() {
    NX = d.nx; NY = d.ny; NZ = d.nz;
    N = int64_t(NX) * NY * NZ;
    type.assign(N, T_BULK);
    tag.assign(N, 0);
    active.assign(N, 1);
    // initialize_nodeType, solver/initializeLBM.cuh:58-80
    for (int z = 0; z < NZ; ++z)
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) type[idx(x, y, z)] = boundary_definition(d, x, y, z);
    if (d.kind == MRLBM_CASE_PERIODIC_BOX) {
        std::fill(type.begin(), type.end(), T_BULK);
    }
    if (d.kind == MRLBM_CASE_CYLINDER || d.kind == MRLBM_CASE_ANNULUS) {
        const bool ann = d.kind == MRLBM_CASE_ANNULUS;
        ngroups = ann ? 2 : 1;
        grp[0].radius = d.r_inner;
        grp[1].radius = d.r_outer;
        for (int k = 0; k < 3; ++k) { grp[0].uw[k] = d.u_wall_inner[k]; grp[1].uw[k] = d.u_wall_outer[k]; }
        grp[1].fluid_dir = -1;
        auto radius = [&](int x, int y) {
            const double dx = double(x) - d.xc, dy = double(y) - d.yc;
            return std::sqrt(dx * dx + dy * dy);
        };
        // mark_cylinder_solid (nodeClass.cuh:110-125) / mark_annulus_solid (annulus
        // nodeClass.cuh:109-126)
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y)
                for (int x = 0; x < NX; ++x) {
                    const double r = radius(x, y);
                    if (r <= d.r_inner || (ann && r >= d.r_outer)) type[idx(x, y, z)] = T_SOLID;
                }
        // scan window: the cylinder scans its bounding box +4 (nodeClass.cuh:174-178), the annulus
        // the whole domain
        int x0 = 0, x1 = NX, y0 = 0, y1 = NY;
        if (!ann) {
            x0 = std::max(0, int(std::ceil(d.xc - d.r_inner)) - 4);
            x1 = std::min(NX, int(std::floor(d.xc + d.r_inner)) + 5);
            y0 = std::max(0, int(std::ceil(d.yc - d.r_inner)) - 4);
            y1 = std::min(NY, int(std::floor(d.yc + d.r_inner)) + 5);
        }
        const double r_mid = 0.5 * (d.r_inner + d.r_outer);
        // load_neighbors (nodeClass.cuh:27-60): x/y outside the domain count as SOLID; z wraps
        // when periodic, else outside counts as SOLID too
        auto neighbour = [&](int x, int y, int z, int q) -> uint32_t {
            const int xn = x + CX[q], yn = y + CY[q];
            int zn = z + CZ[q];
            if (d.periodic_z) {
                if (zn < 0) zn += NZ;
                if (zn >= NZ) zn -= NZ;
            } else if (zn < 0 || zn >= NZ) {
                return T_SOLID;
            }
            if (xn < 0 || xn >= NX || yn < 0 || yn >= NY) return T_SOLID;
            return type[idx(xn, yn, zn)];
        };
        // wall and BCFLUID scans skip the end planes without z periodicity (nodeClass.cuh:171-172,
        // 220-221; the end-plane branch at :202 is never reached)
        const int zb = d.periodic_z ? 0 : 1, ze = d.periodic_z ? NZ : NZ - 1;
        auto octant_bits = [&](int x, int y, int z, uint32_t want, int &count) {
            uint32_t bits = 0;
            count = 0;
            for (int b = 0; b < 8; ++b)
                for (int i = 0; i < 7; ++i)
                    if (neighbour(x, y, z, OCTANT_DIRS[b][i]) == want) {
                        bits |= 1u << b;
                        ++count;
                        break;
                    }
            return bits;
        };
        auto any_neighbour = [&](int x, int y, int z, uint32_t want) {
            for (int q = 1; q < Q; ++q)
                if (neighbour(x, y, z, q) == want) return true;
            return false;
        };
        // classify_boundary_nodes_triangular (nodeClass.cuh:169-216): SOLID with BULK
        // neighbours in more than two octants becomes a wall node
        for (int z = zb; z < ze; ++z)
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    const int64_t i = idx(x, y, z);
                    if (type[i] != T_SOLID || !any_neighbour(x, y, z, T_BULK)) continue;
                    int count = 0;
                    octant_bits(x, y, z, T_BULK, count);
                    if (count > 2) type[i] = (ann && radius(x, y) >= r_mid) ? T_OUTER : T_INNER;
                }
        // classify_bcfluid_nodes_triangular (nodeClass.cuh:218-252): BULK with a SOLID
        // neighbour; tag = octants holding SOLID
        for (int z = zb; z < ze; ++z)
            for (int y = y0; y < y1; ++y)
                for (int x = x0; x < x1; ++x) {
                    const int64_t i = idx(x, y, z);
                    if (type[i] != T_BULK || !any_neighbour(x, y, z, T_SOLID)) continue;
                    int count = 0;
                    const uint32_t bits = octant_bits(x, y, z, T_SOLID, count);
                    type[i] = (ann && radius(x, y) >= r_mid) ? T_BCF_OUTER : T_BCF_INNER;
                    tag[i] = bits;
                }
        // classify_bcsolid_nodes_triangular (cylinder nodeClass.cuh:254-287, annulus
        // nodeClass.cuh:309-357): without z periodicity, an end-wall (BACK/FRONT) node with BULK
        // neighbours in fewer than four octants becomes BCSOLID, tag = those octants; the annulus
        // splits inner/outer at the mid radius
        if (!d.periodic_z) {
            for (int z : {0, NZ - 1})
                for (int y = y0; y < y1; ++y)
                    for (int x = x0; x < x1; ++x) {
                        const int64_t i = idx(x, y, z);
                        if (type[i] != T_BACK && type[i] != T_FRONT) continue;
                        int count = 0;
                        const uint32_t bits = octant_bits(x, y, z, T_BULK, count);
                        if (count < 4) {
                            type[i] = (ann && radius(x, y) >= r_mid) ? T_BCS_OUTER : T_BCS_INNER;
                            tag[i] = bits;
                        }
                    }
        }
        // buildBoundaryList_updateBoundaryNodeType + assignBoundaryIndices
        // (cylinder_helpers.cuh:49-132): lists in scan order z, y, x
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y)
                for (int x = 0; x < NX; ++x) {
                    const int64_t i = idx(x, y, z);
                    const uint32_t t = type[i];
                    if (t == T_INNER || t == T_OUTER) {
                        WallGroup &g = grp[t == T_OUTER ? 1 : 0];
                        tag[i] = uint32_t(g.node.size());
                        g.node.push_back(i);
                    } else if (t == T_BCF_INNER || t == T_BCF_OUTER) {
                        WallGroup &g = grp[t == T_BCF_OUTER ? 1 : 0];
                        g.bcf.push_back(i);
                        g.bcf_tag.push_back(tag[i]);
                    }
                }
        for (int gi = 0; gi < ngroups; ++gi) {
            WallGroup &g = grp[gi];
            if (g.node.empty()) return fail(MRLBM_ERR_GEOMETRY, "oracle: empty wall list");
            for (int64_t i : g.node) {
                const int x = int(i % NX), y = int((i / NX) % NY), z = int(i / (int64_t(NX) * NY));
                // compute_unit_vectors_boundary_nodes (cylinder_helpers.cuh:10-47)
                const double dx = double(x) - d.xc, dy = double(y) - d.yc;
                const double inv_r = 1.0 / std::sqrt(dx * dx + dy * dy);
                const double nx = dx * inv_r, ny = dy * inv_r;
                const double xw = d.xc + g.radius * nx, yw = d.yc + g.radius * ny;
                g.nx.push_back(nx);
                g.ny.push_back(ny);
                g.delta.push_back(std::fabs((xw - double(x)) * nx + (yw - double(y)) * ny));
                // find_incomings_outgoings (cylinder_helpers.cuh:134-228)
                uint32_t in = 1u;
                for (int b = 0; b < 8; ++b) {
                    bool valid = true;
                    for (int k = 0; k < 7; ++k)
                        if (neighbour(x, y, z, OCTANT_DIRS[b][k]) == T_SOLID) { valid = false; break; }
                    if (!valid) continue;
                    for (int k = 0; k < 7; ++k) in |= 1u << opposite(OCTANT_DIRS[b][k]);
                }
                uint32_t out = 1u;
                for (int q = 0; q < Q; ++q)
                    if (in & (1u << opposite(q))) out |= 1u << q;
                g.in.push_back(in);
                g.out.push_back(out);
            }
        }
        // setup_bcfluid_masks (cylinder_helpers.cuh:230-299): per tag, from the first node of
        // the list carrying it; the annulus runs it for the inner then the outer list, so an
        // outer node overwrites a tag the inner list already set.
        for (int gi = 0; gi < ngroups; ++gi) {
            bool computed[256] = {false};
            WallGroup &g = grp[gi];
            for (size_t k = 0; k < g.bcf.size(); ++k) {
                const uint32_t t = g.bcf_tag[k];
                if (computed[t]) continue;
                const int64_t i = g.bcf[k];
                const int x = int(i % NX), y = int((i / NX) % NY), z = int(i / (int64_t(NX) * NY));
                uint32_t in = (1u << Q) - 1;
                for (int q = 0; q < Q; ++q)
                    if (neighbour(x, y, z, q) == T_SOLID) in &= ~(1u << opposite(q));
                uint32_t out = 0;
                for (int q = 0; q < Q; ++q)
                    if (in & (1u << opposite(q))) out |= 1u << q;
                bcf_in[t] = in;
                bcf_out[t] = out;
                computed[t] = true;
            }
        }
        // inlet velocity table (cylinder WEST nodes)
        inlet.assign(int64_t(NY) * NZ, 0.0);
        const double H = double(NY - 1);
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y) {
                double u = d.u_max;
                if (d.inlet_profile == MRLBM_INLET_PARABOLIC) u = 4.0 * d.u_max * double(y) * (H - double(y)) / (H * H);
                const double r = uniform01(d.seed, uint64_t(3 * N) + uint64_t(y) + uint64_t(NY) * z);
                inlet[y + int64_t(NY) * z] = u * (1.0 + d.inlet_noise * (2.0 * r - 1.0));
            }
    }
    // find_active_blocks_3d (simulation.cu:174-214) on 8x8x4 tiles; the kernels look the mask
    // up with blockIdx.z ignored (mlbm.cu:87), i.e. the bz = 0 layer decides every bz.
    const int GX = (NX + 7) / 8, GY = (NY + 7) / 8, GZ = (NZ + 3) / 4;
    std::vector<uint8_t> act(size_t(GX) * GY, 0);
    for (int by = 0; by < GY; ++by)
        for (int bx = 0; bx < GX; ++bx) {
            bool fluid = false;
            for (int tz = 0; tz < 4 && !fluid; ++tz)
                for (int ty = 0; ty < 8 && !fluid; ++ty)
                    for (int tx = 0; tx < 8 && !fluid; ++tx) {
                        const int x = bx * 8 + tx, y = by * 8 + ty, z = tz;
                        if (x >= NX || y >= NY || z >= NZ) continue;
                        if (type[idx(x, y, z)] != T_SOLID) fluid = true;
                    }
            act[bx + size_t(GX) * by] = fluid;
        }
    total_tiles = GX * GY * GZ;
    active_tiles = 0;
    for (uint8_t a : act) active_tiles += a * GZ;
    for (int z = 0; z < NZ; ++z)
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) active[idx(x, y, z)] = act[x / 8 + size_t(GX) * (y / 8)];
    return MRLBM_OK;
}

template <class R>
struct Sim : Base {
    std::vector<R> M[10];  // reference convention: M[0] = rho - rho0
    std::vector<R> old[10];
    R omega = 0;

    void load(const std::vector<R> *src, int64_t i, R *m) const {
        m[0] = R(d.rho0) + src[0][i];
        for (int k = 1; k < 10; ++k) m[k] = src[k][i];
    }
    void store(std::vector<R> *dst, int64_t i, const R *m) {
        dst[0][i] = m[0] - R(d.rho0);
        for (int k = 1; k < 10; ++k) dst[k][i] = m[k];
    }

    int init() override {
        for (auto &v : M) v.assign(N, R(0));
        for (int64_t i = 0; i < N; ++i) {
            double u[3];
            for (int a = 0; a < 3; ++a)
                u[a] = d.init_noise * (2.0 * uniform01(d.seed, uint64_t(a) * uint64_t(N) + uint64_t(i)) - 1.0);
            const bool m0 = d.compat_flags & MRLBM_COMPAT_INIT_M_ZERO;
            M[0][i] = R(0);
            M[1][i] = R(u[0]);
            M[2][i] = R(u[1]);
            M[3][i] = R(u[2]);
            M[4][i] = R(m0 ? 0.0 : u[0] * u[0]);
            M[5][i] = R(m0 ? 0.0 : u[1] * u[1]);
            M[6][i] = R(m0 ? 0.0 : u[2] * u[2]);
            M[7][i] = R(m0 ? 0.0 : u[0] * u[1]);
            M[8][i] = R(m0 ? 0.0 : u[0] * u[2]);
            M[9][i] = R(m0 ? 0.0 : u[1] * u[2]);
        }
        steps = 0;
        forces.clear();
        return MRLBM_OK;
    }

    void get(int f, double *h) const override {
        for (int64_t i = 0; i < N; ++i) h[i] = double(M[f][i]);
    }
    void set(int f, const double *h) override {
        for (int64_t i = 0; i < N; ++i) M[f][i] = R(h[i]);
    }

    // Bilinear interpolation of field k at (x, y) in plane z (extrapolation_utils.cuh:5-47).
    R bilinear(const std::vector<R> *src, int k, R x, R y, int z) const {
        const int x0 = int(std::floor(x)), y0 = int(std::floor(y));
        const int x1 = x0 + 1, y1 = y0 + 1;
        const R xd = x - R(x0), yd = y - R(y0);
        const R q00 = src[k][idx(x0, y0, z)], q10 = src[k][idx(x1, y0, z)];
        const R q01 = src[k][idx(x0, y1, z)], q11 = src[k][idx(x1, y1, z)];
        const R q0 = (R(1) - xd) * q00 + xd * q10;
        const R q1 = (R(1) - xd) * q01 + xd * q11;
        return (R(1) - yd) * q0 + yd * q1;
    }

    // compute_surface_pressure (cases/cylinder/cylinder_kernels.cu:238-283) without the running
    // mean: the instantaneous wall pressure of every inner-wall node from the current moments,
    // with fluid points 1, 2 and 3 spacings out along the normal.
    int surface_pressure(double *out) const override {
        if (ngroups < 1) return fail(MRLBM_ERR_INVALID, "oracle: no curved wall");
        const WallGroup &g = grp[0];
        for (size_t k = 0; k < g.node.size(); ++k) {
            const int64_t i = g.node[k];
            const int z = int(i / (int64_t(NX) * NY));
            const double nx = g.nx[k], ny = g.ny[k];
            const double xw = d.xc + g.radius * nx, yw = d.yc + g.radius * ny;
            double xs[3], ys[3], rho[3];
            for (int p = 0; p < 3; ++p) {
                xs[p] = xw + double(p + 1) * 1.0 * nx;
                ys[p] = yw + double(p + 1) * 1.0 * ny;
                if (!in_bounds(R(xs[p]), R(ys[p]))) return fail(MRLBM_ERR_GEOMETRY, "oracle: pressure stencil leaves the grid");
                rho[p] = d.rho0 + double(bilinear(M, 0, R(xs[p]), R(ys[p]), z));
            }
            out[k] = surface_pressure_extrapolation(d.xc, d.yc, xw, yw, xs[0], ys[0], xs[1], ys[1], xs[2], ys[2], rho[0],
                                                    rho[1], rho[2]);
        }
        return MRLBM_OK;
    }

    // initerpolate_and_rotate + rotate_velocity_moments (extrapolation_utils.cuh:63-107).
    void interp_rotate(const std::vector<R> *src, R nx, R ny, R x, R y, int z, R *vm) const {
        R v[9];
        for (int k = 0; k < 9; ++k) v[k] = bilinear(src, k + 1, x, y, z);
        const R ux = v[0], uy = v[1], uz = v[2], mxx = v[3], myy = v[4], mzz = v[5], mxy = v[6], mxz = v[7], myz = v[8];
        const R c = nx, s = ny, s2 = R(2) * s * c, c2 = c * c - s * s;
        vm[0] = ux * c + uy * s;
        vm[1] = -ux * s + uy * c;
        vm[2] = uz;
        vm[3] = mxx * c * c + myy * s * s + mxy * s2;
        vm[4] = mxx * s * s + myy * c * c - mxy * s2;
        vm[5] = mzz;
        vm[6] = mxy * c2 + R(0.5) * (myy - mxx) * s2;
        vm[7] = mxz * c + myz * s;
        vm[8] = myz * c - mxz * s;
    }

    bool in_bounds(R x, R y) const {
        const int x0 = int(std::floor(x)), y0 = int(std::floor(y));
        return x0 >= 0 && y0 >= 0 && x0 + 1 < NX && y0 + 1 < NY;
    }

    void force_sum(const R *m, uint32_t mask, double sign, double *acc) const {
        for (int q = 0; q < Q; ++q) {
            if (!(mask & (1u << q))) continue;
            const double fq = double(pop_q(m, q));
            acc[0] += sign * fq * CX[q];
            acc[1] += sign * fq * CY[q];
            acc[2] += sign * fq * CZ[q];
            acc[3] += sign * fq;
        }
    }

    void force_pass(double sign, bool incoming, double *acc) const {
        // compute_force_mass_kernel (statistics/forces.cu:25-133): wall list then BCFLUID list
        R m[10];
        for (int gi = 0; gi < ngroups; ++gi) {
            const WallGroup &g = grp[gi];
            for (size_t k = 0; k < g.node.size(); ++k) {
                load(M, g.node[k], m);
                force_sum(m, incoming ? g.in[k] : g.out[k], sign, acc);
            }
            for (size_t k = 0; k < g.bcf.size(); ++k) {
                load(M, g.bcf[k], m);
                const uint32_t t = g.bcf_tag[k];
                force_sum(m, incoming ? bcf_in[t] : bcf_out[t], sign, acc);
            }
        }
    }

    int step(int nsteps, uint32_t flags) override {
        for (int s = 0; s < nsteps; ++s) {
            int rc = step1(flags);
            if (rc) return rc;
        }
        return MRLBM_OK;
    }

    int step1(uint32_t flags) {
        for (int k = 0; k < 10; ++k) old[k] = M[k];
        const bool cyl = d.kind == MRLBM_CASE_CYLINDER;
        // ---- K1: streaming_and_evaluate_Mom (solver/mlbm.cu:83-164)
#pragma omp parallel for schedule(static)
        for (int z = 0; z < NZ; ++z) {
            R f[Q], m[10], mo[10], mx[10];
            for (int y = 0; y < NY; ++y)
                for (int x = 0; x < NX; ++x) {
                    const int64_t i = idx(x, y, z);
                    if (!active[i]) continue;
                    for (int q = 0; q < Q; ++q) {  // periodic pull in all three directions
                        const int xs = (x - CX[q] + NX) % NX, ys = (y - CY[q] + NY) % NY, zs = (z - CZ[q] + NZ) % NZ;
                        R ms[10];
                        load(old, idx(xs, ys, zs), ms);
                        f[q] = pop_q(ms, q);
                    }
                    load(old, i, mo);
                    std::copy(mo, mo + 10, m);
                    const uint32_t t = type[i];
                    if (t == T_BULK) {
                        moments(f, m);
                    } else if (t == T_INNER || t == T_OUTER) {
                        const WallGroup &g = grp[t == T_OUTER ? 1 : 0];
                        const uint32_t k = tag[i];
                        incoming_rotated(R(g.nx[k]), R(g.ny[k]), g.in[k], f, m[0], m + 4);
                        // m holds (rho_I, u_old, rotated incoming m'); store order xx yy zz xy xz yz
                    } else if (t == T_BCF_INNER || t == T_BCF_OUTER) {
                        bcfluid_bc(tag[i], bcf_in[tag[i]], omega, bool(d.compat_flags & MRLBM_COMPAT_BCF_MASS_STRONG),
                                   f, m);
                    } else if (t == T_BCS_INNER || t == T_BCS_OUTER) {
                        bcsolid_bc(tag[i], omega, f, m);
                    } else if (cyl) {
                        // the outlet types read the x-1 neighbour's moments
                        if (t == T_EAST || t == T_NORTH_EAST || t == T_SOUTH_EAST || t == T_EAST_FRONT ||
                            t == T_EAST_BACK || t == T_NEF || t == T_NEB || t == T_SEF || t == T_SEB)
                            load(old, idx(x - 1, y, z), mx);
                        face_bc(t, omega, R(inlet.empty() ? 0.0 : inlet[y + int64_t(NY) * z]), f, mo, mx, m);
                    }
                    store(M, i, m);
                }
        }
        // ---- K3/K4: curved wall, Jacobi (every stencil reads the K1 state)
        for (int gi = 0; gi < ngroups; ++gi) {
            WallGroup &g = grp[gi];
            std::vector<std::array<R, 10>> res(g.node.size());
            for (size_t k = 0; k < g.node.size(); ++k) {
                const int64_t i = g.node[k];
                const int x = int(i % NX), y = int((i / NX) % NY), z = int(i / (int64_t(NX) * NY));
                (void)x; (void)y;
                const R nx = R(g.nx[k]), ny = R(g.ny[k]), delta = R(g.delta[k]);
                const R xw = R(d.xc) + R(g.radius) * nx, yw = R(d.yc) + R(g.radius) * ny;
                const R dir = R(g.fluid_dir);
                const R x1 = xw + dir * R(1) * nx, y1 = yw + dir * R(1) * ny;
                const R x2 = xw + dir * R(2) * R(1) * nx, y2 = yw + dir * R(2) * R(1) * ny;
                if (!in_bounds(x1, y1) || !in_bounds(x2, y2))
                    return fail(MRLBM_ERR_GEOMETRY, "oracle: wall stencil outside the domain");
                R vm1[9], vm2[9];
                interp_rotate(M, nx, ny, x1, y1, z, vm1);
                interp_rotate(M, nx, ny, x2, y2, z, vm2);
                const R uw[3] = {R(g.uw[0]), R(g.uw[1]), R(g.uw[2])};
                // wall values: u'_w and m'_w = u'_w u'_w (curvedLBM.cu:93-101)
                const R vw[7] = {uw[0], uw[1], uw[2], uw[0] * uw[0], uw[1] * uw[1], uw[2] * uw[2], uw[1] * uw[2]};
                const int sel[7] = {0, 1, 2, 3, 4, 5, 8};
                R c3[3];
                extrap_coeffs(delta, c3);
                R p7[7];
                for (int j = 0; j < 7; ++j) p7[j] = c3[0] * vw[j] + c3[1] * vm1[sel[j]] + c3[2] * vm2[sel[j]];
                R mi[10];
                load(M, i, mi);
                R out[10];
                if (d.compat_flags & MRLBM_COMPAT_MASS_STRONG)
                    curved_solution_strong(nx, ny, g.in[k], g.out[k], omega, p7, mi[0], mi[7], mi[8], out);
                else
                    curved_solution(nx, ny, g.in[k], p7, mi[0], mi[7], mi[8], out);
                std::copy(out, out + 10, res[k].begin());
            }
            for (size_t k = 0; k < g.node.size(); ++k) store(M, g.node[k], res[k].data());
        }
        double acc[4] = {0, 0, 0, 0};
        const bool want_forces = (flags & MRLBM_STEP_FORCES) && ngroups > 0;
        if (want_forces) force_pass(+1.0, true, acc);
        // ---- K2: collision_halo_update (solver/mlbm.cu:186-242), every node of active tiles
        const R om = omega;
#pragma omp parallel for schedule(static)
        for (int z = 0; z < NZ; ++z) {
            R m[10];
            for (int y = 0; y < NY; ++y)
                for (int x = 0; x < NX; ++x) {
                    const int64_t i = idx(x, y, z);
                    if (!active[i]) continue;
                    load(M, i, m);
                    collide(om, m);
                    store(M, i, m);
                }
        }
        if (want_forces) {
            force_pass(-1.0, false, acc);
            forces.push_back({acc[0], acc[1], acc[2], acc[3]});
        } else {
            forces.push_back({0.0, 0.0, 0.0, 0.0});
        }
        // inlet_avg_density (cylinder_kernels.cu:285-338): mean of RHO_0 + rho over x = 0;
        // outlet_avg_ux (mlbm.cu:18-81): mean u_x over x = NX - 2; both of the state after the step
        std::array<double, 2> pm = {0.0, 0.0};
        if ((flags & MRLBM_STEP_STATS) && d.kind == MRLBM_CASE_CYLINDER) {
            for (int z = 0; z < NZ; ++z)
                for (int y = 0; y < NY; ++y) {
                    pm[0] += d.rho0 + double(M[0][idx(0, y, z)]);
                    pm[1] += double(M[1][idx(NX - 2, y, z)]);
                }
            pm[0] /= double(NY) * double(NZ);
            pm[1] /= double(NY) * double(NZ);
        }
        planes.push_back(pm);
        ++steps;
        return MRLBM_OK;
    }
};

int check_desc(const mrlbm_case_desc &d) {
    if (d.nx < 3 || d.ny < 3 || d.nz < 1) return fail(MRLBM_ERR_INVALID, "oracle: bad grid");
    if (!(d.tau > 0.5 && d.tau < 2.5)) return fail(MRLBM_ERR_TAU, "oracle: tau out of range");
    if (d.u_max / std::sqrt(1.0 / 3.0) > 0.33) return fail(MRLBM_ERR_MACH, "oracle: Mach too high");
    return MRLBM_OK;
}

}  // namespace

struct oracle_sim {
    std::unique_ptr<Base> impl;
};

extern "C" {

int oracle_create(const mrlbm_case_desc *desc, int32_t real_bytes, oracle_sim **out) {
    if (!desc || !out) return fail(MRLBM_ERR_INVALID, "oracle: null argument");
    int rc = check_desc(*desc);
    if (rc) return rc;
    auto s = new oracle_sim;
    if (real_bytes == 4) {
        auto p = new Sim<float>;
        p->omega = float(1.0 / desc->tau);
        s->impl.reset(p);
    } else {
        auto p = new Sim<double>;
        p->omega = 1.0 / desc->tau;
        s->impl.reset(p);
    }
    s->impl->d = *desc;
    rc = s->impl->setup();
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return MRLBM_OK;
}

int oracle_init(oracle_sim *s) { return s->impl->init(); }
int oracle_surface_pressure(oracle_sim *s, double *out) { return s->impl->surface_pressure(out); }
double oracle_surface_pressure_extrapolation(double xc, double yc, const double *xy8, const double *rho3) {
    return surface_pressure_extrapolation(xc, yc, xy8[0], xy8[1], xy8[2], xy8[3], xy8[4], xy8[5], xy8[6], xy8[7],
                                          rho3[0], rho3[1], rho3[2]);
}
int oracle_step(oracle_sim *s, int32_t n, uint32_t flags) { return s->impl->step(n, flags); }

int oracle_get_moments(oracle_sim *s, int32_t f, double *h) {
    if (f < 0 || f >= 10) return fail(MRLBM_ERR_INVALID, "oracle: bad field");
    s->impl->get(f, h);
    return MRLBM_OK;
}
int oracle_set_moments(oracle_sim *s, int32_t f, const double *h) {
    if (f < 0 || f >= 10) return fail(MRLBM_ERR_INVALID, "oracle: bad field");
    s->impl->set(f, h);
    return MRLBM_OK;
}

int oracle_get_forces(oracle_sim *s, int64_t first, int32_t count, double *out4) {
    auto &F = s->impl->forces;
    if (first < 0 || first + count > int64_t(F.size())) return fail(MRLBM_ERR_INVALID, "oracle: force range");
    for (int32_t k = 0; k < count; ++k)
        for (int j = 0; j < 4; ++j) out4[4 * k + j] = F[first + k][j];
    return MRLBM_OK;
}

int oracle_get_plane_means(oracle_sim *s, int64_t first, int32_t count, double *out2) {
    auto &P = s->impl->planes;
    if (first < 0 || first + count > int64_t(P.size())) return fail(MRLBM_ERR_INVALID, "oracle: plane range");
    for (int32_t k = 0; k < count; ++k)
        for (int j = 0; j < 2; ++j) out2[2 * k + j] = P[first + k][j];
    return MRLBM_OK;
}

int oracle_get_node_types(oracle_sim *s, uint32_t *out) {
    const Base &b = *s->impl;
    for (int64_t i = 0; i < b.N; ++i) out[i] = (b.type[i] << 20) | (b.tag[i] & ((1u << 20) - 1));
    return MRLBM_OK;
}

int oracle_get_wall_table(oracle_sim *s, int32_t group, int64_t *node, double *nx, double *ny, double *delta,
                          uint32_t *in_mask, uint32_t *out_mask) {
    const Base &b = *s->impl;
    if (group < 0 || group >= 2) return fail(MRLBM_ERR_INVALID, "oracle: bad group");
    const WallGroup &g = b.grp[group];
    for (size_t k = 0; k < g.node.size(); ++k) {
        node[k] = g.node[k];
        nx[k] = g.nx[k];
        ny[k] = g.ny[k];
        delta[k] = g.delta[k];
        in_mask[k] = g.in[k];
        out_mask[k] = g.out[k];
    }
    return MRLBM_OK;
}

int oracle_get_bcfluid_table(oracle_sim *s, int32_t group, int64_t *node, uint32_t *tag, uint32_t *in256,
                             uint32_t *out256) {
    const Base &b = *s->impl;
    if (group < 0 || group >= 2) return fail(MRLBM_ERR_INVALID, "oracle: bad group");
    const WallGroup &g = b.grp[group];
    for (size_t k = 0; k < g.bcf.size(); ++k) {
        node[k] = g.bcf[k];
        tag[k] = g.bcf_tag[k];
    }
    for (int t = 0; t < 256; ++t) {
        in256[t] = b.bcf_in[t];
        out256[t] = b.bcf_out[t];
    }
    return MRLBM_OK;
}

int oracle_get_info(oracle_sim *s, mrlbm_info *info) {
    const Base &b = *s->impl;
    std::memset(info, 0, sizeof(*info));
    info->num_nodes = b.N;
    info->nb_inner = int32_t(b.grp[0].node.size());
    info->nb_outer = int32_t(b.grp[1].node.size());
    info->nb_fluid_inner = int32_t(b.grp[0].bcf.size());
    info->nb_fluid_outer = int32_t(b.grp[1].bcf.size());
    info->active_tiles = b.active_tiles;
    info->total_tiles = b.total_tiles;
    info->steps_done = b.steps;
    info->precision = b.d.precision;
    return MRLBM_OK;
}

int oracle_destroy(oracle_sim *s) {
    delete s;
    return MRLBM_OK;
}

// ---- VTK output, src/output/saveData.cu. `real` is double in the FP64 build, so RHO_0 + rho
// is a double sum rounded to float (:105-106, :211-212), velocities are rounded to float and
// NaN at SOLID nodes (:107-112, :220-221), and get_cylindrical (:180-208) runs in float with
// dx = (float)x - XC evaluated in double (XC is a double constant) and rounded on assignment.
void oracle_vti_points(int32_t layout, int32_t nx, int32_t ny, int32_t nz, double rho0, double xc, double yc,
                       double vel_norm, const double *rho_dev, const double *ux, const double *uy,
                       const double *uz, const uint8_t *solid, float *out4) {
    const size_t n = size_t(nx) * size_t(ny) * size_t(nz);
    float *rho_out = out4, *a1 = out4 + n, *a2 = out4 + 2 * n, *a3 = out4 + 3 * n;
    auto cylindrical = [&](size_t idx, int x, int y, bool get_theta) -> float {
        if (solid[idx]) return NAN;
        const float vx = float(ux[idx]);
        const float vy = float(uy[idx]);
        const float dx = float(double(float(x)) - xc);
        const float dy = float(double(float(y)) - yc);
        const float r2 = dx * dx + dy * dy;
        if (r2 < 1e-12f) return 0.0f;
        const float inv_r = 1.0f / std::sqrt(r2);
        const float cos_t = dx * inv_r;
        const float sin_t = dy * inv_r;
        if (get_theta) return (-vx * sin_t + vy * cos_t) / float(vel_norm);
        return (vx * cos_t + vy * sin_t) / float(vel_norm);
    };
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t i = size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * size_t(z));
                rho_out[i] = float(rho0 + rho_dev[i]);
                if (layout == 0) {
                    a1[i] = solid[i] ? NAN : float(ux[i]);
                    a2[i] = solid[i] ? NAN : float(uy[i]);
                    a3[i] = solid[i] ? NAN : float(uz[i]);
                } else {
                    a1[i] = solid[i] ? NAN : float(uz[i]);
                    a2[i] = cylindrical(i, x, y, true);
                    a3[i] = cylindrical(i, x, y, false);
                }
            }
}

// write_vti_3d / write_vti_3d_annulus file layout (saveData.cu:42-116, 118-230): XML header with
// four appended Float32 arrays at offsets k * (8 + 4n), '_', then per array its UInt64 byte
// count and the values in x-fastest order, then the closing tags.
int oracle_write_vti(const char *path, int32_t layout, int32_t nx, int32_t ny, int32_t nz, const float *pts4) {
    static const char *names[2][4] = {{"rho", "ux", "uy", "uz"}, {"rho", "uz", "u_theta", "u_radial"}};
    std::ofstream f(path, std::ios::binary);
    if (!f) return fail(MRLBM_ERR_IO, "oracle: cannot open vti file");
    const uint64_t points = uint64_t(nx) * ny * nz;
    const uint64_t bytes = points * sizeof(float);
    f << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
      << "  <ImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1
      << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
      << "    <Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nz - 1 << "\">\n"
      << "      <PointData>\n";
    uint64_t offset = 0;
    for (int k = 0; k < 4; ++k) {
        f << "        <DataArray type=\"Float32\" Name=\"" << names[layout ? 1 : 0][k]
          << "\" format=\"appended\" offset=\"" << offset << "\"/>\n";
        offset += 8 + bytes;
    }
    f << "      </PointData>\n    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n_";
    for (int k = 0; k < 4; ++k) {
        f.write(reinterpret_cast<const char *>(&bytes), 8);
        f.write(reinterpret_cast<const char *>(pts4 + size_t(k) * points), std::streamsize(bytes));
    }
    f << "\n  </AppendedData>\n</VTKFile>\n";
    return f ? MRLBM_OK : fail(MRLBM_ERR_IO, "oracle: vti write failed");
}

// write_master_pvd (saveData.cu:3-40): one DataSet per saved step, timestep = iter / MACR_SAVE.
int oracle_write_master_pvd(const char *path, int32_t max_iter, int32_t save_every) {
    if (save_every <= 0) return fail(MRLBM_ERR_INVALID, "oracle: save interval must be positive");
    std::ofstream out(path);
    if (!out) return fail(MRLBM_ERR_IO, "oracle: cannot open pvd file");
    out << "<VTKFile type=\"Collection\" version=\"0.1\">\n\t<Collection>\n";
    for (int iter = 0; iter <= max_iter; iter++)
        if (iter % save_every == 0)
            out << "\t\t<DataSet timestep=\"" << iter / save_every << "\" file=\"data" << (10000000 + iter)
                << ".vti\" />\n";
    out << "\t</Collection>\n</VTKFile>\n";
    return out ? MRLBM_OK : fail(MRLBM_ERR_IO, "oracle: pvd write failed");
}

int32_t oracle_num_threads(void) { return omp_get_max_threads(); }

// ---- per-node building blocks in double, for closed-form tests
void oracle_pop_reconstruction(const double *m10, double *f27) {
    for (int q = 0; q < Q; ++q) f27[q] = pop_q(m10, q);
}
void oracle_eval_moments(const double *f27, double *m10) { moments(f27, m10); }
void oracle_collide(double omega, double *m10) { collide(omega, m10); }
int oracle_node_bc(uint32_t type, uint32_t tag, uint32_t in_mask, double omega, double u_in, const double *f27,
                   const double *old10, const double *old10_xm1, double *out10) {
    std::copy(old10, old10 + 10, out10);
    if (type == T_BCF_INNER || type == T_BCF_OUTER) return bcfluid_bc(tag, in_mask, omega, false, f27, out10) ? 1 : 0;
    if (type == T_BCS_INNER || type == T_BCS_OUTER) return bcsolid_bc(tag, omega, f27, out10) ? 1 : 0;
    return face_bc(type, omega, u_in, f27, old10, old10_xm1, out10) ? 1 : 0;
}
void oracle_incoming_rotated(double nx, double ny, uint32_t in_mask, const double *f27, double *rho, double *m6) {
    incoming_rotated(nx, ny, in_mask, f27, *rho, m6);
}
void oracle_extrapolation_coeffs(double delta, double *c3) { extrap_coeffs(delta, c3); }
void oracle_curved_solution_strong(double nx, double ny, uint32_t in_mask, uint32_t out_mask, double omega,
                                   const double *p7, double rho_i, double mxy_i, double mxz_i, double *out10) {
    curved_solution_strong(nx, ny, in_mask, out_mask, omega, p7, rho_i, mxy_i, mxz_i, out10);
}
int oracle_bcfluid(uint32_t tag, uint32_t in_mask, double omega, int32_t strong, const double *f27, double *out10) {
    for (int k = 0; k < 10; ++k) out10[k] = 0.0;
    return bcfluid_bc(tag, in_mask, omega, strong != 0, f27, out10) ? 1 : 0;
}
void oracle_curved_solution(double nx, double ny, uint32_t in_mask, const double *p7, double rho_i, double mxy_i,
                            double mxz_i, double *out10) {
    curved_solution(nx, ny, in_mask, p7, rho_i, mxy_i, mxz_i, out10);
}

}  // extern "C"
