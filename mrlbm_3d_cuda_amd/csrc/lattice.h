// lattice.h - D3Q27 velocity set shared by host setup and device kernels.
//
// The velocity ordering is the reference's (core/D3Q27/latticeProperties.cuh:17-21): 0 rest,
// 1-6 axes, 7-18 face diagonals, 19-26 corners, with q and q+1 opposite for odd q. Wall masks
// and BCFLUID tags are bit sets over this ordering, so it is part of the interface.
#pragma once

#include <cstdint>

namespace mrlbm {

constexpr int Q = 27;

struct LatticeTables {
    int cx[Q], cy[Q], cz[Q], opp[Q];
};

constexpr LatticeTables make_lattice() {
    // (x, y, z) of the odd member of each opposite pair; the even member is its negation
    constexpr int odd[13][3] = {{1, 0, 0},  {0, 1, 0},  {0, 0, 1}, {1, 1, 0},  {1, 0, 1},
                                {0, 1, 1},  {1, -1, 0}, {1, 0, -1}, {0, 1, -1}, {1, 1, 1},
                                {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}};
    LatticeTables t{};
    t.cx[0] = t.cy[0] = t.cz[0] = 0;
    t.opp[0] = 0;
    for (int k = 0; k < 13; ++k) {
        const int q = 2 * k + 1;
        t.cx[q] = odd[k][0];
        t.cy[q] = odd[k][1];
        t.cz[q] = odd[k][2];
        t.cx[q + 1] = -odd[k][0];
        t.cy[q + 1] = -odd[k][1];
        t.cz[q + 1] = -odd[k][2];
        t.opp[q] = q + 1;
        t.opp[q + 1] = q;
    }
    return t;
}

constexpr LatticeTables LATTICE = make_lattice();

// Weights as the product of the 1-D weights, w = w1(cx) w1(cy) w1(cz) with w1(0) = 2/3 and
// w1(+-1) = 1/6 (8/27, 2/27, 1/54, 1/216 of latticeProperties.cuh:7-10). Evaluated left to right
// in double, so every user of the table gets the same rounded values.
constexpr double weight1(int c) { return c == 0 ? 2.0 / 3.0 : 1.0 / 6.0; }
constexpr double weight(int q) { return weight1(LATTICE.cx[q]) * weight1(LATTICE.cy[q]) * weight1(LATTICE.cz[q]); }
constexpr double W_REST = weight(0);
constexpr double W_AXIS = weight(1);
constexpr double W_FACE = weight(7);
constexpr double W_CORNER = weight(19);

// Populations grouped by the z component of their velocity. A population with c_z = +1 at
// plane z was pulled from plane z-1, c_z = 0 from plane z, c_z = -1 from plane z+1.
constexpr int GROUP_ZP[9] = {5, 9, 11, 16, 18, 19, 22, 23, 25};
constexpr int GROUP_Z0[9] = {0, 1, 2, 3, 4, 7, 8, 13, 14};
constexpr int GROUP_ZM[9] = {6, 10, 12, 15, 17, 20, 21, 24, 26};

}  // namespace mrlbm
