// kernels.h - device kernels of one time step (fused stream/collide and curved wall + forces).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "geometry.h"
#include "node_math.h"
#include "storage.h"

namespace mrlbm {

// Arguments of the fused stream/collide pass over one ping-pong buffer pair.
template <class R, class S>
struct FusedArgs {
    const S *__restrict__ src;  // 10 fields, field f at src + f * stride
    S *__restrict__ dst;
    S *__restrict__ donor;  // pre-collision snapshot of donor nodes (same layout)
    size_t stride;
    int nx, ny, nz;
    int tiles_x;     // tiles along x
    int nblocks;     // work items, one block each
    int xcd_groups;  // blocks are spread round-robin over this many XCDs (1: no remap)
    // work item b: x = tile index | (special << 31), y = z0 | (z1 << 16); special items (tiles
    // with boundary nodes, which need info[]) come first and march shorter z ranges
    const uint2 *__restrict__ items;
    const uint32_t *__restrict__ info;  // per-node NodeCode | flags | payload
    const R *__restrict__ wall_nx;
    const R *__restrict__ wall_ny;
    const uint32_t *__restrict__ wall_in;
    const uint32_t *__restrict__ bcf_in;  // 256 per-tag incoming masks
    const R *__restrict__ inlet;          // u_x per (y, z)
    // steps completed on the device (advanced by wall_step); seeds stochastic rounding of the
    // 16-bit forms, null for round-to-nearest
    const unsigned long long *__restrict__ step_counter;
    StoreScale<R> scale;
    R omega, rho0;
};

// Arguments of the curved-wall + force pass (one thread per wall node, then one per BCFLUID).
template <class R, class S>
struct WallArgs {
    S *__restrict__ mom;          // buffer the fused pass just wrote
    const S *__restrict__ donor;  // its pre-collision snapshot
    size_t stride;
    int nwall, nb_inner, nbcf;
    const uint32_t *__restrict__ node;
    const R *__restrict__ wnx;
    const R *__restrict__ wny;
    const R *__restrict__ coef;        // 3 extrapolation coefficients per wall
    const R *__restrict__ wconst;      // WALL_CONSTS closure constants per wall
    const uint32_t *__restrict__ in_mask;
    const uint32_t *__restrict__ out_mask;
    const uint32_t *__restrict__ st_base;  // 2 per wall: index of the (x0, y0) stencil corner
    const R *__restrict__ st_frac;         // 4 per wall: xd1, yd1, xd2, yd2
    int nx;                                // row pitch of the stencil corners
    const uint32_t *__restrict__ bcf_node;
    const uint8_t *__restrict__ bcf_tag;
    const uint32_t *__restrict__ bcf_in;
    const uint32_t *__restrict__ bcf_out;
    R uw[2][3];
    StoreScale<R> scale;
    R omega, rho0;
    int forces;
    double *__restrict__ partials;  // 4 per block
    unsigned *__restrict__ ticket;
    double *__restrict__ ring;      // 4 per step, FORCE_RING_STEPS records
    unsigned long long *__restrict__ step_counter;  // steps completed on the device
};

// Steps of force history kept in the device ring.
constexpr int FORCE_RING_STEPS = 1 << 16;

// variant 0: register-window kernel (default); 1: LDS ring-buffer kernel
template <class R, class S>
void launch_fused(const FusedArgs<R, S> &a, int bx, int by, int variant, hipStream_t stream);

template <class R, class S>
void launch_wall(const WallArgs<R, S> &a, hipStream_t stream);

template <class R, class S>
void launch_encode(const double *in, S *out, size_t stride, uint32_t n, const StoreScale<R> &sc, hipStream_t stream);

template <class R, class S>
void launch_decode(const S *in, size_t stride, double *out, uint32_t n, const StoreScale<R> &sc, hipStream_t stream);

// Tile shapes (BX x BY) the fused kernel is instantiated for.
inline bool fused_tile_supported(int bx, int by) {
    return (bx == 32 && by == 8) || (bx == 16 && by == 8) || (bx == 32 && by == 4) || (bx == 16 && by == 16) ||
           (bx == 64 && by == 4) || (bx == 16 && by == 4);
}

// Bytes of dynamic LDS one fused block uses over the tile and its ring: one plane of 27 pulled
// populations (variant 0), or three planes of the c_z=+1 group, two of the c_z=0 group and one of
// the c_z=-1 group (variant 1).
inline size_t fused_lds_bytes(int bx, int by, size_t real_bytes, int variant = 0) {
    return size_t(variant == 1 ? 54 : 27) * size_t(bx + 2) * size_t(by + 2) * real_bytes;
}

}  // namespace mrlbm
