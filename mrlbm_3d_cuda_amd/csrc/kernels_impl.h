// kernels_impl.h - the kernels of one time step (instantiated by kernels_f64.hip and
// kernels_f32.hip).
//
// fused_step: one pass per node that replaces the reference's streaming_and_evaluate_Mom and
// collision_halo_update (solver/mlbm.cu:83-242). A block owns a BX x BY column of nodes and
// marches it through a z range. For every plane it loads the 10 moments of the tile and a
// one-node ring around it, reconstructs the 27 regularised populations of each loaded node
// into LDS, pulls the 27 populations of its own nodes from the neighbours' LDS entries and
// recollides. Populations are kept in LDS grouped by c_z: the c_z=+1 group of a plane is used
// one plane later, the c_z=0 group in the same plane and the c_z=-1 group one plane earlier,
// so three planes of the +1 group, two of the 0 group and one of the -1 group are live.
// The ghost-plane exchange of the reference (core/D3Q27/halo_interface.cuh) becomes the ring
// load of neighbour moments through HBM/L2 and the z march.
//
// wall_step: curved-wall closure of the wall nodes (apply_bc_cylinder / apply_bc_annulus,
// cases/cylinder/cylinder_kernels.cu:68-128, cases/annulus/annulus_kernels.cu:75-139) with
// deferred collision, and the momentum-exchange force / mass-flux reduction
// (statistics/forces.cu:25-133) in FP64, as per-block partials that the next launch on the stream
// (one block of the next fused pass, or record_kernel) sums into the per-step record.
#pragma once

#include <algorithm>
#include <vector>

#include "kernels.h"

namespace mrlbm {

namespace {

// Barrier that orders LDS only: global loads issued before it may stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Files the pending force / mass record of a case: the wall blocks' partials summed in a fixed
// order (lane t adds blocks t, t + 64, ..., then the wave tree) into the ring slot of the step
// that left them. Run by one whole wave, in the launch after that wall pass (stream order makes
// the partials visible to plain loads); the pending mark is then cleared.
__device__ __forceinline__ void file_record(const RecordArgs &r) {
    const unsigned long long pend = *r.pending;
    if (pend == 0) return;
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    for (int blk = int(threadIdx.x); blk < r.nblocks; blk += 64)
        for (int k = 0; k < 4; ++k) tot[k] += r.partials[4 * blk + k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot[k] += __shfl_xor(tot[k], off, 64);
    if (threadIdx.x == 0) {
        const size_t slot = size_t((pend - 1ull) % FORCE_RING_STEPS);
#pragma unroll
        for (int k = 0; k < 4; ++k) r.ring[4 * slot + k] = tot[k];
        *r.pending = 0;
    }
}

// LDS slot of plane zp (zp >= -1) among the three planes of west-ring-column moments. Plane z is
// written in the iteration before the one that reads it; with one barrier per plane (two-plane
// window march) a fast wave may already write plane z+2 while a slow one still reads plane z, so
// the two must not share a slot.
__device__ __forceinline__ int west_slot(int zp) { return (zp + 3) % 3; }

__device__ __forceinline__ int wrap(int v, int n) {
    v = v < 0 ? v + n : v;
    return v >= n ? v - n : v;
}

// Frame of a tile: the BX x BY tile nodes and a one-node ring around them, as a (BX+2) x (BY+2)
// grid of positions. Thread t owns tile node (t % BX, t / BX) as its first position; the ring
// positions -- south row, north row, west column, east column -- are the second position of the
// first NR threads.
template <int BX, int BY>
struct Frame {
    static constexpr int W = BX + 2;
    static constexpr int NH = W * (BY + 2);
    static constexpr int NT = BX * BY;
    static constexpr int NR = 2 * W + 2 * BY;
    __device__ static void ring(int r, int &hx, int &hy) {
        if (r < W) {
            hx = r;
            hy = 0;
        } else if (r < 2 * W) {
            hx = r - W;
            hy = BY + 1;
        } else if (r < 2 * W + BY) {
            hx = 0;
            hy = r - 2 * W + 1;
        } else {
            hx = W - 1;
            hy = r - 2 * W - BY + 1;
        }
    }
};

// v mod n in [0, n)
__device__ __forceinline__ int pmod(int v, int n) {
    const int r = v % n;
    return r < 0 ? r + n : r;
}

// NaN/Inf watchdog: a wave in which some node's new density is not finite raises the case's flag
// (every lane of such a wave stores the same 1). A non-finite moment anywhere reaches the density of
// its own node within a step: the rest population is part of the node's density sum.
template <class R, class S>
__device__ __forceinline__ void watch_finite(const FusedArgs<R, S> &a, R rho) {
    if (__builtin_expect(__any(!__builtin_isfinite(rho)), 0))
        if (a.nonfinite) *a.nonfinite = 1u;
}

// Collision of a node that is stored next (do_collide false: stored as it is). The 16-bit forms
// store m - u u, which the collision scales by 1 - omega: m* - u u = (1 - omega)(m - u u). They
// keep that non-equilibrium part in M[4..9] (two operations per moment instead of the collision's
// two and the encoder's subtraction), which store_collided then encodes as it is.
template <class R, class S>
__device__ __forceinline__ void collide_for_store(R omega, bool do_collide, R *M) {
    if constexpr (Conv<S>::scaled) {
        const R fac = do_collide ? R(1) - omega : R(1);
        M[4] = fac * (M[4] - M[1] * M[1]);
        M[5] = fac * (M[5] - M[2] * M[2]);
        M[6] = fac * (M[6] - M[3] * M[3]);
        M[7] = fac * (M[7] - M[1] * M[2]);
        M[8] = fac * (M[8] - M[1] * M[3]);
        M[9] = fac * (M[9] - M[2] * M[3]);
    } else {
        if (do_collide) collide(omega, M);
    }
}

// Store of a node finished by collide_for_store.
template <class R, class S>
__device__ __forceinline__ void store_collided(const FieldSet<S> &fs, uint32_t idx, const StoreScale<R> &sc, R rho0,
                                               const R *M, uint32_t seed) {
    S v[10];
    encode_node<R, S, Conv<S>::scaled>(idx, sc, rho0, M, seed, v);
    store_raw<S>(fs, idx, v);
}

// Per-node table values of a boundary node, loaded before the plane's prefetch with in-range
// indices whatever the node code (so every thread issues the same loads).
template <class R>
struct NodeTabs {
    R wnx, wny, inlet;
    uint32_t win, bcfin;
};

template <class R, class S>
__device__ __forceinline__ void load_tabs(const FusedArgs<R, S> &a, uint32_t info, uint32_t inlet_idx, NodeTabs<R> &t) {
    const uint32_t payload = info >> INFO_PAYLOAD_SHIFT;
    const uint32_t wi = min(payload, uint32_t(a.wall_count - 1));
    t.wnx = a.wall_nx[wi];
    t.wny = a.wall_ny[wi];
    t.win = a.wall_in[wi];
    t.bcfin = a.bcf_in[payload & 0xFFu];
    t.inlet = a.inlet[min(inlet_idx, uint32_t(a.inlet_count - 1))];
}

// Post-stream moments of one node of a boundary tile from its 27 pulled populations f (in the
// LDS form of reconstruct_lds), its moments at the start of the step (old) and those of its x-1
// neighbour (oxm), with the node boundary condition (the per-node part of
// streaming_and_evaluate_Mom and collision_halo_update, solver/mlbm.cu:83-164, 186-242). Returns
// in M the pre-collision moments (the snapshot of donor nodes), in do_collide whether the node
// collides now (not when the wall pass finishes it), and store = false for frozen nodes.
template <class R, class S>
__device__ __forceinline__ void boundary_node(const FusedArgs<R, S> &a, R *f, const R *old, const R *oxm,
                                              uint32_t info, const NodeTabs<R> &t, R *M, bool &store, bool &donor,
                                              bool &do_collide) {
    const uint32_t code = info & INFO_CODE_MASK;
    const uint32_t payload = info >> INFO_PAYLOAD_SHIFT;
    do_collide = true;
    store = true;
    if (code == NC_BULK) {
        moments_lds(f, M);
    } else {
        lds_to_pop(f);
        if (code == NC_FROZEN || code == NC_KEEP) {
#pragma unroll
            for (int k = 0; k < 10; ++k) M[k] = old[k];
            store = code == NC_KEEP;  // frozen: both buffers already hold its value
        } else if (code == NC_WALL) {
            // u keeps its old value; rho and m become the incoming rotated moments, handed to the
            // wall pass through the node's donor snapshot (the wall pass stores the node)
#pragma unroll
            for (int k = 0; k < 10; ++k) M[k] = old[k];
            incoming_rotated(t.wnx, t.wny, t.win, f, M);
            do_collide = false;
            store = false;
        } else if (code == NC_BCF) {
            bc_fluid(payload, t.bcfin, a.omega, a.bcf_strong != 0, f, M);
        } else if (code == NC_WEST) {
            bc_west(f, old, t.inlet, M);
        } else if (code == NC_NORTH) {
            bc_ywall<face::K_N>(f, M);
        } else if (code == NC_SOUTH) {
            bc_ywall<face::K_S>(f, M);
        } else if (code == NC_NORTH_WEST) {
            bc_west_edge<face::K_W & face::K_N, -1>(f, a.omega, M);
        } else if (code == NC_SOUTH_WEST) {
            bc_west_edge<face::K_W & face::K_S, +1>(f, a.omega, M);
        } else if (code == NC_EAST) {
            bc_east(f, old, oxm, M);
        } else if (code == NC_NORTH_EAST) {
            bc_east_edge<face::K_E & face::K_N, +1>(f, old, oxm, M);
        } else if (code == NC_SOUTH_EAST) {
            bc_east_edge<face::K_E & face::K_S, -1>(f, old, oxm, M);
        } else if (code == NC_FRONT) {
            bc_zwall<face::K_F>(f, M);
        } else if (code == NC_BACK) {
            bc_zwall<face::K_B>(f, M);
        } else if (code == NC_NORTH_FRONT) {
            bc_yz_edge<face::K_N & face::K_F, +1>(f, a.omega, M);
        } else if (code == NC_NORTH_BACK) {
            bc_yz_edge<face::K_N & face::K_B, -1>(f, a.omega, M);
        } else if (code == NC_SOUTH_FRONT) {
            bc_yz_edge<face::K_S & face::K_F, -1>(f, a.omega, M);
        } else if (code == NC_SOUTH_BACK) {
            bc_yz_edge<face::K_S & face::K_B, +1>(f, a.omega, M);
        } else if (code == NC_WEST_FRONT) {
            bc_west_edge<face::K_W & face::K_F, -1, 8>(f, a.omega, M);
        } else if (code == NC_WEST_BACK) {
            bc_west_edge<face::K_W & face::K_B, +1, 8>(f, a.omega, M);
        } else if (code == NC_EAST_FRONT) {
            bc_east_edge<face::K_E & face::K_F, +1, 8>(f, old, oxm, M);
        } else if (code == NC_EAST_BACK) {
            bc_east_edge<face::K_E & face::K_B, -1, 8>(f, old, oxm, M);
        } else if (code == NC_WEST_CORNER) {
            bc_west_corner(payload, f, M);
        } else if (code == NC_EAST_CORNER) {
            bc_east_corner(old, oxm, M);
        } else {
            bc_solid(payload, a.omega, f, M);  // NC_BCS
        }
    }
    donor = (store || code == NC_WALL) && (info & INFO_DONOR);
}

// State and steps of the boundary-tile path shared by both marches. Per thread: the stored
// values of its own node for the plane being updated and the next one, the node codes of those
// planes and the table values of the plane being updated. The results of a plane are stored as
// soon as they are computed (the bulk path keeps its results in registers until the next
// iteration; here that would hold 20 more values through the boundary closed forms, which then
// spill). The density and velocity of the west ring column go through LDS (wold, three planes) so
// that a node in the tile's first column can read its x-1 neighbour.
template <class R, class S, int BX, int BY>
struct BoundaryPath {
    static constexpr int W = BX + 2;
    RawNode<S> own_cur, own_nx;
    uint32_t info_cur = 0, info_nx = 0;
    NodeTabs<R> tab;

    __device__ void load_info(const FusedArgs<R, S> &a, uint32_t node, uint32_t &out) { out = a.info[node]; }
    __device__ void copy_own(const RawNode<S> &raw, RawNode<S> &dst) { dst = raw; }
    // the x-1 neighbour's density and velocity at the plane being updated (all the outlet closures
    // read of it): lane - 1 inside the tile row, the west ring column (LDS) for the first column
    __device__ void neighbour_xm1(const R *old, int tx, int ty, const R *wold, R *oxm) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const R v = __shfl(old[k], int(threadIdx.x) - 1, 64);
            oxm[k] = tx > 0 ? v : wold[k * BY + ty];
        }
    }
    // Updates the thread's node of the plane and stores it, with the donor snapshot (the
    // pre-collision moments) first. The snapshot is encoded only in waves that hold a donor node;
    // the others still issue its stores (into the padding slot), so every path keeps the same
    // memory operations.
    __device__ void update(const FusedArgs<R, S> &a, R *f, int tx, int ty, bool inside, bool east_tile, uint32_t node,
                           const R *wold, uint32_t seed) {
        // the node's own and x-1 moments are only read by boundary nodes: waves made of bulk
        // nodes skip their decode
        R old[10], oxm[4];
        if (__any((info_cur & INFO_CODE_MASK) != NC_BULK)) {
            decode_node<R, S>(own_cur, a.scale, a.rho0, old);
#pragma unroll
            for (int k = 0; k < 4; ++k) oxm[k] = old[k];
            if (east_tile) neighbour_xm1(old, tx, ty, wold, oxm);  // uniform per block
        } else {
#pragma unroll
            for (int k = 0; k < 10; ++k) old[k] = R(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) oxm[k] = R(0);
        }
        R M[10];
        bool store, donor, do_collide;
        boundary_node(a, f, old, oxm, info_cur, tab, M, store, donor, do_collide);
        watch_finite(a, M[0]);
        const uint32_t idx = (inside && store) ? node : a.pad;
        const uint32_t didx = (inside && donor) ? node : a.pad;
        S v[10];
        if (__any(didx != a.pad)) {
            encode_node<R, S>(didx, a.scale, a.rho0, M, seed ^ 0x5A5A5A5Au, v);
        } else {
#pragma unroll
            for (int k = 0; k < 10; ++k) v[k] = Conv<S>::from(R(0));
        }
        store_raw_nm(a.donor, didx, v);
        collide_for_store<R, S>(a.omega, do_collide, M);
        store_collided<R, S>(a.dst, idx, a.scale, a.rho0, M, seed);
    }
};

// z march of one work item. Populations of the last planes are kept in LDS grouped by c_z
// (3 planes of the c_z=+1 group, 2 of c_z=0, 1 of c_z=-1). The stored values of the next plane
// are in flight in registers while a plane is updated.
//
// Vector-memory loads and stores share one counter, counted in issue order. Every path of an
// iteration issues the same loads and stores: bulk results of a plane are kept in registers and
// stored at the start of the next iteration (after the wait for the prefetched plane, before the
// next prefetch), boundary-tile results right after the update (after the prefetch), threads
// outside the grid, frozen nodes and non-donors store into the padding slot after the last node,
// second-position loads of threads without one repeat the first, and the prefetch past the item
// wraps around in z. A wait for a prefetched plane therefore never waits for stores or loads
// issued after it.
template <class R, class S, int BX, int BY, bool SPECIAL>
__device__ __forceinline__ void fused_march(const FusedArgs<R, S> &a, R *lds, int tile, int z0, int z1) {
    using F = Frame<BX, BY>;
    constexpr int W = F::W;
    constexpr int NH = F::NH;
    // group planes: zp[3][9][NH], z0[2][9][NH], zm[9][NH]; then the west-column moments of three planes
    R *const gp = lds;
    R *const g0 = lds + 27 * NH;
    R *const gm = lds + 45 * NH;
    R *const wold = lds + 54 * NH;

    const int x0 = (tile % a.tiles_x) * BX;
    const int y0 = (tile / a.tiles_x) * BY;
    const int tid = threadIdx.x;
    const int tx = tid % BX, ty = tid / BX;
    const uint32_t plane = uint32_t(a.nx) * uint32_t(a.ny);
    const int h = (ty + 1) * W + (tx + 1);  // first position: the thread's own node

    const bool has1 = tid < F::NR;
    int hx1 = 0, hy1 = 0;
    if (has1) F::ring(tid, hx1, hy1);
    const int p1 = hy1 * W + hx1;
    const uint32_t off0 = uint32_t(pmod(x0 + tx, a.nx)) + uint32_t(a.nx) * uint32_t(pmod(y0 + ty, a.ny));
    const uint32_t off1 = has1 ? uint32_t(pmod(x0 - 1 + hx1, a.nx)) + uint32_t(a.nx) * uint32_t(pmod(y0 - 1 + hy1, a.ny))
                               : off0;
    const bool west1 = has1 && hx1 == 0 && hy1 >= 1 && hy1 <= BY;  // west ring column

    RawNode<S> m0, m1;  // stored values of the plane in flight, per position
    auto load_plane = [&](int z) {
        const uint32_t base = uint32_t(pmod(z, a.nz)) * plane;
        load_raw(a.src, base + off0, m0);
        load_raw(a.src, base + off1, m1);
    };
    // write the populations of a loaded plane (index zp) into the requested group slots
    auto put = [&](const RawNode<S> &raw, int p, int sp, int s0, bool want0, bool wantm, bool west, int zp) {
        R m[10], f[Q];
        decode_node<R, S>(raw, a.scale, a.rho0, m);
        if (SPECIAL && west) {
#pragma unroll
            for (int k = 0; k < 4; ++k) wold[(west_slot(zp) * 4 + k) * BY + (hy1 - 1)] = m[k];
        }
        reconstruct_lds(m, f);
#pragma unroll
        for (int i = 0; i < 9; ++i) gp[(sp * 9 + i) * NH + p] = f[GROUP_ZP[i]];
        if (want0) {
#pragma unroll
            for (int i = 0; i < 9; ++i) g0[(s0 * 9 + i) * NH + p] = f[GROUP_Z0[i]];
        }
        if (wantm) {
#pragma unroll
            for (int i = 0; i < 9; ++i) gm[i * NH + p] = f[GROUP_ZM[i]];
        }
    };
    auto put_plane = [&](int sp, int s0, bool want0, bool wantm, int zp) {
        put(m0, h, sp, s0, want0, wantm, false, zp);
        if (has1) put(m1, p1, sp, s0, want0, wantm, west1, zp);
    };

    const int x = x0 + tx, y = y0 + ty;
    const bool inside = x < a.nx && y < a.ny;
    const bool east_tile = x0 + BX >= a.nx;  // holds the last x column
    uint32_t seed = 0;  // stochastic rounding of the 16-bit forms
    if constexpr (Conv<S>::scaled) {
        if (a.step_counter) seed = uint32_t(*a.step_counter) * 2u + 1u;
    }
    BoundaryPath<R, S, BX, BY> bp;

    // prologue: plane z0-1 (c_z=+1 group) and plane z0 (c_z=+1 and 0 groups), then z0+1 in flight
    load_plane(z0 - 1);
    put_plane(0, 0, false, false, z0 - 1);
    load_plane(z0);
    put_plane(1, 0, true, false, z0);
    if constexpr (SPECIAL) {
        bp.copy_own(m0, bp.own_cur);
        bp.load_info(a, off0 + uint32_t(z0) * plane, bp.info_cur);
    }
    load_plane(z0 + 1);

    int sp_prev = 0, sp_cur = 1, sp_next = 2;  // slots of planes z-1, z, z+1 in the +1 group
    int s0_cur = 0;
    R Mp[10];  // bulk: results of the previous plane, stored one iteration late
    uint32_t idxp = a.pad;

    auto pull = [&](R *f) {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int q = GROUP_ZP[i];
            f[q] = gp[(sp_prev * 9 + i) * NH + h - LATTICE.cx[q] - LATTICE.cy[q] * W];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int q = GROUP_Z0[i];
            f[q] = g0[(s0_cur * 9 + i) * NH + h - LATTICE.cx[q] - LATTICE.cy[q] * W];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int q = GROUP_ZM[i];
            f[q] = gm[i * NH + h - LATTICE.cx[q] - LATTICE.cy[q] * W];
        }
    };
    auto rotate = [&]() {
        const int t = sp_prev;
        sp_prev = sp_cur;
        sp_cur = sp_next;
        sp_next = t;
        s0_cur ^= 1;
    };

    for (int z = z0; z < z1; ++z) {
        put_plane(sp_next, s0_cur ^ 1, true, true, z + 1);  // plane z+1
        if constexpr (SPECIAL) {
            bp.copy_own(m0, bp.own_nx);
            load_tabs(a, bp.info_cur, uint32_t(y) + uint32_t(a.ny) * uint32_t(z), bp.tab);
            bp.load_info(a, off0 + uint32_t(pmod(z + 1, a.nz)) * plane, bp.info_nx);
        } else {
            store_collided<R, S>(a.dst, idxp, a.scale, a.rho0, Mp, seed);
        }
        load_plane(z + 2);  // prefetch, consumed next iteration (wraps past the item)
        lds_barrier();
        R f[Q];
        pull(f);
        const uint32_t node = off0 + uint32_t(z) * plane;
        if constexpr (SPECIAL) {
            bp.update(a, f, tx, ty, inside, east_tile, node, wold + west_slot(z) * 4 * BY, seed);
        } else {
            moments_lds(f, Mp);
            collide_for_store<R, S>(a.omega, true, Mp);
            watch_finite(a, Mp[0]);
            idxp = inside ? node : a.pad;
        }
        lds_barrier();
        rotate();
        if constexpr (SPECIAL) {
            bp.copy_own(bp.own_nx, bp.own_cur);
            bp.info_cur = bp.info_nx;
        }
    }
    if constexpr (!SPECIAL) store_collided<R, S>(a.dst, idxp, a.scale, a.rho0, Mp, seed);
}

// Registers are capped at four waves per SIMD for the one-plane march of 16-bit storage in 32 x 8
// tiles (128 VGPRs: the plain and face marches fit, the boundary path spills about 44 bytes per
// lane), three waves for the other one-plane marches (about 156 VGPRs in use, no spills: capped at
// four, the boundary-heavy 16 x 8 FP32 and batched forms spilled 64-132 bytes per lane and ran up to
// 60% longer) and two for the two-plane march (512-thread blocks: one per CU). A cap of five waves
// (96 VGPRs) made the plain march spill too.
template <class S, int BX, int BY, int NBUF>
constexpr int window_min_blocks() {
    return BX * BY > 256 ? 1 : (NBUF == 2 ? 2 : ((Conv<S>::scaled && BX * BY == 256) ? 4 : 3));
}

// LDS plane of the register-window march, laid out by the rows each population is pulled from:
// for every (c_x, c_z) a block of the c_y = +1 populations (frame rows 0..BY-1), one spare row, the
// c_y = -1 populations (rows 2..BY+1) and the c_y = 0 populations (rows 1..BY). Population q of
// frame position p sits at off(q) + p. Writes that no node pulls -- the c_y = +1 populations of the
// last tile row and of the west / east ring column's last row, the c_y = -1 ones of the first rows --
// land in the spare row between the c_y = +1 and c_y = -1 groups of the block.
template <int BX, int BY>
struct WindowLds {
    static constexpr int W = BX + 2;
    static constexpr int NP = BY * W;     // one population's values of a plane
    static constexpr int B = 3 * NP + W;  // one (c_x, c_z) block
    static constexpr int PLANE = 9 * B;   // one LDS plane
    static constexpr int cyoff(int cy) { return cy > 0 ? 0 : (cy < 0 ? NP - W : 2 * NP); }
    static constexpr int off(int q) {
        return (3 * (LATTICE.cz[q] + 1) + (LATTICE.cx[q] + 1)) * B + cyoff(LATTICE.cy[q]);
    }
};

// Register-window march: LDS holds NBUF planes of the 27 populations pulled in one iteration;
// every position keeps in registers the c_z=+1 populations of its last two planes and the c_z=0
// populations of its last plane, and writes them into LDS in the iteration that pulls them.
// NBUF = 1 needs two barriers per plane and a quarter of the LDS of the ring march (more blocks
// per CU); NBUF = 2 alternates two LDS planes and needs one barrier per plane. The z loop is
// unrolled by two so the two c_z=+1 windows swap roles without register moves. Memory
// operations are ordered and made unconditional as in fused_march. The prefetch past the item's
// last halo plane z1 re-reads plane z1 (an L2 hit) instead of fetching a plane nobody uses.
// FACE (with SPECIAL false): a tile whose only boundary nodes form one domain face runs the plain
// march, and its face lanes replace the moments of their pulled populations with the face's
// closed form before the collision (the density and velocity of the node and, at the outlet, of its
// x-1 neighbour are kept from the decode of the plane, the x-1 values by lane shuffle).
template <class R, class S, int BX, int BY, int NBUF, bool SPECIAL, int FACE = FACE_NONE>
__device__ __forceinline__ void window_march(const FusedArgs<R, S> &a, R *lds, int tile, int z0, int z1, int rs) {
    static_assert(!SPECIAL || FACE == FACE_NONE, "a face march runs the plain path");
    using F = Frame<BX, BY>;
    constexpr int W = F::W;
    using L = WindowLds<BX, BY>;
    constexpr bool kReloadOwn = window_min_blocks<S, BX, BY, NBUF>() == 4;
    R *const wold = lds + NBUF * L::PLANE;

    const int x0 = (tile % a.tiles_x) * BX;
    const int y0 = (tile / a.tiles_x) * BY;
    const int tid = threadIdx.x;
    const int tx = tid % BX, ty = tid / BX;
    const uint32_t plane = uint32_t(a.nx) * uint32_t(a.ny);
    const int h = (ty + 1) * W + (tx + 1);

    // ring positions are held by threads rs .. rs + NR - 1
    const bool has1 = tid >= rs && tid < rs + F::NR;
    int hx1 = 0, hy1 = 0;
    if (has1) F::ring(tid - rs, hx1, hy1);
    const int p1 = hy1 * W + hx1;
    const uint32_t off0 = uint32_t(pmod(x0 + tx, a.nx)) + uint32_t(a.nx) * uint32_t(pmod(y0 + ty, a.ny));
    const uint32_t off1 = has1 ? uint32_t(pmod(x0 - 1 + hx1, a.nx)) + uint32_t(a.nx) * uint32_t(pmod(y0 - 1 + hy1, a.ny))
                               : off0;
    const bool west1 = has1 && hx1 == 0 && hy1 >= 1 && hy1 <= BY;
    // FP32 arithmetic: a ring position reconstructs only the nine populations that stream from it
    // into the tile -- c_y = +1 from the south row, -1 from the north row (corners included), c_x =
    // +1 / -1 from the west / east column -- and keeps windows of three per c_z group. roff[j]: its
    // LDS offset of the c_z = -1 population with c_b = j - 1 (+ 3 B per c_z step).
    constexpr bool RING9 = !kExactOrder<R>;
    constexpr int RW = RING9 ? 3 : 9;
    const bool ring_row = hy1 == 0 || hy1 == BY + 1;
    const int ring_si = ring_row ? (hy1 == 0 ? 1 : -1) : (hx1 == 0 ? 1 : -1);
    int roff[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        roff[j] = (ring_row ? j * L::B + L::cyoff(ring_si) : (ring_si + 1) * L::B + L::cyoff(j - 1)) + p1;

    // stored values of the plane in flight, per position (set B only holds plane z0 in the
    // prologue, which loads it together with plane z0-1)
    RawNode<S> ra0, ra1, rb0, rb1;
    // planes of whole chunks and a grid of whole tiles (every thread's node inside): loads and the
    // plain march's stores address a position by its fixed offsets from the plane's first element
    // (two-plane march only: in the one-plane march the offsets took 23 more registers and gained
    // nothing)
    const bool fast = NBUF == 2 && a.plane_aligned && a.nx % BX == 0 && a.ny % BY == 0;
    const size_t plane_elems = size_t(plane) * 10;
    const LaneOffsets lo0 = lane_offsets<S>(off0), lo1 = lane_offsets<S>(off1);
    auto load_plane_to = [&](int z, RawNode<S> &d0, RawNode<S> &d1) {
        const uint32_t zz = uint32_t(pmod(z, a.nz));
        NodePtrs q0, q1;
        if (fast) {
            const S *pz = a.src.base + size_t(zz) * plane_elems;
            q0 = node_ptrs_at(pz, lo0);
            q1 = node_ptrs_at(pz, lo1);
        } else {
            q0 = node_ptrs(a.src.base, zz * plane + off0);
            q1 = node_ptrs(a.src.base, zz * plane + off1);
        }
        load_raw_p(q0, d0);
        load_raw_p(q1, d1);
    };
    // windows, per position: c_z=+1 groups of two planes (A, B), c_z=0 group of one plane
    R pa0[9], pb0[9], q0[9], pa1[RW], pb1[RW], q1[RW];
    // face march: density and velocity of the own node of the plane being updated and of the next
    R fo_cur[4] = {R(0), R(0), R(0), R(0)}, fo_nx[4] = {R(0), R(0), R(0), R(0)};
    auto decode_pos = [&](const RawNode<S> &raw, R *m, bool west, int zp, R *keep) {
        decode_node<R, S>(raw, a.scale, a.rho0, m);
        if (FACE == FACE_WEST || FACE == FACE_EAST) {
            if (keep) {
#pragma unroll
                for (int k = 0; k < 4; ++k) keep[k] = m[k];
            }
        }
        if (SPECIAL && west) {
#pragma unroll
            for (int k = 0; k < 4; ++k) wold[(west_slot(zp) * 4 + k) * BY + (hy1 - 1)] = m[k];
        }
    };
    auto recon = [&](const RawNode<S> &raw, R *f, bool west, int zp, R *keep = nullptr) {
        R m[10];
        decode_pos(raw, m, west, zp, keep);
        reconstruct_lds(m, f);
    };
    // a ring position's populations of plane zp by c_z group (+1, 0, -1; RW values each; null: not
    // wanted)
    auto recon_ring = [&](const RawNode<S> &raw, bool west, int zp, R *gzp, R *gz0, R *gzm) {
        if constexpr (RING9) {
            R m[10], g[9];
            decode_pos(raw, m, west, zp, nullptr);
            reconstruct_ring(m, ring_row, R(ring_si), g);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                gzp[j] = g[6 + j];
                if (gz0) gz0[j] = g[3 + j];
                if (gzm) gzm[j] = g[j];
            }
        } else {
            R f[Q];
            recon(raw, f, west, zp);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                gzp[i] = f[GROUP_ZP[i]];
                if (gz0) gz0[i] = f[GROUP_Z0[i]];
                if (gzm) gzm[i] = f[GROUP_ZM[i]];
            }
        }
    };

    BoundaryPath<R, S, BX, BY> bp;
    // prologue: +1 group of plane z0-1 into A, +1 and 0 groups of plane z0 into B and q (the
    // two planes are loaded together, then plane z0+1 while plane z0 is reconstructed)
    {
        R f[Q];
        load_plane_to(z0 - 1, ra0, ra1);
        load_plane_to(z0, rb0, rb1);
        recon(ra0, f, false, z0 - 1);
#pragma unroll
        for (int i = 0; i < 9; ++i) pa0[i] = f[GROUP_ZP[i]];
        if (has1) recon_ring(ra1, false, z0 - 1, pa1, nullptr, nullptr);
        load_plane_to(z0 + 1, ra0, ra1);
        recon(rb0, f, false, z0, fo_cur);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            pb0[i] = f[GROUP_ZP[i]];
            q0[i] = f[GROUP_Z0[i]];
        }
        if (has1) recon_ring(rb1, west1, z0, pb1, q1, nullptr);
        if constexpr (SPECIAL) {
            if constexpr (!kReloadOwn) bp.copy_own(rb0, bp.own_cur);
            bp.load_info(a, off0 + uint32_t(z0) * plane, bp.info_cur);
        }
    }

    const int x = x0 + tx, y = y0 + ty;
    const bool inside = x < a.nx && y < a.ny;
    const bool east_tile = x0 + BX >= a.nx;
    uint32_t seed = 0;
    if constexpr (Conv<S>::scaled) {
        if (a.step_counter) seed = uint32_t(*a.step_counter) * 2u + 1u;
    }
    // the plain march stores a plane's results at the start of the next iteration (after the wait
    // for the prefetched plane, before the next prefetch)
    R Mp[10];
    uint32_t idxp = a.pad;
    int zprev = z0;  // plane of Mp
    // the plain march's store of the last results (none before the first plane)
    auto store_prev = [&](bool any) {
        if (fast && !any) return;
        NodePtrs q = fast ? node_ptrs_at(static_cast<const S *>(a.dst.base + size_t(zprev) * plane_elems), lo0)
                          : node_ptrs(static_cast<const S *>(a.dst.base), idxp);
        S v[10];
        encode_node<R, S, Conv<S>::scaled>(idxp, a.scale, a.rho0, Mp, seed, v);
        store_raw_p<S>(q, v);
    };

    // face lanes of a face march: fixed per thread (the outlet's x-1 neighbour is lane - 1, which
    // the host guarantees: the outlet column is not a tile's first column)
    const bool face_lane = FACE != FACE_NONE && inside &&
                           (FACE == FACE_WEST    ? x == 0
                            : FACE == FACE_EAST  ? x == a.nx - 1
                            : FACE == FACE_NORTH ? y == a.ny - 1
                                                 : y == 0);
    // one position: reconstruct plane z+1, write the pulled plane's groups, advance the windows
    // whether a ring position hands population q to the tile (the other populations of a ring
    // position are never pulled, and the LDS plane has no room for them: WindowLds)
    auto ring_takes = [&](int qd) { return ring_row ? LATTICE.cy[qd] == ring_si : LATTICE.cx[qd] == ring_si; };
    auto put = [&](const RawNode<S> &raw, int p, R *G, R *pold, R *q, bool west, int zp, R *keep = nullptr,
                   bool ring = false) {
        R f[Q];
        recon(raw, f, west, zp, keep);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            if (!ring || ring_takes(GROUP_ZP[i])) G[L::off(GROUP_ZP[i]) + p] = pold[i];          // plane z-1
            if (!ring || ring_takes(GROUP_Z0[i])) G[L::off(GROUP_Z0[i]) + p] = q[i];             // plane z
            if (!ring || ring_takes(GROUP_ZM[i])) G[L::off(GROUP_ZM[i]) + p] = f[GROUP_ZM[i]];   // plane z+1
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            pold[i] = f[GROUP_ZP[i]];  // becomes plane z+1's window
            q[i] = f[GROUP_Z0[i]];
        }
    };
    // the same for the thread's ring position
    auto put_ring = [&](const RawNode<S> &raw, R *G, R *pold, R *q, bool west, int zp) {
        if constexpr (RING9) {
            R gzp[3], gz0[3], gzm[3];
            recon_ring(raw, west, zp, gzp, gz0, gzm);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                G[roff[j] + 6 * L::B] = pold[j];
                G[roff[j] + 3 * L::B] = q[j];
                G[roff[j]] = gzm[j];
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                pold[j] = gzp[j];
                q[j] = gz0[j];
            }
        } else {
            put(raw, p1, G, pold, q, west, zp, nullptr, true);
        }
    };
    auto iteration = [&](int z, R *G, R *w0, R *w1, bool stored_before) {
        if (NBUF == 1 && z > z0) lds_barrier();  // previous plane's pulls are done
        put(ra0, h, G, w0, q0, false, z + 1, fo_nx);
        if (has1) put_ring(ra1, G, w1, q1, west1, z + 1);
        if constexpr (SPECIAL) {
            // the own node's stored values of plane z: kept from the prefetch of plane z (two
            // planes of them held in registers), or, in the register-capped march, loaded again
            // (an L2 hit: loaded two iterations ago) before the prefetch, so that the wait for
            // them does not cover it
            if constexpr (kReloadOwn)
                load_raw(a.src, off0 + uint32_t(z) * plane, bp.own_cur);
            else
                bp.copy_own(ra0, bp.own_nx);
            load_tabs(a, bp.info_cur, uint32_t(y) + uint32_t(a.ny) * uint32_t(z), bp.tab);
            bp.load_info(a, off0 + uint32_t(pmod(z + 1, a.nz)) * plane, bp.info_nx);
        } else {
            store_prev(stored_before);
        }
        // the inlet velocity of the plane (every thread loads an in-range entry)
        R u_in = R(0);
        if constexpr (FACE == FACE_WEST)
            u_in = a.inlet[min(uint32_t(y) + uint32_t(a.ny) * uint32_t(z), uint32_t(a.inlet_count - 1))];
        load_plane_to(min(z + 2, z1), ra0, ra1);  // consumed in the next iteration
        lds_barrier();
        R f[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = G[L::off(q) + h - LATTICE.cx[q] - LATTICE.cy[q] * W];
        const uint32_t node = off0 + uint32_t(z) * plane;
        if constexpr (SPECIAL) {
            bp.update(a, f, tx, ty, inside, east_tile, node, wold + west_slot(z) * 4 * BY, seed);
            if constexpr (!kReloadOwn) bp.copy_own(bp.own_nx, bp.own_cur);
            bp.info_cur = bp.info_nx;
        } else {
            moments_lds(f, Mp);
            if constexpr (FACE != FACE_NONE) {
                if (__any(face_lane)) {  // the face's closed form (cases/cylinder/boundaries.cuh:201-315)
                    R g[Q], Mf[10], old[10], oxm[4];
#pragma unroll
                    for (int q = 0; q < Q; ++q) g[q] = f[q];
                    lds_to_pop(g);
#pragma unroll
                    for (int k = 0; k < 10; ++k) old[k] = R(0);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        old[k] = fo_cur[k];
                        oxm[k] = __shfl(fo_cur[k], int(threadIdx.x) - 1, 64);
                    }
                    if constexpr (FACE == FACE_WEST) bc_west(g, old, u_in, Mf);
                    else if constexpr (FACE == FACE_EAST) bc_east(g, old, oxm, Mf);
                    else if constexpr (FACE == FACE_NORTH) bc_ywall<face::K_N>(g, Mf);
                    else bc_ywall<face::K_S>(g, Mf);
                    if (face_lane) {
#pragma unroll
                        for (int k = 0; k < 10; ++k) Mp[k] = Mf[k];
                    }
                }
            }
            collide_for_store<R, S>(a.omega, true, Mp);
            watch_finite(a, Mp[0]);
            idxp = inside ? node : a.pad;
            zprev = z;
        }
        if constexpr (FACE == FACE_WEST || FACE == FACE_EAST) {
#pragma unroll
            for (int k = 0; k < 4; ++k) fo_cur[k] = fo_nx[k];
        }
    };
    R *const G0 = lds;
    R *const G1 = NBUF == 2 ? lds + L::PLANE : lds;
    int z = z0;
    iteration(z, G0, pa0, pa1, false);
    ++z;
    while (z < z1) {
        iteration(z, G1, pb0, pb1, true);
        if (++z >= z1) break;
        iteration(z, G0, pa0, pa1, true);
        ++z;
    }
    if constexpr (!SPECIAL) store_prev(true);
    if (NBUF == 1) lds_barrier();  // the block's LDS is free again for the next item
}

// One work item of the register-window march from its tile field: the boundary path for special
// tiles (when BOUNDARY), a face march for face tiles, the plain march otherwise.
// The ring positions (more work per thread) go to the first waves of the block, or, for odd tiles,
// to its last waves: blocks resident on one CU then spread that work over the SIMDs instead of
// loading the SIMDs of the first waves with all of it.
template <int BX, int BY>
constexpr int ring_alt_start() {
    constexpr int nt = BX * BY, nr = Frame<BX, BY>::NR, nrw = (nr + 63) / 64 * 64;
    return nt - nrw >= 64 ? nt - nrw : 0;
}

template <class R, class S, int BX, int BY, int NBUF, bool BOUNDARY>
__device__ __forceinline__ void window_item(const FusedArgs<R, S> &a, R *lds, uint32_t field, int z0, int z1) {
    const int tile = int(field & kItemTileMask);
    const int face = int((field >> kItemFaceShift) & 7u);
    const int rs = (tile & 1) ? ring_alt_start<BX, BY>() : 0;
    if (!BOUNDARY || !(field & kItemSpecial))
        window_march<R, S, BX, BY, NBUF, false>(a, lds, tile, z0, z1, rs);
    else if (face == FACE_WEST)
        window_march<R, S, BX, BY, NBUF, false, FACE_WEST>(a, lds, tile, z0, z1, rs);
    else if (face == FACE_EAST)
        window_march<R, S, BX, BY, NBUF, false, FACE_EAST>(a, lds, tile, z0, z1, rs);
    else if (face == FACE_NORTH)
        window_march<R, S, BX, BY, NBUF, false, FACE_NORTH>(a, lds, tile, z0, z1, rs);
    else if (face == FACE_SOUTH)
        window_march<R, S, BX, BY, NBUF, false, FACE_SOUTH>(a, lds, tile, z0, z1, rs);
    else
        window_march<R, S, BX, BY, NBUF, true>(a, lds, tile, z0, z1, rs);
}

// PART: which launch of the fused pass (PART_ALL, PART_INTERIOR, PART_WALL_SIDE). The interior
// launch has no boundary tile, so the boundary path is not compiled into it; the two parts are
// separate instantiations so that profiles tell the launches apart.
template <class R, class S, int BX, int BY, int NBUF, int PART>
__global__ __launch_bounds__(BX *BY, (window_min_blocks<S, BX, BY, NBUF>())) void fused_window_kernel(FusedArgs<R, S> a) {
    extern __shared__ unsigned char lds_raw[];
    R *lds = reinterpret_cast<R *>(lds_raw);
    int b = blockIdx.x;
    if (b >= a.nblocks) {  // the block after the work items files the previous step's record
        if (threadIdx.x < 64) file_record(a.rec);
        return;
    }
    if (a.xcd_groups > 1) b = (b % a.xcd_groups) * (a.nblocks / a.xcd_groups) + b / a.xcd_groups;
    const uint2 item = a.items[b];
    const int z0 = int(item.y & 0xFFFFu);
    const int z1 = int(item.y >> 16);
    if (z0 >= z1) return;  // padding item (uniform per block)
    window_item<R, S, BX, BY, NBUF, PART != PART_INTERIOR>(a, lds, item.x, z0, z1);
}

template <class R, class S, int BX, int BY, int PART>
__global__ __launch_bounds__(BX *BY) void fused_step_ring6_kernel(FusedArgs<R, S> a) {
    extern __shared__ unsigned char lds_raw[];
    R *lds = reinterpret_cast<R *>(lds_raw);
    // block -> work item; consecutive items share an XCD so that neighbouring tiles reuse each
    // other's ring loads from the same L2
    int b = blockIdx.x;
    if (b >= a.nblocks) {  // the block after the work items files the previous step's record
        if (threadIdx.x < 64) file_record(a.rec);
        return;
    }
    if (a.xcd_groups > 1) b = (b % a.xcd_groups) * (a.nblocks / a.xcd_groups) + b / a.xcd_groups;
    const uint2 item = a.items[b];
    const int tile = int(item.x & kItemTileMask);
    const bool special = (item.x & kItemSpecial) != 0;
    const int z0 = int(item.y & 0xFFFFu);
    const int z1 = int(item.y >> 16);
    if (z0 >= z1) return;  // padding item (uniform per block)
    if (PART != PART_INTERIOR && special)
        fused_march<R, S, BX, BY, true>(a, lds, tile, z0, z1);
    else
        fused_march<R, S, BX, BY, false>(a, lds, tile, z0, z1);
}

// The stored words of a node folded into one word (keeps every loaded word alive).
template <class T>
__device__ __forceinline__ uint32_t fold_word(const T &v) {
    struct Words {
        uint32_t w[sizeof(T) / 4];
    };
    const Words x = __builtin_bit_cast(Words, v);
    uint32_t h = 0;
#pragma unroll
    for (size_t k = 0; k < sizeof(T) / 4; ++k) h ^= x.w[k];
    return h;
}

template <class S>
__device__ __forceinline__ uint32_t fold_bits(const RawNode<S> &r) {
    if constexpr (kQuads<S>) {
        return fold_word(r.q0) ^ fold_word(r.q1) ^ fold_word(r.p);
    } else {
        uint32_t h = 0;
#pragma unroll
        for (int j = 0; j < 5; ++j) h ^= fold_word(r.p[j]);
        return h;
    }
}

// Memory-only twin of the fused pass: the same work items and XCD remap, and per plane the same
// loads of the tile and its one-node ring (planes z0 - 1 .. z1, the next plane in flight while a
// plane is stored) and the same store of every tile node, without arithmetic, LDS or barriers.
// Its duration is the floor that this access pattern sets on this device, the ceiling the fused
// pass is measured against beside HBM peak. It copies the state into the other buffer (which the
// next step overwrites; frozen nodes hold the same value in both), so it can run on a live
// simulation. The ring's values are folded into the stored bits under a mask that is zero at run
// time (a.nx > 0), so no load is dropped.
template <class R, class S, int BX, int BY>
__global__ __launch_bounds__(BX *BY) void memory_twin_kernel(FusedArgs<R, S> a) {
    int b = blockIdx.x;
    if (b >= a.nblocks) return;
    if (a.xcd_groups > 1) b = (b % a.xcd_groups) * (a.nblocks / a.xcd_groups) + b / a.xcd_groups;
    const uint2 item = a.items[b];
    const int tile = int(item.x & kItemTileMask);
    const int z0 = int(item.y & 0xFFFFu);
    const int z1 = int(item.y >> 16);
    if (z0 >= z1) return;
    using F = Frame<BX, BY>;
    const int x0 = (tile % a.tiles_x) * BX;
    const int y0 = (tile / a.tiles_x) * BY;
    const int tid = threadIdx.x;
    const int tx = tid % BX, ty = tid / BX;
    const uint32_t plane = uint32_t(a.nx) * uint32_t(a.ny);
    const bool has1 = tid < F::NR;
    int hx1 = 0, hy1 = 0;
    if (has1) F::ring(tid, hx1, hy1);
    const uint32_t off0 = uint32_t(pmod(x0 + tx, a.nx)) + uint32_t(a.nx) * uint32_t(pmod(y0 + ty, a.ny));
    const uint32_t off1 = has1 ? uint32_t(pmod(x0 - 1 + hx1, a.nx)) + uint32_t(a.nx) * uint32_t(pmod(y0 - 1 + hy1, a.ny))
                               : off0;
    const bool inside = x0 + tx < a.nx && y0 + ty < a.ny;
    const uint32_t keep = a.nx > 0 ? 0u : ~0u;
    // the fused march's addressing of whole-chunk planes of a grid of whole tiles
    const bool fast = a.plane_aligned && a.nx % BX == 0 && a.ny % BY == 0;
    const size_t plane_elems = size_t(plane) * 10;
    const LaneOffsets lo0 = lane_offsets<S>(off0), lo1 = lane_offsets<S>(off1);
    RawNode<S> c0, c1, n0, n1;
    auto load_plane_to = [&](int z, RawNode<S> &d0, RawNode<S> &d1) {
        const uint32_t zz = uint32_t(pmod(z, a.nz));
        NodePtrs q0, q1;
        if (fast) {
            const S *pz = a.src.base + size_t(zz) * plane_elems;
            q0 = node_ptrs_at(pz, lo0);
            q1 = node_ptrs_at(pz, lo1);
        } else {
            q0 = node_ptrs(a.src.base, zz * plane + off0);
            q1 = node_ptrs(a.src.base, zz * plane + off1);
        }
        load_raw_p(q0, d0);
        load_raw_p(q1, d1);
    };
    load_plane_to(z0 - 1, c0, c1);
    uint32_t fold = fold_bits(c0) ^ fold_bits(c1);
    load_plane_to(z0, c0, c1);
    for (int z = z0; z < z1; ++z) {
        load_plane_to(z + 1, n0, n1);  // in flight while plane z is stored
        fold ^= fold_bits(c1);
        RawNode<S> v = c0;
        if (fold & keep) v = c1;  // never at run time
        store_raw_p<S>(fast ? node_ptrs_at(static_cast<const S *>(a.dst.base + size_t(z) * plane_elems), lo0)
                            : node_ptrs(static_cast<const S *>(a.dst.base), inside ? off0 + uint32_t(z) * plane : a.pad),
                       v);
        c0 = n0;
        c1 = n1;
    }
    if ((fold ^ fold_bits(c0) ^ fold_bits(c1)) & keep) store_raw<S>(a.dst, a.pad, c0);
}

// One work item of a batch {case, tile | special << 31, z0 | z1 << 16, 0}; the case's arguments
// come from a device array. VARIANT as in launch_fused.
template <class R, class S, int BX, int BY, int VARIANT>
__device__ __forceinline__ void batch_item(const FusedArgs<R, S> *__restrict__ cases, const uint4 &item, R *lds) {
    const int z0 = int(item.z & 0xFFFFu);
    const int z1 = int(item.z >> 16);
    if (z0 >= z1) return;  // padding item (uniform per block)
    const FusedArgs<R, S> a = cases[item.x];
    if constexpr (VARIANT == 1) {
        const int tile = int(item.y & kItemTileMask);
        if (item.y & kItemSpecial)
            fused_march<R, S, BX, BY, true>(a, lds, tile, z0, z1);
        else
            fused_march<R, S, BX, BY, false>(a, lds, tile, z0, z1);
    } else {
        window_item<R, S, BX, BY, VARIANT == 4 ? 2 : 1, true>(a, lds, item.y, z0, z1);
    }
}

// Fused pass of a batch: one block per work item, or q.pblocks persistent blocks that pull the
// items of their XCD group through the group's counter until the group's list is exhausted (every
// block stops at the end of the list; no block waits for another). Then one block per case that
// files the case's previous record.
template <class R, class S, int BX, int BY, int VARIANT>
__global__ __launch_bounds__(BX *BY, (VARIANT == 3 ? window_min_blocks<S, BX, BY, 1>() : 2)) void fused_batch_kernel(
    const FusedArgs<R, S> *__restrict__ cases, const uint4 *__restrict__ items, int nitems, int groups, BatchQueue q) {
    extern __shared__ unsigned char lds_raw[];
    R *lds = reinterpret_cast<R *>(lds_raw);
    int b = blockIdx.x;
    const int nwork = q.pblocks > 0 ? q.pblocks : nitems;
    if (b >= nwork) {  // blocks after the work blocks file the cases' previous records
        const RecordArgs r = cases[b - nwork].rec;
        if (r.pending && threadIdx.x < 64) file_record(r);
        return;
    }
    if (q.pblocks == 0) {
        if (groups > 1) b = (b % groups) * (nitems / groups) + b / groups;
        batch_item<R, S, BX, BY, VARIANT>(cases, items[b], lds);
        return;
    }
    // the next step (the other buffer parity) starts from zeroed counters
    if (b == 0 && int(threadIdx.x) < groups) q.counters[(q.parity ^ 1) * groups + int(threadIdx.x)] = 0u;
    const int g = b % groups;
    const int len = nitems / groups;
    unsigned *const counter = q.counters + q.parity * groups + g;
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = int(atomicAdd(counter, 1u));
        __syncthreads();
        const int k = next;
        __syncthreads();  // every thread has read the index before thread 0 takes the next one
        if (k >= len) break;
        batch_item<R, S, BX, BY, VARIANT>(cases, items[g * len + k], lds);
        __syncthreads();  // the item's LDS reads are done before the next item's writes
    }
}

// Bilinear blend of four corner values (bilinear_interpolation, extrapolation_utils.cuh:5-47).
template <class R>
__device__ __forceinline__ R bilinear(R q00, R q10, R q01, R q11, R xd, R yd) {
    const R q0 = (R(1) - xd) * q00 + xd * q10;
    const R q1 = (R(1) - xd) * q01 + xd * q11;
    return (R(1) - yd) * q0 + yd * q1;
}

// Interpolate u and m at one stencil point and rotate into the wall frame
// (initerpolate_and_rotate + rotate_velocity_moments, extrapolation_utils.cuh:63-107). The
// donor snapshot is node-major: each corner's nine values come from one cache line or two.
template <class R, class S>
__device__ __forceinline__ void interp_rotate(const WallArgs<R, S> &a, uint32_t base, R xd, R yd, R c, R s, R *vm) {
    const uint32_t corner[4] = {base, base + 1, base + uint32_t(a.nx), base + uint32_t(a.nx) + 1};
    R q[4][9];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // the corner's ten stored values as five pair words (a node starts on a pair boundary)
        const pair_t<S> *pw = reinterpret_cast<const pair_t<S> *>(a.donor + size_t(corner[j]) * 10);
        S raw[10];
#pragma unroll
        for (int w = 0; w < 5; ++w) split_pair<S>(pw[w], raw[2 * w], raw[2 * w + 1]);
#pragma unroll
        for (int k = 0; k < 9; ++k) q[j][k] = stored_value<R, S>(raw[1 + k], k + 1, a.scale);
        if constexpr (Conv<S>::scaled) {
            q[j][0] += a.scale.uref;
            // stored second moments are non-equilibrium parts: m = (m - u u) + u u is restored at
            // each corner before the corners are blended
            q[j][3] += q[j][0] * q[j][0];
            q[j][4] += q[j][1] * q[j][1];
            q[j][5] += q[j][2] * q[j][2];
            q[j][6] += q[j][0] * q[j][1];
            q[j][7] += q[j][0] * q[j][2];
            q[j][8] += q[j][1] * q[j][2];
        }
    }
    R v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = bilinear(q[0][k], q[1][k], q[2][k], q[3][k], xd, yd);
    const R ux = v[0], uy = v[1], uz = v[2], mxx = v[3], myy = v[4], mzz = v[5], mxy = v[6], mxz = v[7], myz = v[8];
    const R s2 = R(2) * s * c, c2 = c * c - s * s;
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


// Sum of v over the block, the same fixed tree on every call (deterministic).
__device__ __forceinline__ double wall_block_sum(double v) {
    constexpr int NW = WALL_THREADS / 64;
    __shared__ double red[NW];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if constexpr (NW == 1) {
        return __shfl(v, 0, 64);
    } else {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        __syncthreads();
        if (lane == 0) red[wave] = v;
        __syncthreads();
        double t = 0.0;
        for (int w = 0; w < NW; ++w) t += red[w];
        return t;
    }
}

// One block of the wall pass of one case: block `lb` of the case's a.nblocks blocks.
template <class R, class S>
__device__ __forceinline__ void wall_body(const WallArgs<R, S> &a, int lb) {
    const int i = lb * WALL_THREADS + threadIdx.x;
    double F[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < a.nwall) {
        const uint32_t idx = a.node[i];
        R M[10];
        load_node_nm<R, S>(a.donor, idx, a.scale, a.rho0, M);  // rho_I, u_old, m'_I (the fused pass's snapshot)
        const R c = a.wnx[i], s = a.wny[i];
        // record r of the node (component-major, nw records) and the plane offset of its stencils
        int r = i, nw = a.nwall;
        uint32_t zoff = 0;
        if (a.col0 > 0) {
            const bool outer = i >= a.nb_inner;
            const int local = outer ? i - a.nb_inner : i;
            const int nwp = outer ? a.col1 : a.col0;
            const int z = local / nwp;
            r = (outer ? a.col0 : 0) + (local - z * nwp);
            nw = a.col0 + a.col1;
            zoff = uint32_t(z) * a.plane;
        }
        R vm1[9], vm2[9];
        interp_rotate(a, a.st_base[r] + zoff, a.st_frac[r], a.st_frac[nw + r], c, s, vm1);
        interp_rotate(a, a.st_base[nw + r] + zoff, a.st_frac[2 * nw + r], a.st_frac[3 * nw + r], c, s, vm2);
        const int g = i < a.nb_inner ? 0 : 1;
        const R u0 = a.uw[g][0], u1 = a.uw[g][1], u2 = a.uw[g][2];
        // wall values u'_w and m'_w = u'_w u'_w (curvedLBM.cu:93-101)
        const R vw[7] = {u0, u1, u2, u0 * u0, u1 * u1, u2 * u2, u1 * u2};
        constexpr int sel[7] = {0, 1, 2, 3, 4, 5, 8};
        const R c0 = a.coef[r], c1 = a.coef[nw + r], c2 = a.coef[2 * nw + r];
        R p[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) p[j] = c0 * vw[j] + c1 * vm1[sel[j]] + c2 * vm2[sel[j]];
        R out[10];
        if (a.strong) {
            wall_closure_strong(c, s, a.in_mask[i], a.out_mask[r], a.omega, p, M[0], M[7], M[8], out);
        } else if constexpr (kExactOrder<R>) {
            wall_closure_loop(c, s, a.in_mask[i], p, M[0], M[7], M[8], out);
        } else {
            R k[WALL_CONSTS];
#pragma unroll
            for (int j = 0; j < WALL_CONSTS; ++j) k[j] = a.wconst[size_t(j) * nw + r];
            wall_closure(c, s, k, p, M[0], M[7], M[8], out);
        }
        if (a.forces) masked_force(out, a.in_mask[i], 1.0, F);
        collide(a.omega, out);
        if (a.forces) masked_force(out, a.out_mask[r], -1.0, F);
        uint32_t seed = 0;
        if constexpr (Conv<S>::scaled) seed = uint32_t(*a.step_counter) * 2u + 2u;
        store_node<R, S>(a.mom, idx, a.scale, a.rho0, out, seed);
    } else if (a.forces && i < a.nwall + a.nbcf) {
        const int k = i - a.nwall;
        const uint32_t idx = a.bcf_node[k];
        const uint32_t t = a.bcf_tag[k];
        R M[10];
        load_node_nm<R, S>(a.donor, idx, a.scale, a.rho0, M);
        masked_force(M, a.bcf_in[t], 1.0, F);
        collide(a.omega, M);
        masked_force(M, a.bcf_out[t], -1.0, F);
    }
    // block partial of the force / mass records in FP64 (wave shuffles, then across the waves)
    if (a.forces) {
        const double p0 = wall_block_sum(F[0]), p1 = wall_block_sum(F[1]), p2 = wall_block_sum(F[2]),
                     p3 = wall_block_sum(F[3]);
        if (threadIdx.x == 0) {
            a.partials[4 * lb + 0] = p0;
            a.partials[4 * lb + 1] = p1;
            a.partials[4 * lb + 2] = p2;
            a.partials[4 * lb + 3] = p3;
        }
    }
    // the next step's index, and this step's record marked pending
    if (lb == 0 && threadIdx.x == 0) {
        const unsigned long long step = *a.step_counter;
        *a.next_counter = step + 1;
        if (a.forces) *a.pending = step + 1;
    }
}

__global__ __launch_bounds__(64) void record_kernel(RecordArgs r) { file_record(r); }

template <class R, class S>
__global__ __launch_bounds__(64) void record_batch_kernel(const FusedArgs<R, S> *__restrict__ cases) {
    const RecordArgs r = cases[blockIdx.x].rec;
    if (r.pending) file_record(r);
}

template <class R, class S>
__global__ __launch_bounds__(WALL_THREADS) void wall_step_kernel(WallArgs<R, S> a) {
    wall_body(a, xcd_contiguous(int(blockIdx.x), int(gridDim.x)));
}

// Wall pass of a batch: block b runs block blocks[b].y of case blocks[b].x.
template <class R, class S>
__global__ __launch_bounds__(WALL_THREADS) void wall_batch_kernel(const WallArgs<R, S> *__restrict__ cases,
                                                                  const int2 *__restrict__ blocks) {
    const int2 bc = blocks[blockIdx.x];
    const WallArgs<R, S> a = cases[bc.x];
    wall_body(a, bc.y);
}

// Wall pressure of every inner-wall node (compute_surface_pressure,
// cases/cylinder/cylinder_kernels.cu:238-283): density bilinearly interpolated at fluid points
// 1, 2 and 3 spacings out along the normal, p = rho cs2, and the quadratic in the distance from
// the centre through them evaluated at the wall radius (surface_pressure_extrapolation,
// boundary/extrapolation_utils.cuh:109-154). Evaluated in FP64 and added to a per-node FP64 sum
// (the reference keeps a running mean in the working precision).
template <class R, class S>
__global__ __launch_bounds__(256) void pressure_kernel(PressureArgs<R, S> a) {
    const int i = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (i >= a.n) return;
    const uint32_t idx = a.node[i];
    const uint32_t plane = uint32_t(a.nx) * uint32_t(a.ny);
    const uint32_t zoff = (idx / plane) * plane;
    const double nx = double(a.wnx[i]), ny = double(a.wny[i]);
    const double xw = a.xc + a.radius * nx, yw = a.yc + a.radius * ny;
    double xs[3], ys[3], rho[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        xs[p] = xw + double(p + 1) * 1.0 * nx;
        ys[p] = yw + double(p + 1) * 1.0 * ny;
        const int x0 = int(floor(xs[p])), y0 = int(floor(ys[p]));
        const double xd = xs[p] - double(x0), yd = ys[p] - double(y0);
        const uint32_t b = zoff + uint32_t(x0) + uint32_t(a.nx) * uint32_t(y0);
        double q[4];
        const uint32_t c[4] = {b, b + 1, b + uint32_t(a.nx), b + uint32_t(a.nx) + 1};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            q[j] = double(Conv<S>::to(ld_field(a.mom, 0, c[j])));
            if constexpr (!Conv<S>::unit_scale) q[j] *= double(a.scale.dec[0]);
        }
        const double q0 = (1.0 - xd) * q[0] + xd * q[1];
        const double q1 = (1.0 - xd) * q[2] + xd * q[3];
        rho[p] = a.rho0 + ((1.0 - yd) * q0 + yd * q1);
    }
    const double cs2 = 1.0 / 3.0;
    const double xw_diff = xw - a.xc, yw_diff = yw - a.yc;
    const double x1_diff = xs[0] - a.xc, y1_diff = ys[0] - a.yc;
    const double x2_diff = xs[1] - a.xc, y2_diff = ys[1] - a.yc;
    const double x3_diff = xs[2] - a.xc, y3_diff = ys[2] - a.yc;
    const double rw = sqrt(xw_diff * xw_diff + yw_diff * yw_diff);
    const double r1 = sqrt(x1_diff * x1_diff + y1_diff * y1_diff);
    const double r2 = sqrt(x2_diff * x2_diff + y2_diff * y2_diff);
    const double r3 = sqrt(x3_diff * x3_diff + y3_diff * y3_diff);
    const double denom = (r1 - r2) * (r1 - r3) * (r2 - r3);
    const double p1 = rho[0] * cs2, p2 = rho[1] * cs2, p3 = rho[2] * cs2;
    const double a0 =
        (r1 * r3 * p2 * (r3 - r1) + (r2 * r2) * (r3 * p1 - r1 * p3) + r2 * ((r1 * r1) * p3 - (r3 * r3) * p1)) / denom;
    const double a1 = ((r3 * r3) * (p1 - p2) + (r1 * r1) * (p2 - p3) + (r2 * r2) * (p3 - p1)) / denom;
    const double a2 = (r3 * (p2 - p1) + r2 * (p1 - p3) + r1 * (p3 - p2)) / denom;
    a.sum[i] += a0 + a1 * rw + a2 * (rw * rw);
}

// Field-major double arrays (field 0 = rho - rho0, the stored deviation) <-> storage.
template <class R, class S>
__global__ void encode_kernel(const double *__restrict__ in, FieldSet<S> out, uint32_t n, StoreScale<R> sc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) v[f] = in[size_t(f) * n + i];
    if constexpr (!Conv<S>::scaled) {
#pragma unroll
        for (int f = 0; f < 10; ++f) st_field(out, f, i, Conv<S>::from(v[f]));
    } else {
        R M[10];
        M[0] = R(1) + R(v[0]);
#pragma unroll
        for (int f = 1; f < 10; ++f) M[f] = R(v[f]);
        store_node<R, S>(out, i, sc, R(1), M);
        st_field(out, 0, i, Conv<S>::from(R(v[0]) * sc.enc[0]));
    }
}

template <class R, class S>
__global__ void decode_kernel(FieldSet<const S> in, double *__restrict__ out, uint32_t n, StoreScale<R> sc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (!Conv<S>::scaled) {
#pragma unroll
        for (int f = 0; f < 10; ++f) out[size_t(f) * n + i] = double(Conv<S>::to(ld_field(in, f, i)));
    } else {
        R M[10];
        load_node<R, S>(in, i, sc, R(0), M);
        out[i] = double(R(Conv<S>::to(ld_field(in, 0, i))) * sc.dec[0]);
#pragma unroll
        for (int f = 1; f < 10; ++f) out[size_t(f) * n + i] = double(M[f]);
    }
}

}  // namespace

template <class R, class S>
bool fused_part_supported(int bx, int by, int variant) {
    return ((bx == 32 && by == 8) || (bx == 16 && by == 8)) && (variant == 1 || variant == 4);
}

// Kernel of the fused pass for a tile, march variant and launch part (null if not instantiated).
template <class R, class S>
auto fused_kernel_for(int bx, int by, int variant, int part) -> void (*)(FusedArgs<R, S>) {
#define MRLBM_FUSED_TILE(TX, TY, P)                                                  \
    if (bx == TX && by == TY && part == P) {                                         \
        if (variant == 3) return fused_window_kernel<R, S, TX, TY, 1, P>;            \
        if (variant == 4) return fused_window_kernel<R, S, TX, TY, 2, P>;            \
        return fused_step_ring6_kernel<R, S, TX, TY, P>;                             \
    }
    MRLBM_FUSED_TILE(32, 8, PART_ALL)
    MRLBM_FUSED_TILE(16, 8, PART_ALL)
    MRLBM_FUSED_TILE(32, 4, PART_ALL)
    MRLBM_FUSED_TILE(16, 16, PART_ALL)
    MRLBM_FUSED_TILE(64, 4, PART_ALL)
    MRLBM_FUSED_TILE(16, 4, PART_ALL)
    MRLBM_FUSED_TILE(32, 16, PART_ALL)
#undef MRLBM_FUSED_TILE
    // the two-stream schedule's launches: tiles 32x8 and 16x8, ring or two-plane window march
#define MRLBM_FUSED_PART(TX, TY, P)                                                  \
    if (bx == TX && by == TY && part == P) {                                         \
        if (variant == 4) return fused_window_kernel<R, S, TX, TY, 2, P>;            \
        if (variant == 1) return fused_step_ring6_kernel<R, S, TX, TY, P>;           \
        return nullptr;                                                              \
    }
    MRLBM_FUSED_PART(32, 8, PART_INTERIOR)
    MRLBM_FUSED_PART(16, 8, PART_INTERIOR)
    MRLBM_FUSED_PART(32, 8, PART_WALL_SIDE)
    MRLBM_FUSED_PART(16, 8, PART_WALL_SIDE)
#undef MRLBM_FUSED_PART
    return nullptr;
}

template <class R, class S>
int prepare_fused(int bx, int by, int variant, int part) {
    auto kernel = fused_kernel_for<R, S>(bx, by, variant, part);
    if (!kernel) return -1;
    const size_t lds = fused_lds_bytes(bx, by, sizeof(R), variant);
    if (lds <= 65536) return 0;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(lds)) == hipSuccess
               ? 0
               : -1;
}

template <class R, class S>
void launch_memory_twin(const FusedArgs<R, S> &a, int bx, int by, hipStream_t stream) {
#define MRLBM_TWIN_TILE(TX, TY)                                                                        \
    if (bx == TX && by == TY) {                                                                        \
        hipLaunchKernelGGL((memory_twin_kernel<R, S, TX, TY>), dim3(a.nblocks), dim3(TX * TY), 0, stream, a); \
        return;                                                                                        \
    }
    MRLBM_TWIN_TILE(32, 8)
    MRLBM_TWIN_TILE(16, 8)
    MRLBM_TWIN_TILE(32, 4)
    MRLBM_TWIN_TILE(16, 16)
    MRLBM_TWIN_TILE(64, 4)
    MRLBM_TWIN_TILE(16, 4)
    MRLBM_TWIN_TILE(32, 16)
#undef MRLBM_TWIN_TILE
}

// The kernel's dynamic LDS limit is set by prepare_fused when the configuration is chosen.
template <class R, class S>
void launch_fused(const FusedArgs<R, S> &a, int bx, int by, int variant, int part, hipStream_t stream) {
    const size_t lds = fused_lds_bytes(bx, by, sizeof(R), variant);
    auto kernel = fused_kernel_for<R, S>(bx, by, variant, part);
    if (kernel)
        hipLaunchKernelGGL(kernel, dim3(a.nblocks + (a.rec.pending ? 1 : 0)), dim3(bx * by), lds, stream, a);
}

// Kernel of the batched fused pass for a tile and variant (null if not instantiated).
template <class R, class S>
auto fused_batch_kernel_for(int bx, int by, int variant)
    -> void (*)(const FusedArgs<R, S> *, const uint4 *, int, int, BatchQueue) {
#define MRLBM_BATCH_TILE(TX, TY)                                                     \
    if (bx == TX && by == TY) {                                                      \
        if (variant == 3) return fused_batch_kernel<R, S, TX, TY, 3>;                \
        if (variant == 4) return fused_batch_kernel<R, S, TX, TY, 4>;                \
        return fused_batch_kernel<R, S, TX, TY, 1>;                                  \
    }
    MRLBM_BATCH_TILE(32, 8)
    MRLBM_BATCH_TILE(16, 8)
    MRLBM_BATCH_TILE(32, 4)
#undef MRLBM_BATCH_TILE
    return nullptr;
}

template <class R, class S>
int prepare_fused_batch(int bx, int by, int variant) {
    auto kernel = fused_batch_kernel_for<R, S>(bx, by, variant);
    if (!kernel) return -1;
    const size_t lds = fused_lds_bytes(bx, by, sizeof(R), variant);
    if (lds <= 65536) return 0;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(lds)) == hipSuccess
               ? 0
               : -1;
}

template <class R, class S>
void launch_fused_batch(const FusedArgs<R, S> *cases, int ncases, const uint4 *items, int nitems, int groups, int bx,
                        int by, int variant, const BatchQueue &queue, hipStream_t stream) {
    const size_t lds = fused_lds_bytes(bx, by, sizeof(R), variant);
    auto kernel = fused_batch_kernel_for<R, S>(bx, by, variant);
    const int nwork = queue.pblocks > 0 ? queue.pblocks : nitems;
    if (kernel)
        hipLaunchKernelGGL(kernel, dim3(nwork + ncases), dim3(bx * by), lds, stream, cases, items, nitems, groups,
                           queue);
}

template <class R, class S>
void launch_record_batch(const FusedArgs<R, S> *cases, int ncases, hipStream_t stream) {
    if (ncases > 0) hipLaunchKernelGGL((record_batch_kernel<R, S>), dim3(ncases), dim3(64), 0, stream, cases);
}

template <class R, class S>
void launch_wall_batch(const WallArgs<R, S> *cases, const int2 *blocks, int nblocks, hipStream_t stream) {
    hipLaunchKernelGGL((wall_batch_kernel<R, S>), dim3(nblocks), dim3(WALL_THREADS), 0, stream, cases, blocks);
}

template <class R, class S>
void launch_wall(const WallArgs<R, S> &a, hipStream_t stream) {
    const int blocks = wall_blocks(a.nwall, a.forces ? a.nbcf : 0);
    hipLaunchKernelGGL((wall_step_kernel<R, S>), dim3(blocks), dim3(WALL_THREADS), 0, stream, a);
}

template <class R, class S>
void launch_pressure(const PressureArgs<R, S> &a, hipStream_t stream) {
    if (a.n > 0) hipLaunchKernelGGL((pressure_kernel<R, S>), dim3((a.n + 255) / 256), dim3(256), 0, stream, a);
}

// Plane means of one step (inlet_avg_density, src/cases/cylinder/cylinder_kernels.cu:285-338, and
// outlet_avg_ux, src/solver/mlbm.cu:18-81): FP64 sums in a fixed order (strided per thread, then a
// tree), so the records are deterministic.
template <class R, class S>
__global__ __launch_bounds__(256) void plane_kernel(PlaneArgs<R, S> a) {
    __shared__ double red[2][256];
    const int t = int(threadIdx.x);
    const int n = a.ny * a.nz;
    double sr = 0.0, su = 0.0;
    for (int i = t; i < n; i += 256) {
        const int y = i % a.ny, z = i / a.ny;
        const uint32_t row = uint32_t(a.nx) * (uint32_t(y) + uint32_t(a.ny) * uint32_t(z));
        double r = double(Conv<S>::to(ld_field(a.mom, 0, row)));
        double u = double(Conv<S>::to(ld_field(a.mom, 1, row + uint32_t(a.nx - 2))));
        if constexpr (Conv<S>::scaled) {
            r *= double(a.scale.dec[0]);
            u = double(stored_value<R, S>(ld_field(a.mom, 1, row + uint32_t(a.nx - 2)), 1, a.scale)) +
                double(a.scale.uref);
        }
        sr += a.rho0 + r;
        su += u;
    }
    red[0][t] = sr;
    red[1][t] = su;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] += red[0][t + w];
            red[1][t] += red[1][t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const size_t slot = size_t(*a.step_counter % FORCE_RING_STEPS);
        a.ring[2 * slot + 0] = red[0][0] / double(n);
        a.ring[2 * slot + 1] = red[1][0] / double(n);
    }
}

template <class R, class S>
void launch_planes(const PlaneArgs<R, S> &a, hipStream_t stream) {
    hipLaunchKernelGGL((plane_kernel<R, S>), dim3(1), dim3(256), 0, stream, a);
}

template <class R, class S>
void launch_encode(const double *in, const FieldSet<S> &out, uint32_t n, const StoreScale<R> &sc, hipStream_t stream) {
    hipLaunchKernelGGL((encode_kernel<R, S>), dim3((n + 255) / 256), dim3(256), 0, stream, in, out, n, sc);
}

template <class R, class S>
void launch_decode(const FieldSet<const S> &in, double *out, uint32_t n, const StoreScale<R> &sc, hipStream_t stream) {
    hipLaunchKernelGGL((decode_kernel<R, S>), dim3((n + 255) / 256), dim3(256), 0, stream, in, out, n, sc);
}

}  // namespace mrlbm

#define MRLBM_INSTANTIATE(R, S)                                                                             \
    template void launch_fused<R, S>(const FusedArgs<R, S> &, int, int, int, int, hipStream_t);             \
    template int prepare_fused<R, S>(int, int, int, int);                                                   \
    template void launch_memory_twin<R, S>(const FusedArgs<R, S> &, int, int, hipStream_t);                 \
    template bool fused_part_supported<R, S>(int, int, int);                                                \
    template void launch_wall<R, S>(const WallArgs<R, S> &, hipStream_t);                                   \
    template void launch_pressure<R, S>(const PressureArgs<R, S> &, hipStream_t);                           \
    template void launch_planes<R, S>(const PlaneArgs<R, S> &, hipStream_t);                                 \
    template void launch_fused_batch<R, S>(const FusedArgs<R, S> *, int, const uint4 *, int, int, int, int, int, \
                                           const BatchQueue &, hipStream_t);                                \
    template void launch_record_batch<R, S>(const FusedArgs<R, S> *, int, hipStream_t);                      \
    template int prepare_fused_batch<R, S>(int, int, int);                                                  \
    template void launch_wall_batch<R, S>(const WallArgs<R, S> *, const int2 *, int, hipStream_t);            \
    template void launch_encode<R, S>(const double *, const FieldSet<S> &, uint32_t, const StoreScale<R> &,        \
                                      hipStream_t);                                                         \
    template void launch_decode<R, S>(const FieldSet<const S> &, double *, uint32_t, const StoreScale<R> &,  \
                                      hipStream_t);
