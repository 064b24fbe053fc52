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
// (statistics/forces.cu:25-133) in FP64.
#pragma once

#include <algorithm>

#include "kernels.h"

namespace mrlbm {

namespace {

// Barrier that orders LDS only: global loads issued before it may stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int wrap(int v, int n) {
    v = v < 0 ? v + n : v;
    return v >= n ? v - n : v;
}

// Post-stream moments of one node from its 27 pulled populations f, node boundary condition,
// donor snapshot and collision, written to dst (the per-node part of streaming_and_evaluate_Mom
// and collision_halo_update, solver/mlbm.cu:83-164, 186-242). `special` is uniform per block:
// blocks without boundary nodes skip the node-type look-up.
template <class R, class S>
__device__ __forceinline__ void update_node(const FusedArgs<R, S> &a, const R *f, int x, int y, int z,
                                            uint32_t plane, bool special) {
    const uint32_t idx = uint32_t(x) + uint32_t(a.nx) * uint32_t(y) + uint32_t(z) * plane;
    uint32_t seed = 0;
    if constexpr (Conv<S>::scaled) {
        if (a.step_counter) seed = uint32_t(*a.step_counter) * 2u + 1u;
    }
    R M[10];
    if (!special) {
        moments(f, M);
        collide(a.omega, M);
        store_node<R, S>(a.dst, a.stride, idx, a.scale, a.rho0, M, seed);
    } else {
        const uint32_t info = a.info[idx];
        const uint32_t code = info & INFO_CODE_MASK;
        const uint32_t payload = info >> INFO_PAYLOAD_SHIFT;
        bool do_collide = true;
        if (code == NC_BULK) {
            moments(f, M);
        } else if (code == NC_KEEP || code == NC_FROZEN) {
            load_node<R, S>(a.src, a.stride, idx, a.scale, a.rho0, M);
            do_collide = code == NC_KEEP;
        } else if (code == NC_WALL) {
            // u keeps its old value; rho and m become the incoming rotated moments
            load_node<R, S>(a.src, a.stride, idx, a.scale, a.rho0, M);
            incoming_rotated(a.wall_nx[payload], a.wall_ny[payload], a.wall_in[payload], f, M);
            do_collide = false;  // finished by wall_step
        } else if (code == NC_BCF) {
            bc_fluid(payload, a.bcf_in[payload], f, M);
        } else if (code == NC_WEST) {
            R old[10];
            load_node<R, S>(a.src, a.stride, idx, a.scale, a.rho0, old);
            bc_west(f, old, a.inlet[uint32_t(y) + uint32_t(a.ny) * uint32_t(z)], M);
        } else if (code == NC_NORTH) {
            bc_ywall<face::K_N>(f, M);
        } else if (code == NC_SOUTH) {
            bc_ywall<face::K_S>(f, M);
        } else if (code == NC_NORTH_WEST) {
            bc_west_edge<face::K_W & face::K_N, -1>(f, a.omega, M);
        } else if (code == NC_SOUTH_WEST) {
            bc_west_edge<face::K_W & face::K_S, +1>(f, a.omega, M);
        } else {  // NC_EAST, NC_NORTH_EAST, NC_SOUTH_EAST
            R old[10], oxm[10];
            load_node<R, S>(a.src, a.stride, idx, a.scale, a.rho0, old);
            load_node<R, S>(a.src, a.stride, idx - 1, a.scale, a.rho0, oxm);
            if (code == NC_EAST)
                bc_east(f, old, oxm, M);
            else if (code == NC_NORTH_EAST)
                bc_east_edge<face::K_E & face::K_N, +1>(f, old, oxm, M);
            else
                bc_east_edge<face::K_E & face::K_S, -1>(f, old, oxm, M);
        }
        if (info & INFO_DONOR) store_node<R, S>(a.donor, a.stride, idx, a.scale, a.rho0, M, seed ^ 0x5A5A5A5Au);
        if (do_collide) collide(a.omega, M);
        store_node<R, S>(a.dst, a.stride, idx, a.scale, a.rho0, M, seed);
    }
}

// Ring-buffer variant: populations of the last planes kept in LDS (3 planes of the c_z=+1
// group, 2 of c_z=0, 1 of c_z=-1); one thread per tile node plus the ring spread over threads.
template <class R, class S, int BX, int BY>
__global__ __launch_bounds__(BX *BY) void fused_step_ring6_kernel(FusedArgs<R, S> a) {
    constexpr int NT = BX * BY;
    constexpr int W = BX + 2;
    constexpr int NH = W * (BY + 2);
    extern __shared__ unsigned char lds_raw[];
    R *lds = reinterpret_cast<R *>(lds_raw);
    // group planes: zp[3][9][NH], z0[2][9][NH], zm[9][NH]
    R *const gp = lds;
    R *const g0 = lds + 27 * NH;
    R *const gm = lds + 45 * NH;

    // block -> work item; consecutive items share an XCD so that neighbouring tiles reuse each
    // other's ring loads from the same L2
    int b = blockIdx.x;
    if (a.xcd_groups > 1) b = (b % a.xcd_groups) * (a.nblocks / a.xcd_groups) + b / a.xcd_groups;
    const uint2 item = a.items[b];
    const int tile = int(item.x & 0x7FFFFFFFu);
    const bool special = (item.x >> 31) != 0;
    const int x0 = (tile % a.tiles_x) * BX;
    const int y0 = (tile / a.tiles_x) * BY;
    const int z0 = int(item.y & 0xFFFFu);
    const int z1 = int(item.y >> 16);
    if (z0 >= z1) return;  // padding item (uniform per block)

    const int tid = threadIdx.x;
    const int tx = tid % BX, ty = tid / BX;
    const uint32_t plane = uint32_t(a.nx) * uint32_t(a.ny);

    // ring/tile positions loaded by this thread: p0 = tid, p1 = tid + NT (if inside the ring)
    const int p1 = tid + NT;
    const bool has1 = p1 < NH;
    uint32_t off0, off1 = 0;
    {
        const int hx = tid % W, hy = tid / W;
        off0 = uint32_t(wrap(x0 - 1 + hx, a.nx)) + uint32_t(a.nx) * uint32_t(wrap(y0 - 1 + hy, a.ny));
        if (has1) {
            const int hx1 = p1 % W, hy1 = p1 / W;
            off1 = uint32_t(wrap(x0 - 1 + hx1, a.nx)) + uint32_t(a.nx) * uint32_t(wrap(y0 - 1 + hy1, a.ny));
        }
    }

    R m0[10], m1[10];
    auto load_plane = [&](int z) {
        const uint32_t base = uint32_t(wrap(z, a.nz)) * plane;
        load_node<R, S>(a.src, a.stride, base + off0, a.scale, a.rho0, m0);
        if (has1) load_node<R, S>(a.src, a.stride, base + off1, a.scale, a.rho0, m1);
    };
    // write the populations of the loaded plane into the requested group slots
    auto put = [&](const R *m, int p, int sp, int s0, bool want0, bool wantm) {
        R f[Q];
        reconstruct(m, f);
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
    auto put_plane = [&](int sp, int s0, bool want0, bool wantm) {
        put(m0, tid, sp, s0, want0, wantm);
        if (has1) put(m1, p1, sp, s0, want0, wantm);
    };

    // prologue: plane z0-1 (c_z=+1 group) and plane z0 (c_z=+1 and 0 groups)
    load_plane(z0 - 1);
    put_plane(0, 0, false, false);
    load_plane(z0);
    put_plane(1, 0, true, false);
    load_plane(z0 + 1);

    const int x = x0 + tx, y = y0 + ty;
    const bool inside = x < a.nx && y < a.ny;
    const int h = (ty + 1) * W + (tx + 1);
    int sp_prev = 0, sp_cur = 1, sp_next = 2;  // slots of planes z-1, z, z+1 in the +1 group
    int s0_cur = 0;

    for (int z = z0; z < z1; ++z) {
        put_plane(sp_next, s0_cur ^ 1, true, true);  // plane z+1
        if (z + 1 < z1) load_plane(z + 2);            // prefetch, consumed next iteration
        lds_barrier();
        if (inside) {
            R f[Q];
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
            update_node(a, f, x, y, z, plane, special);
        }
        lds_barrier();
        const int t = sp_prev;
        sp_prev = sp_cur;
        sp_cur = sp_next;
        sp_next = t;
        s0_cur ^= 1;
        if (z + 1 < z1) {
            // registers now hold plane z+2; nothing to do until the next put
        }
    }
}

// Register-window variant (default): one thread per tile-or-ring position. Each thread keeps
// the c_z=+1 populations of its column for planes z-1 and z, the c_z=0 ones for plane z and the
// moments of plane z+1 in registers, so LDS holds one plane of 27 pulled populations per
// position; ring positions sit in their own waves after the tile's.
template <int BX, int BY>
constexpr int fused_v2_threads() {
    return ((BX + 2) * (BY + 2) + 63) / 64 * 64;
}

template <class R, class S, int BX, int BY>
__global__ __launch_bounds__((fused_v2_threads<BX, BY>())) void fused_step_kernel(FusedArgs<R, S> a) {
    constexpr int NT = BX * BY;
    constexpr int W = BX + 2;
    constexpr int NH = W * (BY + 2);
    extern __shared__ unsigned char lds_raw[];
    R *const G = reinterpret_cast<R *>(lds_raw);  // G[q * NH + p]: population q pulled from p

    int b = blockIdx.x;
    if (a.xcd_groups > 1) b = (b % a.xcd_groups) * (a.nblocks / a.xcd_groups) + b / a.xcd_groups;
    const uint2 item = a.items[b];
    const int tile = int(item.x & 0x7FFFFFFFu);
    const bool special = (item.x >> 31) != 0;
    const int x0 = (tile % a.tiles_x) * BX;
    const int y0 = (tile / a.tiles_x) * BY;
    const int z0 = int(item.y & 0xFFFFu);
    const int z1 = int(item.y >> 16);
    if (z0 >= z1) return;  // padding item (uniform per block)

    const int tid = threadIdx.x;
    const uint32_t plane = uint32_t(a.nx) * uint32_t(a.ny);
    // position of this thread in the (BX+2) x (BY+2) frame: tile nodes first, then the ring
    int hx = 0, hy = 0;
    const bool interior = tid < NT;
    const bool has_pos = tid < NH;
    if (interior) {
        hx = tid % BX + 1;
        hy = tid / BX + 1;
    } else if (has_pos) {
        const int r = tid - NT;
        if (r < BX) {
            hx = r + 1;
            hy = 0;
        } else if (r < 2 * BX) {
            hx = r - BX + 1;
            hy = BY + 1;
        } else if (r < 2 * BX + BY) {
            hx = 0;
            hy = r - 2 * BX + 1;
        } else if (r < 2 * BX + 2 * BY) {
            hx = BX + 1;
            hy = r - 2 * BX - BY + 1;
        } else {
            const int k = r - 2 * BX - 2 * BY;
            hx = (k & 1) ? BX + 1 : 0;
            hy = (k & 2) ? BY + 1 : 0;
        }
    }
    const int p = hy * W + hx;
    const uint32_t off =
        uint32_t(wrap(x0 - 1 + hx, a.nx)) + uint32_t(a.nx) * uint32_t(wrap(y0 - 1 + hy, a.ny));

    R mnext[10];
    auto load_plane = [&](int z) {
        load_node<R, S>(a.src, a.stride, uint32_t(wrap(z, a.nz)) * plane + off, a.scale, a.rho0, mnext);
    };
    R pprev[9], pcur[9], zcur[9];  // c_z=+1 of planes z-1, z; c_z=0 of plane z
    if (has_pos) {
        R f[Q];
        load_plane(z0 - 1);
        reconstruct(mnext, f);
#pragma unroll
        for (int i = 0; i < 9; ++i) pprev[i] = f[GROUP_ZP[i]];
        load_plane(z0);
        reconstruct(mnext, f);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            pcur[i] = f[GROUP_ZP[i]];
            zcur[i] = f[GROUP_Z0[i]];
        }
        load_plane(z0 + 1);
    }

    const int tx = hx - 1, ty = hy - 1;
    const int x = x0 + tx, y = y0 + ty;
    const bool update = interior && x < a.nx && y < a.ny;

    for (int z = z0; z < z1; ++z) {
        if (has_pos) {
            R f[Q];
            reconstruct(mnext, f);  // plane z+1
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                G[GROUP_ZP[i] * NH + p] = pprev[i];
                G[GROUP_Z0[i] * NH + p] = zcur[i];
                G[GROUP_ZM[i] * NH + p] = f[GROUP_ZM[i]];
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                pprev[i] = pcur[i];
                pcur[i] = f[GROUP_ZP[i]];
                zcur[i] = f[GROUP_Z0[i]];
            }
            if (z + 1 < z1) load_plane(z + 2);  // consumed next iteration
        }
        lds_barrier();
        if (update) {
            R f[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = G[q * NH + p - LATTICE.cx[q] - LATTICE.cy[q] * W];
            update_node(a, f, x, y, z, plane, special);
        }
        lds_barrier();
    }
}

template <class R, class S>
__device__ __forceinline__ R bilinear(const S *__restrict__ field, uint32_t base, uint32_t pitch, R xd, R yd) {
    const R q00 = R(Conv<S>::to(field[base]));
    const R q10 = R(Conv<S>::to(field[base + 1]));
    const R q01 = R(Conv<S>::to(field[base + pitch]));
    const R q11 = R(Conv<S>::to(field[base + pitch + 1]));
    const R q0 = (R(1) - xd) * q00 + xd * q10;
    const R q1 = (R(1) - xd) * q01 + xd * q11;
    return (R(1) - yd) * q0 + yd * q1;
}

// Interpolate u and m at one stencil point and rotate into the wall frame
// (initerpolate_and_rotate + rotate_velocity_moments, extrapolation_utils.cuh:63-107).
template <class R, class S>
__device__ __forceinline__ void interp_rotate(const WallArgs<R, S> &a, uint32_t base, R xd, R yd, R c, R s, R *vm) {
    R v[9];
    if constexpr (!Conv<S>::scaled) {
#pragma unroll
        for (int k = 0; k < 9; ++k)
            v[k] = bilinear<R, S>(a.donor + size_t(k + 1) * a.stride, base, uint32_t(a.nx), xd, yd);
    } else {
        // stored second moments are non-equilibrium parts: m = (m - u u) + u u is restored at each
        // corner before the corners are blended
        const uint32_t corner[4] = {base, base + 1, base + uint32_t(a.nx), base + uint32_t(a.nx) + 1};
        const R wgt[4] = {(R(1) - xd) * (R(1) - yd), xd * (R(1) - yd), (R(1) - xd) * yd, xd * yd};
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = R(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            R q[9];
#pragma unroll
            for (int k = 0; k < 9; ++k)
                q[k] = R(Conv<S>::to(a.donor[size_t(k + 1) * a.stride + corner[j]])) * a.scale.dec[field_class(k + 1)];
            q[3] += q[0] * q[0];
            q[4] += q[1] * q[1];
            q[5] += q[2] * q[2];
            q[6] += q[0] * q[1];
            q[7] += q[0] * q[2];
            q[8] += q[1] * q[2];
#pragma unroll
            for (int k = 0; k < 9; ++k) v[k] += wgt[j] * q[k];
        }
    }
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

constexpr int WALL_THREADS = 256;

template <class R, class S>
__global__ __launch_bounds__(WALL_THREADS) void wall_step_kernel(WallArgs<R, S> a) {
    const int i = blockIdx.x * WALL_THREADS + threadIdx.x;
    double F[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < a.nwall) {
        const uint32_t idx = a.node[i];
        R M[10];
        load_node<R, S>(a.mom, a.stride, idx, a.scale, a.rho0, M);  // rho_I, u_old, m'_I
        const R c = a.wnx[i], s = a.wny[i];
        R vm1[9], vm2[9];
        interp_rotate(a, a.st_base[2 * i], a.st_frac[4 * i], a.st_frac[4 * i + 1], c, s, vm1);
        interp_rotate(a, a.st_base[2 * i + 1], a.st_frac[4 * i + 2], a.st_frac[4 * i + 3], c, s, vm2);
        const int g = i < a.nb_inner ? 0 : 1;
        const R u0 = a.uw[g][0], u1 = a.uw[g][1], u2 = a.uw[g][2];
        // wall values u'_w and m'_w = u'_w u'_w (curvedLBM.cu:93-101)
        const R vw[7] = {u0, u1, u2, u0 * u0, u1 * u1, u2 * u2, u1 * u2};
        constexpr int sel[7] = {0, 1, 2, 3, 4, 5, 8};
        const R c0 = a.coef[3 * i], c1 = a.coef[3 * i + 1], c2 = a.coef[3 * i + 2];
        R p[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) p[j] = c0 * vw[j] + c1 * vm1[sel[j]] + c2 * vm2[sel[j]];
        R out[10];
        if constexpr (kExactOrder<R>) {
            wall_closure_loop(c, s, a.in_mask[i], p, M[0], M[7], M[8], out);
        } else {
            R k[WALL_CONSTS];
#pragma unroll
            for (int j = 0; j < WALL_CONSTS; ++j) k[j] = a.wconst[size_t(i) * WALL_CONSTS + j];
            wall_closure(c, s, k, p, M[0], M[7], M[8], out);
        }
        if (a.forces) masked_force(out, a.in_mask[i], 1.0, F);
        collide(a.omega, out);
        if (a.forces) masked_force(out, a.out_mask[i], -1.0, F);
        uint32_t seed = 0;
        if constexpr (Conv<S>::scaled) seed = uint32_t(*a.step_counter) * 2u + 2u;
        store_node<R, S>(a.mom, a.stride, idx, a.scale, a.rho0, out, seed);
    } else if (a.forces && i < a.nwall + a.nbcf) {
        const int k = i - a.nwall;
        const uint32_t idx = a.bcf_node[k];
        const uint32_t t = a.bcf_tag[k];
        R M[10];
        load_node<R, S>(a.donor, a.stride, idx, a.scale, a.rho0, M);
        masked_force(M, a.bcf_in[t], 1.0, F);
        collide(a.omega, M);
        masked_force(M, a.bcf_out[t], -1.0, F);
    }
    // block reduction in FP64: wave shuffles, then across the four waves through LDS. The last
    // block to finish (atomic ticket) adds the block partials in block order, so the per-step
    // record is deterministic, files it under the device step counter and advances the counter.
    __shared__ double red[4][WALL_THREADS / 64];
    if (a.forces) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double v = F[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            F[k] = v;
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0)
            for (int k = 0; k < 4; ++k) red[k][wave] = F[k];
    }
    __syncthreads();
    __shared__ bool last;
    if (threadIdx.x == 0) {
        if (a.forces) {
            for (int k = 0; k < 4; ++k) {
                double tot = 0.0;
                for (int w = 0; w < WALL_THREADS / 64; ++w) tot += red[k][w];
                a.partials[4 * blockIdx.x + k] = tot;
            }
        }
        __threadfence();
        const unsigned ticket = atomicAdd(a.ticket, 1u);
        last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        const unsigned long long step = __hip_atomic_load(a.step_counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a.forces) {
            double tot[4] = {0.0, 0.0, 0.0, 0.0};
            for (unsigned blk = 0; blk < gridDim.x; ++blk)
                for (int k = 0; k < 4; ++k)
                    tot[k] += __hip_atomic_load(&a.partials[4 * blk + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const size_t slot = size_t(step % FORCE_RING_STEPS);
            for (int k = 0; k < 4; ++k) a.ring[4 * slot + k] = tot[k];
        }
        *a.step_counter = step + 1;
        *a.ticket = 0u;  // ready for the next launch
    }
}

// Field-major double arrays (field 0 = rho - rho0, the stored deviation) <-> storage.
template <class R, class S>
__global__ void encode_kernel(const double *__restrict__ in, S *__restrict__ out, size_t stride, uint32_t n,
                              StoreScale<R> sc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) v[f] = in[size_t(f) * n + i];
    if constexpr (!Conv<S>::scaled) {
#pragma unroll
        for (int f = 0; f < 10; ++f) out[f * stride + i] = Conv<S>::from(v[f]);
    } else {
        R M[10];
        M[0] = R(1) + R(v[0]);
#pragma unroll
        for (int f = 1; f < 10; ++f) M[f] = R(v[f]);
        store_node<R, S>(out, stride, i, sc, R(1), M);
        out[i] = Conv<S>::from(R(v[0]) * sc.enc[0]);
    }
}

template <class R, class S>
__global__ void decode_kernel(const S *__restrict__ in, size_t stride, double *__restrict__ out, uint32_t n,
                              StoreScale<R> sc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (!Conv<S>::scaled) {
#pragma unroll
        for (int f = 0; f < 10; ++f) out[size_t(f) * n + i] = double(Conv<S>::to(in[f * stride + i]));
    } else {
        R M[10];
        load_node<R, S>(in, stride, i, sc, R(0), M);
        out[i] = double(R(Conv<S>::to(in[i])) * sc.dec[0]);
#pragma unroll
        for (int f = 1; f < 10; ++f) out[size_t(f) * n + i] = double(M[f]);
    }
}

}  // namespace

template <class R, class S>
void launch_fused(const FusedArgs<R, S> &a, int bx, int by, int variant, hipStream_t stream) {
    const size_t lds = fused_lds_bytes(bx, by, sizeof(R), variant);
    auto go = [&](auto kernel, int nt) {
        static bool attr_set = false;
        if (!attr_set) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            attr_set = true;
        }
        hipLaunchKernelGGL(kernel, dim3(a.nblocks), dim3(nt), lds, stream, a);
    };
#define MRLBM_FUSED_TILE(TX, TY)                                                     \
    if (bx == TX && by == TY) {                                                      \
        if (variant == 1)                                                            \
            go(fused_step_ring6_kernel<R, S, TX, TY>, TX * TY);                      \
        else                                                                         \
            go(fused_step_kernel<R, S, TX, TY>, fused_v2_threads<TX, TY>());         \
        return;                                                                      \
    }
    MRLBM_FUSED_TILE(32, 8)
    MRLBM_FUSED_TILE(16, 8)
    MRLBM_FUSED_TILE(32, 4)
    MRLBM_FUSED_TILE(16, 16)
    MRLBM_FUSED_TILE(64, 4)
    MRLBM_FUSED_TILE(16, 4)
#undef MRLBM_FUSED_TILE
}

template <class R, class S>
void launch_wall(const WallArgs<R, S> &a, hipStream_t stream) {
    const int n = a.nwall + (a.forces ? a.nbcf : 0);
    const int blocks = std::max(1, (n + WALL_THREADS - 1) / WALL_THREADS);
    hipLaunchKernelGGL((wall_step_kernel<R, S>), dim3(blocks), dim3(WALL_THREADS), 0, stream, a);
}

template <class R, class S>
void launch_encode(const double *in, S *out, size_t stride, uint32_t n, const StoreScale<R> &sc, hipStream_t stream) {
    hipLaunchKernelGGL((encode_kernel<R, S>), dim3((n + 255) / 256), dim3(256), 0, stream, in, out, stride, n, sc);
}

template <class R, class S>
void launch_decode(const S *in, size_t stride, double *out, uint32_t n, const StoreScale<R> &sc, hipStream_t stream) {
    hipLaunchKernelGGL((decode_kernel<R, S>), dim3((n + 255) / 256), dim3(256), 0, stream, in, stride, out, n, sc);
}

}  // namespace mrlbm

#define MRLBM_INSTANTIATE(R, S)                                                                             \
    template void launch_fused<R, S>(const FusedArgs<R, S> &, int, int, int, hipStream_t);                  \
    template void launch_wall<R, S>(const WallArgs<R, S> &, hipStream_t);                                   \
    template void launch_encode<R, S>(const double *, S *, size_t, uint32_t, const StoreScale<R> &, hipStream_t); \
    template void launch_decode<R, S>(const S *, size_t, double *, uint32_t, const StoreScale<R> &, hipStream_t);
