// kernels.h - device kernels of one time step (fused stream/collide and curved wall + forces).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "geometry.h"
#include "node_math.h"
#include "storage.h"

namespace mrlbm {

// Per-step force / mass records. The wall pass of a step leaves one FP64 partial per block and
// marks the record pending; the next launch on the case's stream (the next fused pass, or the
// record launch that ends a step call) sums the partials in a fixed order into the ring.
struct RecordArgs {
    unsigned long long *__restrict__ pending;  // step + 1 of the record waiting (0: none); null: no records
    const double *__restrict__ partials;       // 4 per wall block
    double *__restrict__ ring;                 // 4 per step, FORCE_RING_STEPS records
    int nblocks;                               // wall blocks of a step with forces
};

// Work item tile field: tile index | face kind << 28 | special << 31. A special tile (one that
// holds boundary or donor nodes) whose only boundary nodes form one domain face -- the inlet or
// outlet column or a y-wall row, without donors -- has a face kind, and runs the plain march with
// that face's closed form on its face lanes.
constexpr uint32_t kItemSpecial = 0x80000000u;
constexpr uint32_t kItemTileMask = 0x0FFFFFFFu;
constexpr int kItemFaceShift = 28;
constexpr int FACE_NONE = 0, FACE_WEST = 1, FACE_EAST = 2, FACE_NORTH = 3, FACE_SOUTH = 4;

// Arguments of the fused stream/collide pass over one ping-pong buffer pair.
template <class R, class S>
struct FusedArgs {
    FieldSet<const S> src;  // moments at the start of the step
    FieldSet<S> dst;        // moments after the step
    S *__restrict__ donor;  // pre-collision snapshot of donor nodes, node-major
    uint32_t pad;           // index of a padding slot after the last node (discarded stores)
    int nx, ny, nz;
    int plane_aligned;  // nx ny is a multiple of the chunk: a node's offsets repeat from plane to plane
    int tiles_x;     // tiles along x
    int nblocks;     // work items, one block each
    int xcd_groups;  // blocks are spread round-robin over this many XCDs (1: no remap)
    // work item b: x = tile field (tile index, face kind, special flag), y = z0 | (z1 << 16);
    // special items (tiles with boundary nodes, which need info[]) come first and may march
    // shorter z ranges
    const uint2 *__restrict__ items;
    const uint32_t *__restrict__ info;  // per-node NodeCode | flags | payload
    const R *__restrict__ wall_nx;      // wall_count entries (at least 1)
    const R *__restrict__ wall_ny;
    const uint32_t *__restrict__ wall_in;
    const uint32_t *__restrict__ bcf_in;  // 256 per-tag incoming masks
    const R *__restrict__ inlet;          // u_x per (y, z); inlet_count entries (at least 1)
    int wall_count, inlet_count;
    // index of this step on the device (the step counter of this buffer parity, set by the
    // previous step's wall pass); seeds stochastic rounding of the 16-bit forms, null for
    // round-to-nearest
    const unsigned long long *__restrict__ step_counter;
    StoreScale<R> scale;
    R omega, rho0;
    int bcf_strong;  // BCFLUID density from the Strong quadratic
    // records of the case: with a non-null `pending` the launch has one block more, which files
    // the previous step's record
    RecordArgs rec;
    // NaN/Inf watchdog: set to 1 by any wave that computes a non-finite density (mapped host
    // memory, read by the host without synchronising; null: no watchdog)
    uint32_t *__restrict__ nonfinite;
};

// Arguments of the curved-wall + force pass (one thread per wall node, then one per BCFLUID).
template <class R, class S>
struct WallArgs {
    FieldSet<S> mom;          // buffer the fused pass just wrote
    const S *__restrict__ donor;  // its pre-collision snapshot, node-major
    int nwall, nb_inner, nbcf;
    const uint32_t *__restrict__ node;
    const R *__restrict__ wnx;
    const R *__restrict__ wny;
    const uint32_t *__restrict__ in_mask;  // per wall node
    // per-wall records, component-major (component j of record r at j * nrec + r): one record per
    // wall node, or, with col0 > 0 (z-extruded walls), one per wall of plane 0 -- col0 inner then
    // col1 outer records -- whose node and stencil indices are offset by z * plane
    const R *__restrict__ coef;        // 3 extrapolation coefficients
    const R *__restrict__ wconst;      // WALL_CONSTS closure constants
    const uint32_t *__restrict__ out_mask;
    const uint32_t *__restrict__ st_base;  // 2: index of the (x0, y0) stencil corner
    const R *__restrict__ st_frac;         // 4: xd1, yd1, xd2, yd2
    int col0, col1;
    uint32_t plane;                        // nx * ny
    int nx;                                // row pitch of the stencil corners
    const uint32_t *__restrict__ bcf_node;
    const uint8_t *__restrict__ bcf_tag;
    const uint32_t *__restrict__ bcf_in;
    const uint32_t *__restrict__ bcf_out;
    R uw[2][3];
    StoreScale<R> scale;
    R omega, rho0;
    int forces;
    int strong;                     // MassBC::Strong closure instead of Equilibrium
    int nblocks;                    // blocks of this case in the wall pass
    double *__restrict__ partials;  // 4 per block
    // Step counters, one per buffer parity: this step reads its index from `step_counter` and
    // sets the other parity's counter, read by the next step, to index + 1. No launch writes a
    // counter that its own blocks read.
    const unsigned long long *__restrict__ step_counter;
    unsigned long long *__restrict__ next_counter;
    unsigned long long *__restrict__ pending;  // set to index + 1 on steps with forces
};

// Arguments of the wall-pressure statistics pass (one thread per inner-wall node).
template <class R, class S>
struct PressureArgs {
    FieldSet<const S> mom;            // state after the step (field 0: stored density)
    const uint32_t *__restrict__ node;  // inner-wall nodes
    const R *__restrict__ wnx;
    const R *__restrict__ wny;
    int n, nx, ny;
    double xc, yc, radius, rho0;
    StoreScale<R> scale;
    double *__restrict__ sum;  // per inner-wall node, FP64 running sum
};

// Arguments of the plane-mean statistics pass (one block): mean density on the inlet plane x = 0
// and mean u_x on the plane x = nx - 2 of the state after the step, filed in the plane ring at the
// step's index.
template <class R, class S>
struct PlaneArgs {
    FieldSet<const S> mom;  // state after the step (fields 0 and 1: density deviation, u_x)
    int nx, ny, nz;
    double rho0;
    StoreScale<R> scale;
    const unsigned long long *__restrict__ step_counter;  // the step's index (its parity's counter)
    double *__restrict__ ring;                             // 2 per step, FORCE_RING_STEPS records
};

// Steps of force history kept in the device ring.
constexpr int FORCE_RING_STEPS = 1 << 16;

// Threads per block of the wall pass (one wall or BCFLUID node each); the host sizes the
// per-block force partials with it.
constexpr int WALL_THREADS = 64;

inline int wall_blocks(int nwall, int nbcf) { return std::max(1, (nwall + nbcf + WALL_THREADS - 1) / WALL_THREADS); }

// Blocks are dispatched round-robin over the 8 XCDs; the wall pass maps dispatch slot b to the
// logical block that keeps consecutive wall nodes (which share stencil cache lines) on one XCD:
// XCD j takes a contiguous range of n / 8 (+1 for the first n mod 8 XCDs) logical blocks.
__host__ __device__ inline int xcd_contiguous(int b, int n) {
    const int q = n / 8, r = n % 8, j = b % 8, k = b / 8;
    return j * q + (j < r ? j : r) + k;
}

// Launches of the fused pass: every work item, or the interior / wall-side items of the
// two-stream schedule.
constexpr int PART_ALL = 0, PART_INTERIOR = 1, PART_WALL_SIDE = 2;

// variant 1: ring march; 3, 4: register-window march with one / two LDS planes. part: PART_ALL,
// or PART_INTERIOR / PART_WALL_SIDE where fused_part_supported.
template <class R, class S>
void launch_fused(const FusedArgs<R, S> &a, int bx, int by, int variant, int part, hipStream_t stream);

template <class R, class S>
bool fused_part_supported(int bx, int by, int variant);

// Memory-only twin of the fused pass over the same work items (no arithmetic): the access
// pattern's floor on this device. Copies src into dst.
template <class R, class S>
void launch_memory_twin(const FusedArgs<R, S> &a, int bx, int by, hipStream_t stream);

// Gives the fused kernel of this configuration the dynamic LDS it needs (once per kernel, before
// any launch or graph capture); 0 on success, -1 for a configuration that is not instantiated.
template <class R, class S>
int prepare_fused(int bx, int by, int variant, int part);

template <class R, class S>
void launch_wall(const WallArgs<R, S> &a, hipStream_t stream);

template <class R, class S>
void launch_pressure(const PressureArgs<R, S> &a, hipStream_t stream);

template <class R, class S>
void launch_planes(const PlaneArgs<R, S> &a, hipStream_t stream);

// Files a pending force record (one block; nothing when none is pending).
void launch_record(const RecordArgs &r, hipStream_t stream);

// Batched launches: one fused pass over the work items of many cases (item {case, tile |
// special << 31, z0 | z1 << 16, 0}; tiles BX x BY in {32x8, 16x8, 32x4}) plus one block per case
// that files the case's pending record, and one wall pass whose block b runs block blocks[b].y
// of case blocks[b].x. Argument arrays live on the device.
// Work queues of a persistent batched launch: pblocks resident blocks (a multiple of the XCD
// groups; block b serves group b mod groups, on XCD b mod 8) pull their group's work items in
// order through one counter per group. A launch uses the counters of its buffer parity and zeroes
// the other parity's, which the next step uses. pblocks = 0: one block per work item.
struct BatchQueue {
    int pblocks;
    unsigned *counters;  // 2 x groups
    int parity;
};

template <class R, class S>
void launch_fused_batch(const FusedArgs<R, S> *cases, int ncases, const uint4 *items, int nitems, int groups, int bx,
                        int by, int variant, const BatchQueue &queue, hipStream_t stream);

// Files the pending records of the cases (one block per case).
template <class R, class S>
void launch_record_batch(const FusedArgs<R, S> *cases, int ncases, hipStream_t stream);

template <class R, class S>
void launch_wall_batch(const WallArgs<R, S> *cases, const int2 *blocks, int nblocks, hipStream_t stream);

// Sets the dynamic-LDS limit of the batched fused kernel for this configuration (once, before any
// launch or graph capture); 0 on success.
template <class R, class S>
int prepare_fused_batch(int bx, int by, int variant);

inline bool batch_tile_supported(int bx, int by) {
    return (bx == 32 && by == 8) || (bx == 16 && by == 8) || (bx == 32 && by == 4);
}

template <class R, class S>
void launch_encode(const double *in, const FieldSet<S> &out, uint32_t n, const StoreScale<R> &sc, hipStream_t stream);

template <class R, class S>
void launch_decode(const FieldSet<const S> &in, double *out, uint32_t n, const StoreScale<R> &sc, hipStream_t stream);

// Point arrays of a VTK output file (output.hip): four float32 arrays, field-major, x fastest.
struct VtiArgs {
    uint32_t n;          // nodes
    int nx, ny;          // grid (x fastest)
    int layout;          // MRLBM_VTI_CARTESIAN or MRLBM_VTI_ANNULUS
    double xc, yc;       // axis of the cylindrical projection
    float vel_norm;      // velocity divisor of the cylindrical components
    const uint32_t *solid;  // one bit per node: SOLID points are written as NaN
    float *out;             // 4 * n values
};

template <class R, class S>
void launch_vti_points(const FieldSet<const S> &in, const StoreScale<R> &sc, R rho0, const VtiArgs &a,
                       hipStream_t stream);

// Launch configuration of the fused pass (chosen by defaults, the environment or the tuner).
struct LaunchConfig {
    int bx = 32, by = 8;    // tile
    int variant = 1;        // 1 ring march, 3 / 4 register-window march with 1 / 2 LDS planes
    int lz = 32;            // z chunk of a work item
    int lz_special = 0;     // z chunk of boundary tiles (0: lz)
    bool xcd_remap = true;  // one slab per XCD
    int blocks_per_cu = 0;  // batches: persistent fused launch with this many blocks per CU (0: one block per item)
    int packing = 0;        // batches: 0 deep cases cut into XCD slabs, 1 whole cases on the least loaded XCD
    int slab_axis = 2;      // axis cut into the XCD slabs: 2 = z, 1 = y (rows of tiles, z chunks in rounds)
    bool face_tiles = true;  // face marches for tiles whose only boundary nodes form one face
    int special_order = 0;  // boundary tiles: 0 first, 1 in tile order, 2 last (in a z slab; in a y slab per z round, or after every round)
    double special_weight = 0;  // y slabs: plain tiles' worth of a boundary-path tile (0: by storage)
};

// Tile shapes (BX x BY) the fused kernel is instantiated for.
inline bool fused_tile_supported(int bx, int by) {
    return (bx == 32 && by == 8) || (bx == 16 && by == 8) || (bx == 32 && by == 4) || (bx == 16 && by == 16) ||
           (bx == 64 && by == 4) || (bx == 16 && by == 4) || (bx == 32 && by == 16);
}

// Bytes of dynamic LDS one fused block uses over the tile and its ring: three planes of the
// c_z=+1 population group, two of the c_z=0 group and one of the c_z=-1 group (ring march,
// variant 1), or one / two planes of all 27 pulled populations (register-window march,
// variants 3 / 4); then the density and velocity of the tile's west ring column for three
// planes.
inline size_t fused_lds_bytes(int bx, int by, size_t real_bytes, int variant) {
    // ring march: 54 population slots of every frame position; register-window march: one or two
    // planes of nine (c_x, c_z) blocks of three population rows-sets and a spare row (WindowLds)
    const size_t values = variant == 1 ? 54 * size_t(bx + 2) * size_t(by + 2)
                                       : size_t(variant == 3 ? 1 : 2) * 9 * size_t(bx + 2) * size_t(3 * by + 1);
    // + the west ring column's density and velocity of three planes (boundary tiles)
    return (values + 12 * size_t(by)) * real_bytes;
}

}  // namespace mrlbm
