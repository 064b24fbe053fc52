// sim.h - one simulation on the GPU: device buffers, case tables, work items and the queueing
// of the step kernels on one HIP stream (the host side of the reference's driver, main.cu:7-40,
// simulation.cu:13-172, solve_mrlbm.cuh). Shared by the single-case C ABI (capi.hip) and the
// batched launcher (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mrlbm.h"
#include "geometry.h"
#include "kernels.h"

namespace mrlbm {

// Last error message of the calling thread (mrlbm_last_error).
inline thread_local std::string g_last_error;

inline int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_error(MRLBM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace mrlbm

namespace mrlbm {

// Reference velocity (x component) of a case: the mean inflow of the cylinder channel (u_max for
// a uniform inlet, 2/3 u_max for the parabolic one), 0 otherwise.
inline double reference_velocity(const mrlbm_case_desc &d) {
    if (d.kind != MRLBM_CASE_CYLINDER) return 0.0;
    return d.inlet_profile == MRLBM_INLET_PARABOLIC ? d.u_max * (2.0 / 3.0) : d.u_max;
}

class SimBase {
  public:
    virtual ~SimBase() = default;
    virtual int setup() = 0;
    virtual int init() = 0;
    virtual int step(int n, uint32_t flags) = 0;
    virtual int get(int f, double *h) = 0;
    virtual int set(int f, const double *h) = 0;
    // the current moments in the storage type, field-major, 10 * n values (checkpoint payload)
    virtual int get_raw(void *h) = 0;
    virtual int set_raw(const void *h) = 0;
    virtual int get_forces(int64_t first, int count, double *out4) = 0;
    virtual int time_steps(int n, uint32_t flags, double *ms_total, double *ms_k) = 0;
    virtual int sync() = 0;
    virtual int set_steps(int64_t n) = 0;
    virtual int stats_reset() = 0;
    virtual int get_surface_pressure(double *mean, int64_t *samples) = 0;
    virtual int get_plane_means(int64_t first, int count, double *out2) = 0;
    // all 10 fields at once, x fastest, field 0 = rho - rho0 (both buffers)
    virtual int set_all(const std::vector<double> &all10) = 0;
    // four float32 point arrays of a VTK file from the current state (MRLBM_VTI_* layout)
    virtual int vti_points(int layout, float vel_norm, float *host) = 0;
    virtual LaunchConfig current_launch() const = 0;
    // Steps as the library runs them with every graph captured beforehand (the timed region holds
    // no capture): ms_total for all n steps, ms_step the median over the graph replays of one
    // replay's duration per step (0 without a replay), replays their count.
    virtual int time_replays(int n, uint32_t flags, double *ms_total, double *ms_step, int *replays) = 0;
    // NaN/Inf watchdog: raised by the fused pass when a node's density is not finite; cleared when
    // a new state is set (init, set_moments, checkpoint reads).
    virtual bool nonfinite() const = 0;
    // u_ref of the 16-bit forms' stored u_x - u_ref (0 for FP64 and FP32)
    virtual double velocity_reference() const = 0;
    // Device tables and work items of a new geometry g, then a fresh initial state.
    virtual int apply_geometry() = 0;
    // Mean duration (ms) of one launch of the fused pass's memory-only twin over `launches`
    // launches (after one untimed); the state is unchanged.
    virtual int time_memory_twin(int launches, double *ms) = 0;
    int device = 0;  // HIP device the simulation's memory lives on
    int check_finite() const {
        return nonfinite() ? set_error(MRLBM_ERR_NONFINITE, "non-finite moments: the NaN/Inf watchdog tripped")
                           : int(MRLBM_OK);
    }
    mrlbm_case_desc d{};
    Geometry g;
    int64_t steps = 0;
    int bytes_per_value = 0;
    bool device_ready = false;
    // stream of the batch that launches this case's kernels (null for a standalone case); reads
    // wait for it first
    hipStream_t batch_stream = nullptr;
    int wait_batch() {
        if (batch_stream) HIP_TRY(hipStreamSynchronize(batch_stream));
        return MRLBM_OK;
    }
};

template <class R, class S>
class Sim final : public SimBase {
  public:
    ~Sim() override { release(); }

    int setup() override {
        bytes_per_value = int(sizeof(S));
        HIP_TRY(hipGetDevice(&device));
        HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ev_wall_side_, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&ev_interior_, hipEventDisableTiming));
        // watchdog flag in mapped, coherent host memory: kernels store to it, the host reads it
        // at any time
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&h_nonfinite_), sizeof(uint32_t),
                              hipHostMallocMapped | hipHostMallocCoherent));
        *h_nonfinite_ = 0u;
        HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void **>(&d_nonfinite_), h_nonfinite_, 0));
        const int64_t n = g.n;
        // at least one padding slot after the last node: bulk threads outside the grid store there
        stride_ = size_t((n + 1 + 63) / 64 * 64);
        // node indices are 32-bit
        if (uint64_t(n) + 1 > uint64_t(0xFFFFFFFFu))
            return set_error(MRLBM_ERR_UNSUPPORTED, "grid too large: node indices must fit 32 bits");
        elems_ = field_buffer_elems<S>(n);
        HIP_TRY(hipMalloc(&buf_[0], elems_ * sizeof(S)));
        HIP_TRY(hipMalloc(&buf_[1], elems_ * sizeof(S)));
        HIP_TRY(hipMalloc(&donor_, 10 * stride_ * sizeof(S)));
        HIP_TRY(hipMemsetAsync(donor_, 0, 10 * stride_ * sizeof(S), stream_));
        omega_ = R(1.0 / d.tau);
        rho0_ = R(d.rho0);
        // 16-bit storage scales of the rho, u and m deviations: FP16 maps deviations of O(1e-2)
        // to O(1); the BF16 form stores rho and m as they are (and its velocity in FP16 bits)
        scale_.enc[0] = R(Conv<S>::unit_scale ? 1 : 64);
        scale_.enc[1] = R(Conv<S>::unit_scale ? 1 : 16);
        scale_.enc[2] = R(Conv<S>::unit_scale ? 1 : 256);
        for (int k = 0; k < 3; ++k) scale_.dec[k] = R(1) / scale_.enc[k];
        // the 16-bit forms store u_x - u_ref (MRLBM_UREF=0: u_x itself)
        const char *ue = std::getenv("MRLBM_UREF");
        scale_.uref = (Conv<S>::scaled && !(ue && std::atoi(ue) == 0)) ? R(reference_velocity(d)) : R(0);

        // ---- launch configuration of the fused pass (defaults, overridden by the environment)
        LaunchConfig lc = default_launch();
        if (const char *e = std::getenv("MRLBM_TILE")) {
            int tx = 0, ty = 0;
            if (std::sscanf(e, "%dx%d", &tx, &ty) != 2 || !fused_tile_supported(tx, ty))
                return set_error(MRLBM_ERR_INVALID, std::string("unsupported MRLBM_TILE ") + e);
            lc.bx = tx;
            lc.by = ty;
        }
        if (const char *e = std::getenv("MRLBM_FUSED_VARIANT")) {
            const int v = std::atoi(e);
            lc.variant = (v == 3 || v == 4) ? v : 1;
        }
        if (const char *e = std::getenv("MRLBM_LZ")) lc.lz = std::max(1, std::atoi(e));
        if (const char *e = std::getenv("MRLBM_LZ_SPECIAL")) lc.lz_special = std::max(1, std::atoi(e));
        if (const char *e = std::getenv("MRLBM_XCD_REMAP")) lc.xcd_remap = std::atoi(e) != 0;
        if (const char *e = std::getenv("MRLBM_SLAB_AXIS")) lc.slab_axis = (e[0] == 'y' || e[0] == '1') ? 1 : 2;
        if (const char *e = std::getenv("MRLBM_FACE_TILES")) lc.face_tiles = std::atoi(e) != 0;
        if (const char *e = std::getenv("MRLBM_SPECIAL_WEIGHT")) lc.special_weight = std::max(0.0, std::atof(e));
        if (const char *e = std::getenv("MRLBM_SPECIAL_ORDER")) lc.special_order = std::max(0, std::min(2, std::atoi(e)));
        if (const char *e = std::getenv("MRLBM_STOCHASTIC_ROUNDING")) stochastic_ = Conv<S>::scaled && std::atoi(e) != 0;
        if (const char *e = std::getenv("MRLBM_GRAPH_STEPS")) graph_steps_ = std::max(0, std::atoi(e) / 2 * 2);
        if (const char *e = std::getenv("MRLBM_TWO_STREAM")) two_stream_env_ = std::atoi(e) != 0 ? 1 : 0;

        HIP_TRY(hipMalloc(&d_ring_, size_t(FORCE_RING_STEPS) * 4 * sizeof(double)));
        HIP_TRY(hipMalloc(&d_plane_ring_, size_t(FORCE_RING_STEPS) * 2 * sizeof(double)));
        if (int rc = build_tables()) return rc;
        HIP_TRY(hipMalloc(&d_step_counter_, 2 * sizeof(unsigned long long)));
        HIP_TRY(hipMemsetAsync(d_step_counter_, 0, 2 * sizeof(unsigned long long), stream_));
        HIP_TRY(hipMalloc(&d_pending_, sizeof(unsigned long long)));
        HIP_TRY(hipMemsetAsync(d_pending_, 0, sizeof(unsigned long long), stream_));
        HIP_TRY(hipMemsetAsync(d_ring_, 0, size_t(FORCE_RING_STEPS) * 4 * sizeof(double), stream_));
        HIP_TRY(hipMemsetAsync(d_plane_ring_, 0, size_t(FORCE_RING_STEPS) * 2 * sizeof(double), stream_));
        HIP_TRY(hipStreamSynchronize(stream_));
        return configure_launch(lc);
    }

    // Device tables of the geometry g: per-node info words with donor flags, wall records and
    // stencils, BCFLUID lists and masks, the inlet profile, force partials and the wall-pressure
    // sums. Tables of an earlier geometry are freed first (mrlbm_set_geometry).
    int build_tables() {
        for (void **pp : {(void **)&d_info_, (void **)&d_wconst_, (void **)&d_wnx_, (void **)&d_wny_, (void **)&d_win_,
                          (void **)&d_wout_, (void **)&d_wnode_, (void **)&d_wcoef_, (void **)&d_stbase_,
                          (void **)&d_stfrac_, (void **)&d_bcfnode_, (void **)&d_bcftag_, (void **)&d_bcfin_,
                          (void **)&d_bcfout_, (void **)&d_inlet_, (void **)&d_partials_, (void **)&d_ps_sum_}) {
            if (*pp) HIP_TRY(hipFree(*pp));
            *pp = nullptr;
        }
        // stencils of the wall nodes and donor marking
        std::vector<uint32_t> info = g.info;
        const int nwall = int(g.wall[0].node.size() + g.wall[1].node.size());
        std::vector<uint32_t> w_node, w_in, w_out, st_base;
        std::vector<R> w_nx, w_ny, w_coef, st_frac, w_const;
        for (int gi = 0; gi < g.ngroups; ++gi) {
            const WallTable &w = g.wall[gi];
            for (size_t k = 0; k < w.node.size(); ++k) {
                const int64_t i = w.node[k];
                const int z = int(i / (int64_t(g.nx) * g.ny));
                const R nx = R(w.nx[k]), ny = R(w.ny[k]), delta = R(w.delta[k]);
                w_node.push_back(uint32_t(i));
                w_nx.push_back(nx);
                w_ny.push_back(ny);
                w_in.push_back(w.in_mask[k]);
                w_out.push_back(w.out_mask[k]);
                // extrapolation_quadratic coefficients (extrapolation_utils.cuh:49-61)
                const R dd = delta / R(1), d2 = dd * dd;
                w_coef.push_back(R(1) - R(0.5) * d2 + R(1.5) * dd);
                w_coef.push_back(d2 - R(2) * dd);
                w_coef.push_back(-R(0.5) * (d2 - dd));
                double kc[WALL_CONSTS];
                wall_constants(w.nx[k], w.ny[k], w.in_mask[k], kc);
                for (int j = 0; j < WALL_CONSTS; ++j) w_const.push_back(R(kc[j]));
                // wall point and the two fluid points along the normal (curvedLBM.cu:78-84)
                const R xw = R(w.xc) + R(w.radius) * nx, yw = R(w.yc) + R(w.radius) * ny;
                const R dir = R(w.fluid_dir);
                const R pts[2][2] = {{xw + dir * R(1) * nx, yw + dir * R(1) * ny},
                                     {xw + dir * R(2) * R(1) * nx, yw + dir * R(2) * R(1) * ny}};
                for (int p = 0; p < 2; ++p) {
                    const int x0 = int(std::floor(pts[p][0])), y0 = int(std::floor(pts[p][1]));
                    if (x0 < 0 || y0 < 0 || x0 + 1 >= g.nx || y0 + 1 >= g.ny)
                        return set_error(MRLBM_ERR_GEOMETRY, "wall interpolation stencil leaves the domain");
                    const int64_t base = int64_t(x0) + int64_t(g.nx) * (int64_t(y0) + int64_t(g.ny) * z);
                    st_base.push_back(uint32_t(base));
                    st_frac.push_back(pts[p][0] - R(x0));
                    st_frac.push_back(pts[p][1] - R(y0));
                    for (int64_t c : {base, base + 1, base + g.nx, base + g.nx + 1}) info[c] |= INFO_DONOR;
                }
            }
        }
        // a wall node hands its incoming moments to the wall pass through its donor snapshot
        for (int gi = 0; gi < g.ngroups; ++gi)
            for (int64_t i : g.wall[gi].node) info[i] |= INFO_DONOR;
        std::vector<uint32_t> bcf_node;
        std::vector<uint8_t> bcf_tag;
        for (int gi = 0; gi < g.ngroups; ++gi) {
            const WallTable &w = g.wall[gi];
            for (size_t k = 0; k < w.bcf_node.size(); ++k) {
                bcf_node.push_back(uint32_t(w.bcf_node[k]));
                bcf_tag.push_back(uint8_t(w.bcf_tag[k]));
                info[w.bcf_node[k]] |= INFO_DONOR;  // force sums need their pre-collision moments
            }
        }
        g.num_donors = 0;
        for (uint32_t v : info) g.num_donors += (v & INFO_DONOR) ? 1 : 0;
        nwall_ = nwall;
        nb_inner_ = int(g.wall[0].node.size());
        nbcf_ = int(bcf_node.size());

        // columns (x, y) that hold a boundary or donor node at some z make a tile special;
        // columns frozen at every z (inactive tiles of the reference) are never updated: both
        // buffers hold their values, and tiles made only of them get no work items
        // (per special column: the set of node codes it holds, bit 31 for a donor node)
        {
            const size_t ncol = size_t(g.nx) * g.ny;
            std::vector<uint32_t> col(ncol, 0);
            frozen_cols_.assign(ncol, 1);
            for (int64_t i = 0; i < g.n; ++i) {
                const size_t c = size_t(i % int64_t(ncol));
                if (info[i] != NC_BULK) col[c] |= (1u << (info[i] & INFO_CODE_MASK)) | ((info[i] & INFO_DONOR) ? 0x80000000u : 0u);
                if ((info[i] & INFO_CODE_MASK) != NC_FROZEN) frozen_cols_[c] = 0;
            }
            special_cols_.clear();
            special_col_codes_.clear();
            for (size_t c = 0; c < ncol; ++c)
                if (col[c]) {
                    special_cols_.push_back(uint32_t(c));
                    special_col_codes_.push_back(col[c]);
                }
        }

        std::vector<R> wall_nx_all(std::max(nwall, 1)), wall_ny_all(std::max(nwall, 1));
        std::vector<uint32_t> wall_in_all(std::max(nwall, 1));
        for (int k = 0; k < nwall; ++k) {
            wall_nx_all[k] = w_nx[k];
            wall_ny_all[k] = w_ny[k];
            wall_in_all[k] = w_in[k];
        }
        std::vector<R> inlet(std::max<size_t>(g.inlet.size(), 1), R(0));
        for (size_t k = 0; k < g.inlet.size(); ++k) inlet[k] = R(g.inlet[k]);

        // z-extruded walls (the cylinder and annulus cases): every plane lists the same wall columns
        // in the same order with the same normals, depths and masks. The wall pass then reads the
        // records of one plane's walls (a few hundred KB that stay in L2) and offsets the node and
        // stencil indices by the plane, instead of reading one record per wall node and step.
        wall_cols_[0] = wall_cols_[1] = 0;
        {
            bool cols = g.ngroups > 0;
            const int64_t plane = int64_t(g.nx) * g.ny;
            for (int gi = 0; gi < g.ngroups && cols; ++gi) {
                const WallTable &w = g.wall[gi];
                const size_t n = w.node.size();
                cols = n % size_t(g.nz) == 0;
                const size_t nwp = cols ? n / size_t(g.nz) : 1;
                for (size_t k = nwp; k < n && cols; ++k) {
                    const size_t k0 = k % nwp;
                    cols = w.node[k] == w.node[k0] + int64_t(k / nwp) * plane && w.nx[k] == w.nx[k0] &&
                           w.ny[k] == w.ny[k0] && w.delta[k] == w.delta[k0] && w.in_mask[k] == w.in_mask[k0] &&
                           w.out_mask[k] == w.out_mask[k0];
                }
                wall_cols_[gi] = int(nwp);
            }
            if (!cols) wall_cols_[0] = wall_cols_[1] = 0;
        }
        // the records of the walls that the wall pass reads (every wall, or one plane's)
        const int nrec = wall_cols_[0] > 0 ? wall_cols_[0] + wall_cols_[1] : nwall;
        if (wall_cols_[0] > 0) {
            auto pick = [&](const auto &v, int per) {
                std::decay_t<decltype(v)> out;
                for (int gi = 0; gi < 2; ++gi)
                    for (int k = 0; k < wall_cols_[gi]; ++k) {
                        const size_t i = size_t(gi == 0 ? 0 : nb_inner_) + size_t(k);
                        for (int j = 0; j < per; ++j) out.push_back(v[i * per + j]);
                    }
                return out;
            };
            w_const = pick(w_const, WALL_CONSTS);
            w_coef = pick(w_coef, 3);
            st_frac = pick(st_frac, 4);
            st_base = pick(st_base, 2);
            w_out = pick(w_out, 1);
        }
        // per-wall records go to the device component-major (entry j of record i at j * nrec + i) so
        // that the wall pass reads them with coalesced loads
        auto soa = [nrec](const auto &v, int per) {
            std::decay_t<decltype(v)> out(v.size());
            for (int i = 0; i < nrec; ++i)
                for (int j = 0; j < per; ++j) out[size_t(j) * nrec + i] = v[size_t(i) * per + j];
            return out;
        };
        w_const = soa(w_const, WALL_CONSTS);
        w_coef = soa(w_coef, 3);
        st_frac = soa(st_frac, 4);
        st_base = soa(st_base, 2);

        int rc;
        if ((rc = upload(d_info_, info))) return rc;
        if ((rc = upload(d_wconst_, w_const))) return rc;
        if ((rc = upload(d_wnx_, wall_nx_all))) return rc;
        if ((rc = upload(d_wny_, wall_ny_all))) return rc;
        if ((rc = upload(d_win_, wall_in_all))) return rc;
        if ((rc = upload(d_wout_, w_out))) return rc;
        if ((rc = upload(d_wnode_, w_node))) return rc;
        if ((rc = upload(d_wcoef_, w_coef))) return rc;
        if ((rc = upload(d_stbase_, st_base))) return rc;
        if ((rc = upload(d_stfrac_, st_frac))) return rc;
        if ((rc = upload(d_bcfnode_, bcf_node))) return rc;
        if ((rc = upload(d_bcftag_, bcf_tag))) return rc;
        std::vector<uint32_t> bin(g.bcf_in, g.bcf_in + 256), bout(g.bcf_out, g.bcf_out + 256);
        if ((rc = upload(d_bcfin_, bin))) return rc;
        if ((rc = upload(d_bcfout_, bout))) return rc;
        if ((rc = upload(d_inlet_, inlet))) return rc;
        // the wall pass launches at most wall_blocks(nwall, nbcf) blocks (forces on)
        HIP_TRY(hipMalloc(&d_partials_, size_t(wall_blocks(nwall, nbcf_)) * 4 * sizeof(double)));
        // wall-pressure statistics of the inner wall: fluid points 1..3 spacings out must lie inside
        stats_ok_ = g.ngroups > 0 && !g.wall[0].node.empty();
        for (size_t k = 0; stats_ok_ && k < g.wall[0].node.size(); ++k) {
            const double xw = g.wall[0].xc + g.wall[0].radius * g.wall[0].nx[k],
                         yw = g.wall[0].yc + g.wall[0].radius * g.wall[0].ny[k];
            for (int p = 1; p <= 3; ++p) {
                const int x0 = int(std::floor(xw + p * g.wall[0].nx[k])), y0 = int(std::floor(yw + p * g.wall[0].ny[k]));
                if (x0 < 0 || y0 < 0 || x0 + 1 >= g.nx || y0 + 1 >= g.ny) stats_ok_ = false;
            }
        }
        HIP_TRY(hipMalloc(&d_ps_sum_, std::max<size_t>(g.wall[0].node.size(), 1) * sizeof(double)));
        HIP_TRY(hipMemsetAsync(d_ps_sum_, 0, std::max<size_t>(g.wall[0].node.size(), 1) * sizeof(double), stream_));
        HIP_TRY(hipStreamSynchronize(stream_));
        stats_samples_ = 0;
        return MRLBM_OK;
    }

    // Fused-pass defaults: 16x8 tiles in FP64 and 32x8 in FP32 arithmetic (LDS-bound; 16x8 for
    // cache-resident FP32 grids), the ring march in FP64 and register windows in FP32 arithmetic.
    LaunchConfig default_launch() const {
        LaunchConfig lc;
        lc.bx = sizeof(R) == 8 ? 16 : 32;
        lc.by = 8;
        // FP32 storage on a grid whose two moment buffers fit in MI355X's 256 MB last-level cache:
        // the march is then bound by its boundary tiles, not by HBM row segments, and 16-wide
        // tiles run faster (256x128x64 cylinder +3.7%); grids that stream from HBM keep 32-wide
        // rows (512x256x256 wake: 16x8 is 5% slower)
        constexpr size_t kLastLevelCacheBytes = size_t(256) << 20;
        const bool cache_resident = 2 * field_buffer_elems<S>(g.n) * sizeof(S) <= kLastLevelCacheBytes;
        if (sizeof(R) == 4 && !Conv<S>::scaled && cache_resident) lc.bx = 16;
        // register-window march for FP32 arithmetic: with 16-bit storage one LDS plane, whose
        // four blocks per CU hide the decode and encode work better than the two-plane form's two;
        // with FP32 storage two LDS planes (one barrier per plane) on grids that stream from HBM
        // (512x256x256 wake: 1% faster than one plane), one plane on cache-resident grids, whose
        // boundary tiles bound them (256x128x64 cylinder: 37.6 against 40.8 us per fused pass)
        lc.variant = sizeof(R) == 8 ? 1 : ((Conv<S>::scaled || cache_resident) ? 3 : 4);
        // z chunks short enough for about 1024 work items (a block's plane-by-plane march is
        // latency-bound, so small grids need many short chunks), at most 32 planes
        const int64_t tiles = int64_t((g.nx + lc.bx - 1) / lc.bx) * ((g.ny + lc.by - 1) / lc.by);
        lc.lz = int(std::max<int64_t>(1, std::min<int64_t>(32, tiles * g.nz / 1024)));
        // boundary-path tiles in chunks of the same length (FP64), or of a quarter of it on
        // cache-resident grids in FP32 arithmetic, whose boundary tiles bound the pass: shorter
        // items of the slower tiles leave a shorter tail (256x128x64 cylinder, A/B on one box:
        // FP32 storage 47,920-48,090 -> 48,602-48,820 MLUPS, BF16 50,944-50,983 -> 51,172-51,481,
        // FP16 45,867-46,072 -> 47,065-47,195; FP64 12,289-12,300 -> 12,056-12,063, so not there)
        lc.lz_special = (sizeof(R) == 4 && cache_resident) ? std::max(1, lc.lz / 4) : 0;
        lc.xcd_remap = true;
        // FP32 arithmetic on grids that stream from HBM: y slabs -- each XCD marches a band of
        // tile rows (cut to equal estimated work) through z in rounds, so a chunk's halo planes
        // were loaded by the same XCD one round before -- with boundary tiles first in each round,
        // those on the boundary path in half chunks (512x256x256 wake, A/B on one box: FP32 fused
        // pass 608.5 -> 592.1 us, BF16 387.4 -> 382.3 us against z slabs; then whole chunks for
        // face marches: FP32 591.5 -> 582.3 us, BF16 368.2 -> 361.2 us)
        if (sizeof(R) == 4 && !cache_resident && (g.ny + lc.by - 1) / lc.by >= 8) {
            lc.slab_axis = 1;
            lc.special_order = 0;
            lc.lz_special = std::max(1, lc.lz / 2);
        }
        return lc;
    }

    // Tile shape, march variant and work items of the fused pass. The grid is cut into one slab
    // per XCD (blocks are dispatched round-robin over the XCDs, so slab j runs on XCD j and its
    // tiles share that L2): z slabs, inside which every tile is cut into z chunks of lz planes
    // (lz_special for tiles on the boundary path), or y slabs marched in z rounds.
    // The boundary tiles go first.
    int configure_launch(const LaunchConfig &want) {
        LaunchConfig lc = want;
        if (!fused_tile_supported(lc.bx, lc.by)) return set_error(MRLBM_ERR_INVALID, "unsupported fused tile");
        int dev = 0, max_lds = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (fused_lds_bytes(lc.bx, lc.by, sizeof(R), lc.variant) > size_t(max_lds))
            return set_error(MRLBM_ERR_UNSUPPORTED, "fused tile needs more LDS than a block may use");
        if (prepare_fused<R, S>(lc.bx, lc.by, lc.variant, PART_ALL) != 0)
            return set_error(MRLBM_ERR_UNSUPPORTED, "fused kernel of this configuration cannot get its LDS");
        const int tiles_x = (g.nx + lc.bx - 1) / lc.bx;
        const int tiles_y = (g.ny + lc.by - 1) / lc.by;
        const int ntiles = tiles_x * tiles_y;
        if (uint32_t(ntiles) > kItemTileMask) return set_error(MRLBM_ERR_UNSUPPORTED, "too many tiles in a z plane");
        std::vector<uint8_t> tile_special(ntiles, 0), tile_live(ntiles, 0);
        std::vector<uint32_t> tile_codes(ntiles, 0);
        for (size_t k = 0; k < special_cols_.size(); ++k) {
            const uint32_t c = special_cols_[k];
            const int x = int(c % uint32_t(g.nx)), y = int(c / uint32_t(g.nx));
            tile_special[x / lc.bx + tiles_x * (y / lc.by)] = 1;
            tile_codes[x / lc.bx + tiles_x * (y / lc.by)] |= special_col_codes_[k];
        }
        // face kind of a special tile: its only boundary nodes (no donor) are inlet nodes, outlet
        // nodes (not in the tile's first column, whose x-1 neighbour is in the ring), north-wall or
        // south-wall nodes; the register-window march runs such tiles as face marches
        auto face_kind = [&](int t) -> uint32_t {
            if (lc.variant == 1 || !lc.face_tiles) return uint32_t(FACE_NONE);
            const uint32_t m = tile_codes[t];
            if (m == (1u << NC_WEST)) return uint32_t(FACE_WEST);
            if (m == (1u << NC_EAST) && (g.nx - 1) % lc.bx != 0) return uint32_t(FACE_EAST);
            if (m == (1u << NC_NORTH)) return uint32_t(FACE_NORTH);
            if (m == (1u << NC_SOUTH)) return uint32_t(FACE_SOUTH);
            return uint32_t(FACE_NONE);
        };
        for (size_t c = 0; c < frozen_cols_.size(); ++c)
            if (!frozen_cols_[c]) {
                const int x = int(c % size_t(g.nx)), y = int(c / size_t(g.nx));
                tile_live[x / lc.bx + tiles_x * (y / lc.by)] = 1;
            }
        // tiles that read a curved-wall node (in the tile or its one-node ring): in the two-stream
        // schedule they wait for the previous step's wall pass, the other (interior) tiles do not
        std::vector<uint8_t> tile_wall(ntiles, 0);
        for (int gi = 0; gi < 2; ++gi)
            for (int64_t nd : g.wall[gi].node) {
                const int64_t c = nd % (int64_t(g.nx) * g.ny);
                const int x = int(c % g.nx), y = int(c / g.nx);
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int xx = (x + dx + g.nx) % g.nx, yy = (y + dy + g.ny) % g.ny;
                        tile_wall[xx / lc.bx + tiles_x * (yy / lc.by)] = 1;
                    }
            }
        // y slabs: each XCD owns a band of tile rows over the full z range and runs it in rounds
        // of one z chunk, so that a chunk's halo planes are still in that XCD's L2 from the round
        // before (boundary tiles cut their part of a round into chunks of lz_special)
        const bool yslab = lc.xcd_remap && lc.slab_axis == 1 && tiles_y >= 8;
        const int groups = yslab ? 8 : ((lc.xcd_remap && g.nz >= 32) ? 8 : 1);
        const int slab = yslab ? g.nz : (g.nz + groups - 1) / groups;
        const int lz = std::max(1, std::min(slab, lc.lz));
        const int lzs = lc.lz_special > 0 ? std::max(1, std::min(yslab ? lz : slab, lc.lz_special)) : lz;
        // lists[0]: every item; lists[1]: interior tiles; lists[2]: wall-side tiles
        std::vector<std::vector<uint2>> lists[3];
        for (auto &l : lists) l.resize(groups);
        auto add = [&](int gi, int t, int z, int z1) {
            const bool sp = tile_special[t];
            const int part = (sp || tile_wall[t]) ? 2 : 1;
            const uint32_t field = uint32_t(t) | (sp ? kItemSpecial | (face_kind(t) << kItemFaceShift) : 0u);
            const uint2 it = make_uint2(field, uint32_t(z) | (uint32_t(z1) << 16));
            lists[0][gi].push_back(it);
            lists[part][gi].push_back(it);
        };
        // tiles marched in chunks of lz_special: those on the boundary path. A face march costs
        // about a plain march, and in a y slab its whole chunks keep step with the plain tiles'
        // (512x256x256 wake with a D=8 cylinder, FP32: 574 us with every boundary tile in half
        // chunks, 561 us in whole chunks)
        auto short_chunks = [&](int t) { return tile_special[t] && face_kind(t) == FACE_NONE; };
        // pass of a tile in its slab (or z round): boundary tiles in pass 0 (first), in pass 1 with
        // the others (tile order), or in pass 2 (last; in a y slab after every round)
        auto pass_of = [&](int t) { return tile_special[t] ? lc.special_order : 1; };
        // y slabs: the tiles, in row-major order, cut into one contiguous range per XCD of equal
        // estimated work, a tile running the boundary path counting 1.5 plain tiles (FP32 storage)
        // or 2 (16-bit storage and FP64), as measured per plane (MRLBM_SPECIAL_WEIGHT overrides;
        // 512x256x256 wake, FP32: 1.0 / 1.25 / 1.5 / 2.0 gave 586 / 584 / 582 / 591 us); rows
        // holding a curved wall would otherwise load their XCDs alone
        std::vector<int> yrange(groups + 1, 0);
        if (yslab) {
            const double ws = lc.special_weight > 0 ? lc.special_weight : (sizeof(R) == 4 && !Conv<S>::scaled) ? 1.5 : 2.0;
            std::vector<double> prefix(ntiles + 1, 0.0);
            for (int t = 0; t < ntiles; ++t)
                prefix[t + 1] = prefix[t] + (!tile_live[t] ? 0.0 : (tile_special[t] && face_kind(t) == FACE_NONE ? ws : 1.0));
            for (int k = 1; k < groups; ++k) {
                const double goal = prefix[ntiles] * k / groups;
                int t = yrange[k - 1];
                while (t < ntiles && prefix[t + 1] <= goal) ++t;
                yrange[k] = t;
            }
            yrange[groups] = ntiles;
        }
        for (int gi = 0; gi < groups; ++gi) {
            if (yslab) {
                const int ta = yrange[gi], tb = yrange[gi + 1];
                // a tile's part [z, ze) of a round, in chunks of lz (lz_special for tiles on the
                // boundary path)
                auto add_round = [&](int t, int z) {
                    const int ze = std::min(g.nz, z + lz), step = short_chunks(t) ? lzs : lz;
                    for (int zz = z; zz < ze; zz += step) add(gi, t, zz, std::min(ze, zz + step));
                };
                for (int z = 0; z < g.nz; z += lz)  // the rounds: passes 0 and 1
                    for (int pass = 0; pass < 2; ++pass)
                        for (int t = ta; t < tb; ++t)
                            if (pass_of(t) == pass && tile_live[t]) add_round(t, z);
                for (int z = 0; z < g.nz; z += lz)  // then pass 2, after every round
                    for (int t = ta; t < tb; ++t)
                        if (pass_of(t) == 2 && tile_live[t]) add_round(t, z);
                continue;
            }
            const int za = gi * slab, zb = std::min(g.nz, za + slab);
            for (int pass = 0; pass < 3; ++pass)
                for (int t = 0; t < ntiles; ++t) {
                    if (pass_of(t) != pass || !tile_live[t]) continue;
                    const int step = short_chunks(t) ? lzs : lz;
                    for (int z = za; z < zb; z += step) add(gi, t, z, std::min(zb, z + step));
                }
        }
        std::vector<uint2> items[3];
        for (int p = 0; p < 3; ++p) {
            size_t len = 0;
            for (auto &l : lists[p]) len = std::max(len, l.size());
            for (auto &l : lists[p]) {
                l.resize(len, make_uint2(0u, 0u));
                items[p].insert(items[p].end(), l.begin(), l.end());
            }
        }
        for (auto &d : d_items_)
            if (d) HIP_TRY(hipFree(d));
        for (auto &d : d_items_) d = nullptr;
        for (int p = 0; p < 3; ++p)
            if (int rc = upload(d_items_[p], items[p])) return rc;
        for (auto &e : graphs_)
            if (e) (void)hipGraphExecDestroy(e);
        for (auto &e : graphs_) e = nullptr;
        bx_ = lc.bx;
        by_ = lc.by;
        variant_ = lc.variant;
        tiles_x_ = tiles_x;
        tiles_y_ = tiles_y;
        ntiles_ = ntiles;
        lz_ = lz;
        lz_special_ = lzs;
        for (int p = 0; p < 3; ++p) nitems_[p] = int(items[p].size());
        xcd_ = groups;
        lists_ = std::move(lists[0]);
        launch_ = lc;
        // two streams per step (MRLBM_TWO_STREAM=1 only) where the fused pass reads no step
        // counter (no stochastic rounding), the case has curved walls and both launches have work.
        // Off by default: with grouped fields it no longer gains in FP64 (256x128x64 channel,
        // 12,444 / 12,428 MLUPS one stream against 12,334 / 12,475 two streams) and loses 18% in
        // FP32.
        const bool two = two_stream_env_ > 0;
        two_stream_ = two && !stochastic_ && nwall_ > 0 && nitems_[1] > 0 && nitems_[2] > 0 &&
                      fused_part_supported<R, S>(lc.bx, lc.by, lc.variant);
        if (two_stream_ && (prepare_fused<R, S>(lc.bx, lc.by, lc.variant, PART_INTERIOR) != 0 ||
                            prepare_fused<R, S>(lc.bx, lc.by, lc.variant, PART_WALL_SIDE) != 0))
            return set_error(MRLBM_ERR_UNSUPPORTED, "two-stream fused kernels cannot get their LDS");
        return MRLBM_OK;
    }

    // Wall-pressure statistics (MRLBM_STEP_STATS): FP64 sums per inner-wall node.
    int stats_reset() override {
        if (int rc = wait_batch()) return rc;
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipMemset(d_ps_sum_, 0, std::max<size_t>(g.wall[0].node.size(), 1) * sizeof(double)));
        stats_samples_ = 0;
        return MRLBM_OK;
    }

    int get_surface_pressure(double *mean, int64_t *samples) override {
        if (!stats_ok_) return set_error(MRLBM_ERR_UNSUPPORTED, "no inner wall with a pressure stencil inside the grid");
        if (int rc = wait_batch()) return rc;
        HIP_TRY(hipStreamSynchronize(stream_));
        const size_t nb = g.wall[0].node.size();
        HIP_TRY(hipMemcpy(mean, d_ps_sum_, nb * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nb; ++k) mean[k] = stats_samples_ ? mean[k] / double(stats_samples_) : 0.0;
        if (samples) *samples = stats_samples_;
        return MRLBM_OK;
    }

    bool stats_on(uint32_t flags) const { return (flags & MRLBM_STEP_STATS) && stats_ok_; }
    // inlet / outlet plane means: channel cases (which always run the wall pass that advances the
    // device step counter)
    bool planes_on(uint32_t flags) const { return (flags & MRLBM_STEP_STATS) && d.kind == MRLBM_CASE_CYLINDER; }

    PlaneArgs<R, S> plane_args(int cur) const {
        PlaneArgs<R, S> p{};
        p.mom = as_const(make_fields(buf_[cur ^ 1]));
        p.nx = g.nx;
        p.ny = g.ny;
        p.nz = g.nz;
        p.rho0 = d.rho0;
        p.scale = scale_;
        p.step_counter = d_step_counter_ + cur;  // the index of the step the planes belong to
        p.ring = d_plane_ring_;
        return p;
    }

    int get_plane_means(int64_t first, int count, double *out2) override {
        if (d.kind != MRLBM_CASE_CYLINDER) return set_error(MRLBM_ERR_UNSUPPORTED, "plane means need an inlet and an outlet");
        if (int rc = wait_batch()) return rc;
        if (count < 0 || first < 0 || first + count > steps || first < steps - FORCE_RING_STEPS)
            return set_error(MRLBM_ERR_INVALID, "plane record range not available");
        HIP_TRY(hipStreamSynchronize(stream_));
        for (int k = 0; k < count;) {
            const int slot = int((first + k) % FORCE_RING_STEPS);
            const int len = std::min(count - k, FORCE_RING_STEPS - slot);
            HIP_TRY(hipMemcpy(out2 + 2 * size_t(k), d_plane_ring_ + 2 * size_t(slot), size_t(len) * 2 * sizeof(double),
                              hipMemcpyDeviceToHost));
            k += len;
        }
        return MRLBM_OK;
    }

    // Arguments of the wall-pressure pass on the state after a step from parity `cur`.
    PressureArgs<R, S> pressure_args(int cur) const {
        PressureArgs<R, S> p{};
        p.mom = as_const(make_fields(buf_[cur ^ 1]));
        p.node = d_wnode_;
        p.wnx = d_wnx_;
        p.wny = d_wny_;
        p.n = int(g.wall[0].node.size());
        p.nx = g.nx;
        p.ny = g.ny;
        p.xc = g.wall[0].xc;
        p.yc = g.wall[0].yc;
        p.radius = g.wall[0].radius;
        p.rho0 = d.rho0;
        p.scale = scale_;
        p.sum = d_ps_sum_;
        return p;
    }

    const LaunchConfig &launch_config() const { return launch_; }
    LaunchConfig current_launch() const override { return launch_; }

    int vti_points(int layout, float vel_norm, float *host) override {
        if (int rc = wait_batch()) return rc;
        const size_t n = size_t(g.n), words = (n + 31) / 32;
        std::vector<uint32_t> bits(words, 0u);
        for (size_t i = 0; i < n; ++i)
            if (g.ref_type[i] == RT_SOLID) bits[i >> 5] |= 1u << (i & 31);
        float *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, 4 * n * sizeof(float) + words * sizeof(uint32_t)));
        const int rc = [&]() -> int {
            VtiArgs a{};
            a.n = uint32_t(n);
            a.nx = g.nx;
            a.ny = g.ny;
            a.layout = layout;
            a.xc = d.xc;
            a.yc = d.yc;
            a.vel_norm = vel_norm;
            a.solid = reinterpret_cast<const uint32_t *>(tmp + 4 * n);
            a.out = tmp;
            HIP_TRY(hipStreamSynchronize(stream_));
            HIP_TRY(hipMemcpy(tmp + 4 * n, bits.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice));
            launch_vti_points<R, S>(as_const(make_fields(buf_[cur_])), scale_, R(d.rho0), a, stream_);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(stream_));
            HIP_TRY(hipMemcpy(host, tmp, 4 * n * sizeof(float), hipMemcpyDeviceToHost));
            return MRLBM_OK;
        }();
        (void)hipFree(tmp);
        return rc;
    }
    // per-XCD work-item lists (each padded to the same length) of the current configuration
    const std::vector<std::vector<uint2>> &item_lists() const { return lists_; }
    int nwall() const { return nwall_; }
    int nbcf() const { return nbcf_; }
    int parity() const { return cur_; }
    hipStream_t stream() const { return stream_; }
    // the buffer holding the current state and its storage scales
    FieldSet<const S> state() const { return as_const(make_fields(buf_[cur_])); }
    const StoreScale<R> &store_scale() const { return scale_; }

    // Kernel arguments of the fused pass from buffer parity `cur`, over every work item or the
    // interior / wall-side items of the two-stream schedule.
    FusedArgs<R, S> fused_args(int cur, int part = PART_ALL) const {
        FusedArgs<R, S> a{};
        a.src = as_const(make_fields(buf_[cur]));
        a.dst = make_fields(buf_[cur ^ 1]);
        a.donor = donor_;
        a.pad = uint32_t(g.n);
        a.nx = g.nx;
        a.ny = g.ny;
        a.nz = g.nz;
        a.plane_aligned = (int64_t(g.nx) * g.ny) % int64_t(kChunk<S>) == 0 ? 1 : 0;
        a.tiles_x = tiles_x_;
        a.nblocks = nitems_[part];
        a.xcd_groups = xcd_;
        a.items = d_items_[part];
        a.info = d_info_;
        a.wall_nx = d_wnx_;
        a.wall_ny = d_wny_;
        a.wall_in = d_win_;
        a.bcf_in = d_bcfin_;
        a.inlet = d_inlet_;
        a.wall_count = std::max(nwall_, 1);
        a.inlet_count = std::max<int>(int(g.inlet.size()), 1);
        a.step_counter = stochastic_ ? d_step_counter_ + cur : nullptr;
        a.rec = part == PART_INTERIOR ? RecordArgs{} : record_args();
        a.scale = scale_;
        a.omega = omega_;
        a.bcf_strong = (d.compat_flags & MRLBM_COMPAT_BCF_MASS_STRONG) ? 1 : 0;
        a.rho0 = rho0_;
        a.nonfinite = d_nonfinite_;
        return a;
    }

    // Kernel arguments of the wall pass on the buffer the fused pass from parity `cur` wrote.
    WallArgs<R, S> wall_args(int cur, bool forces) const {
        WallArgs<R, S> w{};
        w.mom = make_fields(buf_[cur ^ 1]);
        w.donor = donor_;
        w.nwall = nwall_;
        w.nb_inner = nb_inner_;
        w.nbcf = nbcf_;
        w.node = d_wnode_;
        w.wnx = d_wnx_;
        w.wny = d_wny_;
        w.coef = d_wcoef_;
        w.wconst = d_wconst_;
        w.in_mask = d_win_;
        w.out_mask = d_wout_;
        w.st_base = d_stbase_;
        w.st_frac = d_stfrac_;
        w.nx = g.nx;
        w.col0 = wall_cols_[0];
        w.col1 = wall_cols_[1];
        w.plane = uint32_t(int64_t(g.nx) * g.ny);
        w.bcf_node = d_bcfnode_;
        w.bcf_tag = d_bcftag_;
        w.bcf_in = d_bcfin_;
        w.bcf_out = d_bcfout_;
        for (int gi = 0; gi < 2; ++gi)
            for (int k = 0; k < 3; ++k) w.uw[gi][k] = R(g.wall[gi].uw[k]);
        w.scale = scale_;
        w.omega = omega_;
        w.strong = (d.compat_flags & MRLBM_COMPAT_MASS_STRONG) ? 1 : 0;
        w.rho0 = rho0_;
        w.forces = forces ? 1 : 0;
        w.nblocks = wall_blocks(nwall_, forces ? nbcf_ : 0);
        w.partials = d_partials_;
        w.step_counter = d_step_counter_ + cur;
        w.next_counter = d_step_counter_ + (cur ^ 1);
        w.pending = d_pending_;
        return w;
    }

    // The wall pass also advances the device step counter, which seeds the stochastic rounding of
    // the 16-bit forms, so those run it even without walls.
    bool has_wall_launch(uint32_t flags) const {
        return nwall_ > 0 || ((flags & MRLBM_STEP_FORCES) && nbcf_ > 0) || Conv<S>::scaled;
    }
    bool forces_on(uint32_t flags) const { return (flags & MRLBM_STEP_FORCES) && g.ngroups > 0; }

    // Force records of the case: the last wall pass of a step call leaves its record pending (the
    // next fused launch would file it), so every step call ends with a record launch.
    RecordArgs record_args() const {
        RecordArgs r{};
        r.pending = forces_on(MRLBM_STEP_FORCES) ? d_pending_ : nullptr;
        r.partials = d_partials_;
        r.ring = d_ring_;
        r.nblocks = wall_blocks(nwall_, nbcf_);
        return r;
    }
    void queue_record(uint32_t flags) {
        if (forces_on(flags)) launch_record(record_args(), stream_);
    }

    // Buffers swap after every step (driven by a batch that launches the kernels itself).
    void advance(uint32_t flags) {
        cur_ ^= 1;
        ++steps;
        if (stats_on(flags)) ++stats_samples_;
    }

    double velocity_reference() const override { return double(scale_.uref); }

    int time_memory_twin(int launches, double *ms) override {
        if (launches <= 0 || !ms) return set_error(MRLBM_ERR_INVALID, "launch count must be positive");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        FusedArgs<R, S> a = fused_args(cur_);
        launch_memory_twin<R, S>(a, bx_, by_, stream_);
        HIP_TRY(hipEventRecord(e0, stream_));
        for (int k = 0; k < launches; ++k) launch_memory_twin<R, S>(a, bx_, by_, stream_);
        HIP_TRY(hipEventRecord(e1, stream_));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipGetLastError());
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = double(t) / launches;
        return MRLBM_OK;
    }

    int apply_geometry() override {
        HIP_TRY(hipStreamSynchronize(stream_));
        if (int rc = build_tables()) return rc;
        if (int rc = configure_launch(launch_)) return rc;
        return init();
    }
    bool nonfinite() const override { return h_nonfinite_ && __atomic_load_n(h_nonfinite_, __ATOMIC_ACQUIRE) != 0u; }
    void clear_nonfinite() {
        if (h_nonfinite_) __atomic_store_n(h_nonfinite_, 0u, __ATOMIC_RELEASE);
    }

    // Seeded initial state: u_a = a (2 r - 1), r = hash(seed, a N + i); rho = rho0; m = u u
    // (m = 0 with MRLBM_COMPAT_INIT_M_ZERO, as the reference's K9 produces).
    int init() override {
        const int64_t n = g.n;
        std::vector<double> h(size_t(10) * n);
        const bool m0 = d.compat_flags & MRLBM_COMPAT_INIT_M_ZERO;
        for (int64_t i = 0; i < n; ++i) {
            double u[3];
            for (int a = 0; a < 3; ++a) u[a] = d.init_noise * (2.0 * hash_uniform(d.seed, uint64_t(a) * uint64_t(n) + uint64_t(i)) - 1.0);
            const double v[10] = {0.0, u[0], u[1], u[2],
                                  m0 ? 0.0 : u[0] * u[0], m0 ? 0.0 : u[1] * u[1], m0 ? 0.0 : u[2] * u[2],
                                  m0 ? 0.0 : u[0] * u[1], m0 ? 0.0 : u[0] * u[2], m0 ? 0.0 : u[1] * u[2]};
            for (int f = 0; f < 10; ++f) h[size_t(f) * n + i] = v[f];
        }
        int rc = put_all(h);
        if (rc) return rc;
        return set_steps(0);
    }

    // Steps are queued eagerly for short runs; longer runs replay a captured graph of
    // graph_steps_ steps (an even number, so the ping-pong parity is unchanged). A simulation
    // whose watchdog has tripped queues nothing.
    int step(int nsteps, uint32_t flags) override {
        if (int rc = check_finite()) return rc;
        int done = 0;
        if (graph_steps_ > 0) {
            while (nsteps - done >= graph_steps_) {
                hipGraphExec_t exec = nullptr;
                int rc = graph_for(flags, exec);
                if (rc) return rc;
                HIP_TRY(hipGraphLaunch(exec, stream_));
                steps += graph_steps_;
                if (stats_on(flags)) stats_samples_ += graph_steps_;
                done += graph_steps_;
            }
        }
        open_streams();
        for (; done < nsteps; ++done) queue_step(flags, nullptr);
        join_streams();
        queue_record(flags);
        HIP_TRY(hipGetLastError());
        return MRLBM_OK;
    }

    int set_steps(int64_t n) override {
        steps = n;
        const unsigned long long v[2] = {static_cast<unsigned long long>(n), static_cast<unsigned long long>(n)};
        HIP_TRY(hipMemcpy(d_step_counter_, v, sizeof(v), hipMemcpyHostToDevice));
        return MRLBM_OK;
    }

    int sync() override {
        if (int rc = wait_batch()) return rc;
        HIP_TRY(hipStreamSynchronize(stream_));
        return check_finite();
    }

    int get(int f, double *h) override {
        std::vector<double> all;
        int rc = get_all(all);
        if (rc) return rc;
        std::memcpy(h, all.data() + size_t(f) * g.n, size_t(g.n) * sizeof(double));
        return MRLBM_OK;
    }

    // The raw payload is field-major (10 x n values as stored); the device buffers are
    // node-chunked, so the copy goes through a host image of the buffer.
    int get_raw(void *h) override {
        if (int rc = wait_batch()) return rc;
        HIP_TRY(hipStreamSynchronize(stream_));
        std::vector<S> img(elems_);
        HIP_TRY(hipMemcpy(img.data(), buf_[cur_], elems_ * sizeof(S), hipMemcpyDeviceToHost));
        S *out = static_cast<S *>(h);
        for (int f = 0; f < 10; ++f)
            for (int64_t i = 0; i < g.n; ++i)
                out[size_t(f) * g.n + i] = img[field_elem<S>(uint32_t(i), f)];
        return MRLBM_OK;
    }

    int set_raw(const void *h) override {
        HIP_TRY(hipStreamSynchronize(stream_));
        std::vector<S> img(elems_);
        HIP_TRY(hipMemcpy(img.data(), buf_[cur_], elems_ * sizeof(S), hipMemcpyDeviceToHost));  // keeps the padding
        const S *in = static_cast<const S *>(h);
        for (int f = 0; f < 10; ++f)
            for (int64_t i = 0; i < g.n; ++i)
                img[field_elem<S>(uint32_t(i), f)] = in[size_t(f) * g.n + i];
        for (int b = 0; b < 2; ++b)  // both buffers (frozen nodes)
            HIP_TRY(hipMemcpy(buf_[b], img.data(), elems_ * sizeof(S), hipMemcpyHostToDevice));
        clear_nonfinite();  // a new state
        return MRLBM_OK;
    }

    int set(int f, const double *h) override {
        std::vector<double> all;
        int rc = get_all(all);
        if (rc) return rc;
        std::memcpy(all.data() + size_t(f) * g.n, h, size_t(g.n) * sizeof(double));
        return put_all(all);
    }

    int get_forces(int64_t first, int count, double *out4) override {
        if (int rc = wait_batch()) return rc;
        if (count < 0 || first < 0 || first + count > steps || first < steps - FORCE_RING_STEPS)
            return set_error(MRLBM_ERR_INVALID, "force record range not available");
        HIP_TRY(hipStreamSynchronize(stream_));
        // at most two contiguous pieces of the ring
        for (int k = 0; k < count;) {
            const int slot = int((first + k) % FORCE_RING_STEPS);
            const int len = std::min(count - k, FORCE_RING_STEPS - slot);
            HIP_TRY(hipMemcpy(out4 + 4 * size_t(k), d_ring_ + 4 * size_t(slot), size_t(len) * 4 * sizeof(double),
                              hipMemcpyDeviceToHost));
            k += len;
        }
        // the records are written either way; the status tells whether the state blew up
        return check_finite();
    }

    int time_steps(int n, uint32_t flags, double *ms_total, double *ms_k) override {
        if (n <= 0) return set_error(MRLBM_ERR_INVALID, "nsteps must be positive");
        std::vector<hipEvent_t> ev(size_t(4) * n + 2);
        for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
        if (!ms_k && graph_steps_ > 0 && n >= graph_steps_) {
            hipGraphExec_t exec = nullptr;
            if (int rc = graph_for(flags, exec)) return rc;  // captured outside the timed region
        }
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipEventRecord(ev[0], stream_));
        if (ms_k) {
            for (int s = 0; s < n; ++s) queue_step(flags, &ev[2 + 4 * size_t(s)]);
            queue_record(flags);
        } else {
            int rc = step(n, flags);
            if (rc) return rc;
        }
        HIP_TRY(hipEventRecord(ev[1], stream_));
        HIP_TRY(hipEventSynchronize(ev[1]));
        HIP_TRY(hipGetLastError());
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        *ms_total = ms;
        if (ms_k) {
            double acc[2] = {0.0, 0.0};
            int cnt[2] = {0, 0};
            for (int s = 0; s < n; ++s) {
                for (int k = 0; k < 2; ++k) {
                    hipEvent_t a = ev[2 + 4 * size_t(s) + 2 * k], b = ev[3 + 4 * size_t(s) + 2 * k];
                    float t = 0.f;
                    if (k == 1 && !has_wall_launch(flags)) continue;
                    HIP_TRY(hipEventElapsedTime(&t, a, b));
                    acc[k] += t;
                    ++cnt[k];
                }
            }
            ms_k[0] = cnt[0] ? acc[0] / cnt[0] : 0.0;
            ms_k[1] = cnt[1] ? acc[1] / cnt[1] : 0.0;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        return MRLBM_OK;
    }

    int time_replays(int n, uint32_t flags, double *ms_total, double *ms_step, int *replays) override {
        if (n <= 0) return set_error(MRLBM_ERR_INVALID, "nsteps must be positive");
        if (int rc = check_finite()) return rc;
        const int nrep = graph_steps_ > 0 ? n / graph_steps_ : 0;
        hipGraphExec_t exec = nullptr;
        if (nrep > 0)
            if (int rc = graph_for(flags, exec)) return rc;  // an even replay leaves the parity as it is
        std::vector<hipEvent_t> ev(size_t(nrep) + 2, nullptr);
        auto done = [&](int rc) {
            for (auto &e : ev)
                if (e) (void)hipEventDestroy(e);
            return rc;
        };
        for (auto &e : ev)
            if (hipEventCreate(&e) != hipSuccess) return done(set_error(MRLBM_ERR_DEVICE, "hipEventCreate failed"));
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipEventRecord(ev[0], stream_));
        for (int r = 0; r < nrep; ++r) {
            HIP_TRY(hipGraphLaunch(exec, stream_));
            steps += graph_steps_;
            if (stats_on(flags)) stats_samples_ += graph_steps_;
            HIP_TRY(hipEventRecord(ev[size_t(r) + 1], stream_));
        }
        open_streams();
        for (int s = nrep * graph_steps_; s < n; ++s) queue_step(flags, nullptr);
        join_streams();
        queue_record(flags);
        HIP_TRY(hipEventRecord(ev[size_t(nrep) + 1], stream_));
        HIP_TRY(hipEventSynchronize(ev[size_t(nrep) + 1]));
        HIP_TRY(hipGetLastError());
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[size_t(nrep) + 1]));
        *ms_total = ms;
        std::vector<double> per;
        for (int r = 0; r < nrep; ++r) {
            float t = 0.f;
            HIP_TRY(hipEventElapsedTime(&t, ev[size_t(r)], ev[size_t(r) + 1]));
            per.push_back(double(t) / graph_steps_);
        }
        std::sort(per.begin(), per.end());
        if (ms_step) *ms_step = per.empty() ? 0.0 : (per.size() % 2 ? per[per.size() / 2]
                                                                   : 0.5 * (per[per.size() / 2 - 1] + per[per.size() / 2]));
        if (replays) *replays = nrep;
        return done(check_finite());
    }

  private:

    // Executable graph of graph_steps_ steps for these flags from the current buffer parity.
    int graph_for(uint32_t flags, hipGraphExec_t &exec) {
        const int key = int(flags & (MRLBM_STEP_FORCES | MRLBM_STEP_STATS)) * 2 + cur_;
        if (!graphs_[key]) {
            const int cur0 = cur_;
            const int64_t steps0 = steps, samples0 = stats_samples_;
            HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            open_streams();
            for (int s = 0; s < graph_steps_; ++s) queue_step(flags, nullptr);
            join_streams();
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamEndCapture(stream_, &graph));
            cur_ = cur0;
            steps = steps0;
            stats_samples_ = samples0;
            hipGraphExec_t e = nullptr;
            const hipError_t err = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_TRY(err);
            graphs_[key] = e;
        }
        exec = graphs_[key];
        return MRLBM_OK;
    }

    // Queue one time step: fused pass buf[cur] -> buf[cur^1], then the wall pass on buf[cur^1].
    // Steps timed kernel by kernel (events) keep the one-stream order.
    void queue_step(uint32_t flags, hipEvent_t *ev) {
        if (two_stream_ && !ev) {
            queue_step_two_streams(flags);
            return;
        }
        if (ev) (void)hipEventRecord(ev[0], stream_);
        launch_fused<R, S>(fused_args(cur_), bx_, by_, variant_, PART_ALL, stream_);
        if (ev) (void)hipEventRecord(ev[1], stream_);
        if (has_wall_launch(flags)) {
            if (ev) (void)hipEventRecord(ev[2], stream_);
            launch_wall<R, S>(wall_args(cur_, forces_on(flags)), stream_);
            if (ev) (void)hipEventRecord(ev[3], stream_);
        } else if (ev) {
            (void)hipEventRecord(ev[2], stream_);
            (void)hipEventRecord(ev[3], stream_);
        }
        if (stats_on(flags)) {
            launch_pressure<R, S>(pressure_args(cur_), stream_);
            ++stats_samples_;
        }
        if (planes_on(flags)) launch_planes<R, S>(plane_args(cur_), stream_);
        cur_ ^= 1;
        ++steps;
    }

    // One step of the two-stream schedule. The wall-side tiles (boundary tiles and tiles whose
    // ring holds a curved-wall node) read what the previous step's wall pass wrote; the interior
    // tiles read only nodes of the fused pass, so they run on stream2_ while that wall pass runs.
    // Each launch waits for the other one of the previous step, which read the buffer it writes
    // and wrote nodes of its ring. The statistics read interior nodes and wait for both.
    void queue_step_two_streams(uint32_t flags) {
        (void)hipStreamWaitEvent(stream2_, ev_wall_side_, 0);
        (void)hipStreamWaitEvent(stream_, ev_interior_, 0);
        launch_fused<R, S>(fused_args(cur_, PART_WALL_SIDE), bx_, by_, variant_, PART_WALL_SIDE, stream_);
        (void)hipEventRecord(ev_wall_side_, stream_);
        launch_fused<R, S>(fused_args(cur_, PART_INTERIOR), bx_, by_, variant_, PART_INTERIOR, stream2_);
        (void)hipEventRecord(ev_interior_, stream2_);
        if (has_wall_launch(flags)) launch_wall<R, S>(wall_args(cur_, forces_on(flags)), stream_);
        if (stats_on(flags) || planes_on(flags)) {
            (void)hipStreamWaitEvent(stream_, ev_interior_, 0);
            if (stats_on(flags)) {
                launch_pressure<R, S>(pressure_args(cur_), stream_);
                ++stats_samples_;
            }
            if (planes_on(flags)) launch_planes<R, S>(plane_args(cur_), stream_);
        }
        cur_ ^= 1;
        ++steps;
    }

    // A run of two-stream steps starts after everything queued on stream_ and ends with stream_
    // waiting for its last interior launch, so stream_ alone orders what follows.
    void open_streams() {
        if (!two_stream_) return;
        (void)hipEventRecord(ev_wall_side_, stream_);
        (void)hipEventRecord(ev_interior_, stream_);
    }
    void join_streams() {
        if (two_stream_) (void)hipStreamWaitEvent(stream_, ev_interior_, 0);
    }

    template <class T>
    int upload(T *&dptr, const std::vector<T> &h) {
        const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
        HIP_TRY(hipMalloc(&dptr, bytes));
        if (!h.empty()) HIP_TRY(hipMemcpy(dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return MRLBM_OK;
    }

    int set_all(const std::vector<double> &all10) override {
        if (all10.size() != size_t(10) * size_t(g.n)) return set_error(MRLBM_ERR_INVALID, "field data size");
        if (int rc = wait_batch()) return rc;
        HIP_TRY(hipStreamSynchronize(stream_));
        return put_all(all10);  // clears the watchdog
    }

    int put_all(const std::vector<double> &h) {
        double *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, h.size() * sizeof(double)));
        HIP_TRY(hipMemcpy(tmp, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        // both buffers: frozen nodes are never rewritten and must read the same from either
        launch_encode<R, S>(tmp, make_fields(buf_[cur_]), uint32_t(g.n), scale_, stream_);
        launch_encode<R, S>(tmp, make_fields(buf_[cur_ ^ 1]), uint32_t(g.n), scale_, stream_);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipFree(tmp));
        clear_nonfinite();  // a new state
        return MRLBM_OK;
    }

    int get_all(std::vector<double> &h) {
        if (int rc = wait_batch()) return rc;
        h.resize(size_t(10) * g.n);
        double *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, h.size() * sizeof(double)));
        launch_decode<R, S>(as_const(make_fields(buf_[cur_])), tmp, uint32_t(g.n), scale_, stream_);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipMemcpy(h.data(), tmp, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipFree(tmp));
        return MRLBM_OK;
    }

    void release() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (stream2_) (void)hipStreamSynchronize(stream2_);
        for (auto &g : graphs_)
            if (g) (void)hipGraphExecDestroy(g);
        void *ptrs[] = {buf_[0], buf_[1], donor_, d_info_, d_items_[0], d_items_[1], d_items_[2], d_wconst_,
                        d_wnx_, d_wny_, d_win_, d_wout_, d_wnode_,
                        d_wcoef_, d_stbase_, d_stfrac_, d_bcfnode_, d_bcftag_, d_bcfin_, d_bcfout_, d_inlet_,
                        d_partials_, d_ring_, d_step_counter_, d_pending_, d_ps_sum_, d_plane_ring_};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (h_nonfinite_) (void)hipHostFree(h_nonfinite_);
        h_nonfinite_ = d_nonfinite_ = nullptr;
        if (ev_wall_side_) (void)hipEventDestroy(ev_wall_side_);
        if (ev_interior_) (void)hipEventDestroy(ev_interior_);
        if (stream2_) (void)hipStreamDestroy(stream2_);
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = stream2_ = nullptr;
    }

    hipStream_t stream_ = nullptr;
    // two-stream schedule: the interior tiles of step s run on stream2_ after the wall-side tiles
    // of step s-1 (ev_wall_side_), concurrently with its wall pass; the wall-side tiles of step s
    // run on stream_ after the interior tiles of step s-1 (ev_interior_) and that wall pass
    hipStream_t stream2_ = nullptr;
    hipEvent_t ev_wall_side_ = nullptr, ev_interior_ = nullptr;
    bool two_stream_ = false;
    int two_stream_env_ = -1;  // MRLBM_TWO_STREAM: 1 on, 0 off, unset (-1) the precision's default
    S *buf_[2] = {nullptr, nullptr};
    S *donor_ = nullptr;
    size_t stride_ = 0;  // donor snapshot: node slots (n + 1, rounded up)
    size_t elems_ = 0;   // elements of one node-chunked moment buffer
    int cur_ = 0;
    R omega_ = 0, rho0_ = 1;
    StoreScale<R> scale_{};
    int variant_ = 1;
    bool stochastic_ = Conv<S>::scaled;
    int graph_steps_ = 16;
    hipGraphExec_t graphs_[8] = {};
    uint32_t *h_nonfinite_ = nullptr;  // watchdog flag (mapped host memory) and its device address
    uint32_t *d_nonfinite_ = nullptr;
    int bx_ = 32, by_ = 8, lz_ = 16, lz_special_ = 16, tiles_x_ = 1, tiles_y_ = 1, ntiles_ = 1, xcd_ = 1;
    int nitems_[3] = {1, 0, 0};  // work items (blocks) of the three item lists
    std::vector<uint32_t> special_cols_;  // x + nx y of columns holding non-bulk nodes
    std::vector<uint32_t> special_col_codes_;  // their node-code sets (bit 31: a donor node)
    std::vector<uint8_t> frozen_cols_;    // per column: frozen at every z
    std::vector<std::vector<uint2>> lists_;
    LaunchConfig launch_;
    int nwall_ = 0, nb_inner_ = 0, nbcf_ = 0;
    int wall_cols_[2] = {0, 0};  // wall records per plane of each group (0: a record per wall node)
    uint32_t *d_info_ = nullptr;
    uint2 *d_items_[3] = {nullptr, nullptr, nullptr};  // every item, interior items, wall-side items
    R *d_wconst_ = nullptr;
    R *d_wnx_ = nullptr, *d_wny_ = nullptr, *d_wcoef_ = nullptr, *d_stfrac_ = nullptr, *d_inlet_ = nullptr;
    uint32_t *d_win_ = nullptr, *d_wout_ = nullptr, *d_wnode_ = nullptr, *d_stbase_ = nullptr;
    uint32_t *d_bcfnode_ = nullptr, *d_bcfin_ = nullptr, *d_bcfout_ = nullptr;
    uint8_t *d_bcftag_ = nullptr;
    double *d_partials_ = nullptr, *d_ring_ = nullptr, *d_ps_sum_ = nullptr, *d_plane_ring_ = nullptr;
    bool stats_ok_ = false;
    int64_t stats_samples_ = 0;
    unsigned long long *d_step_counter_ = nullptr;  // step index, one counter per buffer parity
    unsigned long long *d_pending_ = nullptr;       // step + 1 of a force record not yet filed
};

}  // namespace mrlbm

namespace mrlbm {

// A simulation of the descriptor's precision with its geometry built; with `device` also its
// device buffers and tables.
// output.hip: file formats of the visualisation output and the run summary
std::string vti_file_name(int64_t timestep);
int write_vti_file(const std::string &path, int nx, int ny, int nz, const char *const names[4], const float *points);
int write_pvd_file(const std::string &path, int32_t max_iter, int32_t save_every);
std::string sim_info_string(const SimBase &s);

inline int make_sim(const mrlbm_case_desc &desc, bool device, std::unique_ptr<SimBase> &out) {
    std::unique_ptr<SimBase> impl;
    switch (desc.precision) {
    case MRLBM_FP64: impl.reset(new Sim<double, double>); break;
    case MRLBM_FP32: impl.reset(new Sim<float, float>); break;
    case MRLBM_FP32_BF16: impl.reset(new Sim<float, __hip_bfloat16>); break;
    case MRLBM_FP32_FP16: impl.reset(new Sim<float, _Float16>); break;
    default: return set_error(MRLBM_ERR_INVALID, "unknown precision");
    }
    impl->d = desc;
    std::string err;
    int rc = build_geometry(desc, impl->g, err);
    if (rc) return set_error(rc, err);
    if (device) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            return set_error(MRLBM_ERR_DEVICE, "no HIP device available");
        rc = impl->setup();
        if (rc) return rc;
        impl->device_ready = true;
    }
    out = std::move(impl);
    return MRLBM_OK;
}

// Device bytes a case of this descriptor allocates: two moment buffers and the donor snapshot
// (10 fields each, padded), node codes and the force ring; wall tables (well below 1%) are left
// out.
inline uint64_t case_device_bytes(const mrlbm_case_desc &d) {
    const uint64_t n = uint64_t(d.nx) * uint64_t(d.ny) * uint64_t(d.nz);
    const uint64_t stride = (n + 1 + 63) / 64 * 64;
    const uint64_t s = d.precision == MRLBM_FP64 ? 8 : (d.precision == MRLBM_FP32 ? 4 : 2);
    return 30 * stride * s + 4 * n + uint64_t(FORCE_RING_STEPS) * 6 * sizeof(double);
}

}  // namespace mrlbm

// C handle of one simulation. A handle that belongs to a batch (mrlbm_batch_case) is stepped by
// the batch; its own step, init, checkpoint_read and destroy calls are refused.
struct mrlbm_sim {
    std::unique_ptr<mrlbm::SimBase> impl;
    bool in_batch = false;
};
