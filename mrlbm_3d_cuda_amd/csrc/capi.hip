// capi.hip - simulation object and the extern "C" entry points of include/mrlbm.h.
//
// Host side of the reference's driver (main.cu:7-40, simulation.cu:13-172, solve_mrlbm.cuh):
// owns the device buffers, uploads the case tables and queues the step kernels on one HIP
// stream. Errors are returned as status codes with a thread-local message instead of exit().
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

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_error(MRLBM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

using mrlbm::FORCE_RING_STEPS;

}  // namespace

namespace mrlbm {

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
    mrlbm_case_desc d{};
    Geometry g;
    int64_t steps = 0;
    int bytes_per_value = 0;
    bool device_ready = false;
};

template <class R, class S>
class Sim final : public SimBase {
  public:
    ~Sim() override { release(); }

    int setup() override {
        bytes_per_value = int(sizeof(S));
        HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        const int64_t n = g.n;
        stride_ = size_t((n + 63) / 64 * 64);
        HIP_TRY(hipMalloc(&buf_[0], 10 * stride_ * sizeof(S)));
        HIP_TRY(hipMalloc(&buf_[1], 10 * stride_ * sizeof(S)));
        HIP_TRY(hipMalloc(&donor_, 10 * stride_ * sizeof(S)));
        HIP_TRY(hipMemsetAsync(donor_, 0, 10 * stride_ * sizeof(S), stream_));
        omega_ = R(1.0 / d.tau);
        rho0_ = R(d.rho0);
        // 16-bit storage scales: deviations of O(1e-2) mapped to O(1)
        scale_.enc[0] = R(64);
        scale_.enc[1] = R(16);
        scale_.enc[2] = R(256);
        for (int k = 0; k < 3; ++k) scale_.dec[k] = R(1) / scale_.enc[k];

        // ---- tiling of the fused pass
        if (sizeof(R) == 8) {
            bx_ = 16;
            by_ = 8;
        } else {
            bx_ = 32;
            by_ = 8;
        }
        if (const char *e = std::getenv("MRLBM_TILE")) {
            int tx = 0, ty = 0;
            if (std::sscanf(e, "%dx%d", &tx, &ty) != 2 || !fused_tile_supported(tx, ty))
                return set_error(MRLBM_ERR_INVALID, std::string("unsupported MRLBM_TILE ") + e);
            bx_ = tx;
            by_ = ty;
        }
        int dev = 0, max_lds = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (const char *e = std::getenv("MRLBM_FUSED_VARIANT")) variant_ = std::atoi(e) == 1 ? 1 : 0;
        if (const char *e = std::getenv("MRLBM_STOCHASTIC_ROUNDING")) stochastic_ = Conv<S>::scaled && std::atoi(e) != 0;
        if (const char *e = std::getenv("MRLBM_GRAPH_STEPS")) graph_steps_ = std::max(0, std::atoi(e) / 2 * 2);
        if (fused_lds_bytes(bx_, by_, sizeof(R), variant_) > size_t(max_lds)) {
            if (std::getenv("MRLBM_TILE"))
                return set_error(MRLBM_ERR_UNSUPPORTED, "fused tile needs more LDS than a block may use");
            // default shape too large for this device: halve the tile height
            bx_ = sizeof(R) == 8 ? 16 : 32;
            by_ = 4;
            if (fused_lds_bytes(bx_, by_, sizeof(R), variant_) > size_t(max_lds))
                return set_error(MRLBM_ERR_UNSUPPORTED, "no fused tile fits the device LDS");
        }
        tiles_x_ = (g.nx + bx_ - 1) / bx_;
        tiles_y_ = (g.ny + by_ - 1) / by_;
        ntiles_ = tiles_x_ * tiles_y_;

        // ---- stencils of the wall nodes and donor marking
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
                const R xw = R(d.xc) + R(w.radius) * nx, yw = R(d.yc) + R(w.radius) * ny;
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

        // ---- work items of the fused pass. The z axis is cut into one slab per XCD (blocks are
        // dispatched round-robin over the XCDs, so slab j runs on XCD j and its tiles share that
        // L2); inside a slab every tile is cut into z chunks, shorter for tiles that hold
        // boundary nodes (their blocks take the slow per-node path), and those chunks go first.
        std::vector<uint8_t> tile_special(ntiles_, 0);
        for (int64_t i = 0; i < g.n; ++i)
            if (info[i] != NC_BULK) {
                const int x = int(i % g.nx), y = int((i / g.nx) % g.ny);
                tile_special[x / bx_ + tiles_x_ * (y / by_)] = 1;
            }
        int groups = (g.nz >= 32) ? 8 : 1;
        if (const char *e = std::getenv("MRLBM_XCD_REMAP")) groups = std::atoi(e) ? groups : 1;
        const int slab = (g.nz + groups - 1) / groups;
        lz_ = std::min(slab, 32);
        if (const char *e = std::getenv("MRLBM_LZ")) lz_ = std::max(1, std::min(slab, std::atoi(e)));
        int lz_special = std::max(2, lz_ / 4);
        if (const char *e = std::getenv("MRLBM_LZ_SPECIAL")) lz_special = std::max(1, std::min(slab, std::atoi(e)));
        std::vector<std::vector<uint2>> lists(groups);
        for (int gi = 0; gi < groups; ++gi) {
            const int za = gi * slab, zb = std::min(g.nz, za + slab);
            for (int pass = 0; pass < 2; ++pass)
                for (int t = 0; t < ntiles_; ++t) {
                    const bool sp = tile_special[t];
                    if (sp != (pass == 0)) continue;
                    const int step = sp ? lz_special : lz_;
                    for (int z = za; z < zb; z += step) {
                        const int z1 = std::min(zb, z + step);
                        lists[gi].push_back(make_uint2(uint32_t(t) | (sp ? 0x80000000u : 0u),
                                                       uint32_t(z) | (uint32_t(z1) << 16)));
                    }
                }
        }
        size_t len = 0;
        for (auto &l : lists) len = std::max(len, l.size());
        std::vector<uint2> items;
        for (auto &l : lists) {
            l.resize(len, make_uint2(0u, 0u));
            items.insert(items.end(), l.begin(), l.end());
        }
        nblocks_ = int(items.size());
        xcd_ = groups;

        std::vector<R> wall_nx_all(std::max(nwall, 1)), wall_ny_all(std::max(nwall, 1));
        std::vector<uint32_t> wall_in_all(std::max(nwall, 1));
        for (int k = 0; k < nwall; ++k) {
            wall_nx_all[k] = w_nx[k];
            wall_ny_all[k] = w_ny[k];
            wall_in_all[k] = w_in[k];
        }
        std::vector<R> inlet(std::max<size_t>(g.inlet.size(), 1), R(0));
        for (size_t k = 0; k < g.inlet.size(); ++k) inlet[k] = R(g.inlet[k]);

        int rc;
        if ((rc = upload(d_info_, info))) return rc;
        if ((rc = upload(d_items_, items))) return rc;
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
        const int wall_blocks = std::max(1, (nwall + nbcf_ + 255) / 256);
        HIP_TRY(hipMalloc(&d_partials_, size_t(wall_blocks) * 4 * sizeof(double)));
        HIP_TRY(hipMalloc(&d_ticket_, sizeof(unsigned)));
        HIP_TRY(hipMemsetAsync(d_ticket_, 0, sizeof(unsigned), stream_));
        HIP_TRY(hipMalloc(&d_ring_, size_t(FORCE_RING_STEPS) * 4 * sizeof(double)));
        HIP_TRY(hipMalloc(&d_step_counter_, sizeof(unsigned long long)));
        HIP_TRY(hipMemsetAsync(d_step_counter_, 0, sizeof(unsigned long long), stream_));
        HIP_TRY(hipMemsetAsync(d_ring_, 0, size_t(FORCE_RING_STEPS) * 4 * sizeof(double), stream_));
        HIP_TRY(hipStreamSynchronize(stream_));
        return MRLBM_OK;
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
    // graph_steps_ steps (an even number, so the ping-pong parity is unchanged).
    int step(int nsteps, uint32_t flags) override {
        int done = 0;
        if (graph_steps_ > 0 && launched_) {
            while (nsteps - done >= graph_steps_) {
                hipGraphExec_t exec = nullptr;
                int rc = graph_for(flags, exec);
                if (rc) return rc;
                HIP_TRY(hipGraphLaunch(exec, stream_));
                steps += graph_steps_;
                done += graph_steps_;
            }
        }
        for (; done < nsteps; ++done) queue_step(flags, nullptr);
        HIP_TRY(hipGetLastError());
        return MRLBM_OK;
    }

    int set_steps(int64_t n) override {
        steps = n;
        const unsigned long long v = static_cast<unsigned long long>(n);
        HIP_TRY(hipMemcpy(d_step_counter_, &v, sizeof(v), hipMemcpyHostToDevice));
        return MRLBM_OK;
    }

    int sync() override {
        HIP_TRY(hipStreamSynchronize(stream_));
        return MRLBM_OK;
    }

    int get(int f, double *h) override {
        std::vector<double> all;
        int rc = get_all(all);
        if (rc) return rc;
        std::memcpy(h, all.data() + size_t(f) * g.n, size_t(g.n) * sizeof(double));
        return MRLBM_OK;
    }

    int get_raw(void *h) override {
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipMemcpy2D(h, size_t(g.n) * sizeof(S), buf_[cur_], stride_ * sizeof(S), size_t(g.n) * sizeof(S), 10,
                            hipMemcpyDeviceToHost));
        return MRLBM_OK;
    }

    int set_raw(const void *h) override {
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipMemcpy2D(buf_[cur_], stride_ * sizeof(S), h, size_t(g.n) * sizeof(S), size_t(g.n) * sizeof(S), 10,
                            hipMemcpyHostToDevice));
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
        return MRLBM_OK;
    }

    int time_steps(int n, uint32_t flags, double *ms_total, double *ms_k) override {
        if (n <= 0) return set_error(MRLBM_ERR_INVALID, "nsteps must be positive");
        std::vector<hipEvent_t> ev(size_t(4) * n + 2);
        for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipEventRecord(ev[0], stream_));
        if (ms_k) {
            for (int s = 0; s < n; ++s) queue_step(flags, &ev[2 + 4 * size_t(s)]);
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

  private:
    // The wall pass also advances the device step counter, which seeds the stochastic rounding of
    // the 16-bit forms, so those run it even without walls.
    bool has_wall_launch(uint32_t flags) const {
        return nwall_ > 0 || ((flags & MRLBM_STEP_FORCES) && nbcf_ > 0) || Conv<S>::scaled;
    }

    // Executable graph of graph_steps_ steps for these flags from the current buffer parity.
    int graph_for(uint32_t flags, hipGraphExec_t &exec) {
        const int key = ((flags & MRLBM_STEP_FORCES) ? 1 : 0) * 2 + cur_;
        if (!graphs_[key]) {
            const int cur0 = cur_;
            const int64_t steps0 = steps;
            HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (int s = 0; s < graph_steps_; ++s) queue_step(flags, nullptr);
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamEndCapture(stream_, &graph));
            cur_ = cur0;
            steps = steps0;
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
    void queue_step(uint32_t flags, hipEvent_t *ev) {
        FusedArgs<R, S> a{};
        a.src = buf_[cur_];
        a.dst = buf_[cur_ ^ 1];
        a.donor = donor_;
        a.stride = stride_;
        a.nx = g.nx;
        a.ny = g.ny;
        a.nz = g.nz;
        a.tiles_x = tiles_x_;
        a.nblocks = nblocks_;
        a.xcd_groups = xcd_;
        a.items = d_items_;
        a.info = d_info_;
        a.wall_nx = d_wnx_;
        a.wall_ny = d_wny_;
        a.wall_in = d_win_;
        a.bcf_in = d_bcfin_;
        a.inlet = d_inlet_;
        a.step_counter = stochastic_ ? d_step_counter_ : nullptr;
        a.scale = scale_;
        a.omega = omega_;
        a.rho0 = rho0_;
        if (ev) (void)hipEventRecord(ev[0], stream_);
        launch_fused<R, S>(a, bx_, by_, variant_, stream_);
        launched_ = true;
        if (ev) (void)hipEventRecord(ev[1], stream_);
        const bool forces = (flags & MRLBM_STEP_FORCES) && g.ngroups > 0;
        if (has_wall_launch(flags)) {
            WallArgs<R, S> w{};
            w.mom = buf_[cur_ ^ 1];
            w.donor = donor_;
            w.stride = stride_;
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
            w.bcf_node = d_bcfnode_;
            w.bcf_tag = d_bcftag_;
            w.bcf_in = d_bcfin_;
            w.bcf_out = d_bcfout_;
            for (int gi = 0; gi < 2; ++gi)
                for (int k = 0; k < 3; ++k) w.uw[gi][k] = R(g.wall[gi].uw[k]);
            w.scale = scale_;
            w.omega = omega_;
            w.rho0 = rho0_;
            w.forces = forces ? 1 : 0;
            w.partials = d_partials_;
            w.ticket = d_ticket_;
            w.ring = d_ring_;
            w.step_counter = d_step_counter_;
            if (ev) (void)hipEventRecord(ev[2], stream_);
            launch_wall<R, S>(w, stream_);
            if (ev) (void)hipEventRecord(ev[3], stream_);
        } else if (ev) {
            (void)hipEventRecord(ev[2], stream_);
            (void)hipEventRecord(ev[3], stream_);
        }
        cur_ ^= 1;
        ++steps;
    }

    template <class T>
    int upload(T *&dptr, const std::vector<T> &h) {
        const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
        HIP_TRY(hipMalloc(&dptr, bytes));
        if (!h.empty()) HIP_TRY(hipMemcpy(dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return MRLBM_OK;
    }

    int put_all(const std::vector<double> &h) {
        double *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, h.size() * sizeof(double)));
        HIP_TRY(hipMemcpy(tmp, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        launch_encode<R, S>(tmp, buf_[cur_], stride_, uint32_t(g.n), scale_, stream_);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipFree(tmp));
        return MRLBM_OK;
    }

    int get_all(std::vector<double> &h) {
        h.resize(size_t(10) * g.n);
        double *tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, h.size() * sizeof(double)));
        launch_decode<R, S>(buf_[cur_], stride_, tmp, uint32_t(g.n), scale_, stream_);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipMemcpy(h.data(), tmp, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipFree(tmp));
        return MRLBM_OK;
    }

    void release() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (auto &g : graphs_)
            if (g) (void)hipGraphExecDestroy(g);
        void *ptrs[] = {buf_[0], buf_[1], donor_, d_info_, d_items_, d_wconst_, d_wnx_, d_wny_, d_win_, d_wout_, d_wnode_,
                        d_wcoef_, d_stbase_, d_stfrac_, d_bcfnode_, d_bcftag_, d_bcfin_, d_bcfout_, d_inlet_,
                        d_partials_, d_ticket_, d_ring_, d_step_counter_};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
    }

    hipStream_t stream_ = nullptr;
    S *buf_[2] = {nullptr, nullptr};
    S *donor_ = nullptr;
    size_t stride_ = 0;
    int cur_ = 0;
    R omega_ = 0, rho0_ = 1;
    StoreScale<R> scale_{};
    int variant_ = 1;
    bool stochastic_ = Conv<S>::scaled;
    int graph_steps_ = 16;
    bool launched_ = false;
    hipGraphExec_t graphs_[4] = {nullptr, nullptr, nullptr, nullptr};
    int bx_ = 32, by_ = 8, lz_ = 16, tiles_x_ = 1, tiles_y_ = 1, ntiles_ = 1, nblocks_ = 1, xcd_ = 1;
    int nwall_ = 0, nb_inner_ = 0, nbcf_ = 0;
    uint32_t *d_info_ = nullptr;
    uint2 *d_items_ = nullptr;
    R *d_wconst_ = nullptr;
    R *d_wnx_ = nullptr, *d_wny_ = nullptr, *d_wcoef_ = nullptr, *d_stfrac_ = nullptr, *d_inlet_ = nullptr;
    uint32_t *d_win_ = nullptr, *d_wout_ = nullptr, *d_wnode_ = nullptr, *d_stbase_ = nullptr;
    uint32_t *d_bcfnode_ = nullptr, *d_bcfin_ = nullptr, *d_bcfout_ = nullptr;
    uint8_t *d_bcftag_ = nullptr;
    double *d_partials_ = nullptr, *d_ring_ = nullptr;
    unsigned long long *d_step_counter_ = nullptr;
    unsigned *d_ticket_ = nullptr;
};

}  // namespace mrlbm

struct mrlbm_sim {
    std::unique_ptr<mrlbm::SimBase> impl;
};

namespace {

uint64_t fnv1a(const void *data, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct CheckpointHeader {
    char magic[8];  // "MRLBMCK1"
    uint32_t version;
    uint32_t precision;
    int32_t nx, ny, nz;
    int32_t fields;
    int64_t steps;
    uint64_t payload_bytes;
    uint64_t checksum;  // FNV-1a over the payload
};

}  // namespace

extern "C" {

int32_t mrlbm_abi_version(void) { return MRLBM_ABI_VERSION; }

const char *mrlbm_last_error(void) { return g_last_error.c_str(); }

int mrlbm_case_desc_default(int32_t kind, mrlbm_case_desc *desc) {
    if (!desc) return set_error(MRLBM_ERR_INVALID, "null descriptor");
    std::memset(desc, 0, sizeof(*desc));
    desc->kind = kind;
    desc->rho0 = 1.0;
    desc->periodic_z = 1;
    desc->init_noise = 1e-4;
    desc->seed = 0x4D524C4200000000ull;
    if (kind == MRLBM_CASE_PERIODIC_BOX) {
        desc->precision = MRLBM_FP64;
        desc->nx = 64;
        desc->ny = 64;
        desc->nz = 16;
        desc->periodic_x = desc->periodic_y = 1;
        desc->tau = 0.8;
        desc->u_max = 0.05;
        desc->seed |= 0xB0;
    } else if (kind == MRLBM_CASE_ANNULUS) {
        // annular Couette flow, 64x64x16, R_in 12, R_out 28, inner wall speed 0.05, tau 0.8
        desc->precision = MRLBM_FP64;
        desc->nx = 64;
        desc->ny = 64;
        desc->nz = 16;
        desc->tau = 0.8;
        desc->u_max = 0.05;
        desc->xc = 32.0;
        desc->yc = 32.0;
        desc->r_inner = 12.0;
        desc->r_outer = 28.0;
        desc->u_wall_inner[1] = 0.05;
        desc->seed |= 0x01;
    } else if (kind == MRLBM_CASE_CYLINDER) {
        // confined cylinder, 256x128x64, D 20 at (64, 64), parabolic inlet u_max 0.05, Re 100
        desc->precision = MRLBM_FP32;
        desc->nx = 256;
        desc->ny = 128;
        desc->nz = 64;
        desc->u_max = 0.05;
        desc->tau = 0.5 + 3.0 * (0.05 * 20.0 / 100.0);
        desc->inlet_profile = MRLBM_INLET_PARABOLIC;
        desc->xc = 64.0;
        desc->yc = 64.0;
        desc->r_inner = 10.0;
        desc->seed |= 0x02;
    } else {
        return set_error(MRLBM_ERR_INVALID, "unknown case kind");
    }
    return MRLBM_OK;
}

int mrlbm_create(const mrlbm_case_desc *desc, mrlbm_sim **out) {
    if (!desc || !out) return set_error(MRLBM_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<mrlbm::SimBase> impl;
    switch (desc->precision) {
    case MRLBM_FP64: impl.reset(new mrlbm::Sim<double, double>); break;
    case MRLBM_FP32: impl.reset(new mrlbm::Sim<float, float>); break;
    case MRLBM_FP32_BF16: impl.reset(new mrlbm::Sim<float, __hip_bfloat16>); break;
    case MRLBM_FP32_FP16: impl.reset(new mrlbm::Sim<float, _Float16>); break;
    default: return set_error(MRLBM_ERR_INVALID, "unknown precision");
    }
    impl->d = *desc;
    std::string err;
    int rc = mrlbm::build_geometry(*desc, impl->g, err);
    if (rc) return set_error(rc, err);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(MRLBM_ERR_DEVICE, "no HIP device available");
    rc = impl->setup();
    if (rc) return rc;
    impl->device_ready = true;
    auto s = new mrlbm_sim;
    s->impl = std::move(impl);
    *out = s;
    return MRLBM_OK;
}

int mrlbm_create_geometry_only(const mrlbm_case_desc *desc, mrlbm_sim **out) {
    if (!desc || !out) return set_error(MRLBM_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<mrlbm::SimBase> impl;
    switch (desc->precision) {
    case MRLBM_FP64: impl.reset(new mrlbm::Sim<double, double>); break;
    case MRLBM_FP32: impl.reset(new mrlbm::Sim<float, float>); break;
    case MRLBM_FP32_BF16: impl.reset(new mrlbm::Sim<float, __hip_bfloat16>); break;
    case MRLBM_FP32_FP16: impl.reset(new mrlbm::Sim<float, _Float16>); break;
    default: return set_error(MRLBM_ERR_INVALID, "unknown precision");
    }
    impl->d = *desc;
    std::string err;
    int rc = mrlbm::build_geometry(*desc, impl->g, err);
    if (rc) return set_error(rc, err);
    auto s = new mrlbm_sim;
    s->impl = std::move(impl);
    *out = s;
    return MRLBM_OK;
}

#define SIM_CHECK(s)                                                                     \
    do {                                                                                 \
        if (!(s) || !(s)->impl) return set_error(MRLBM_ERR_INVALID, "null simulation"); \
    } while (0)

#define DEV_CHECK(s)                                                                                \
    do {                                                                                            \
        SIM_CHECK(s);                                                                               \
        if (!(s)->impl->device_ready) return set_error(MRLBM_ERR_INVALID, "geometry-only simulation"); \
    } while (0)

int mrlbm_init(mrlbm_sim *s) {
    DEV_CHECK(s);
    return s->impl->init();
}

int mrlbm_step(mrlbm_sim *s, int32_t nsteps, uint32_t flags) {
    DEV_CHECK(s);
    if (nsteps < 0) return set_error(MRLBM_ERR_INVALID, "negative step count");
    return s->impl->step(nsteps, flags);
}

int mrlbm_sync(mrlbm_sim *s) {
    DEV_CHECK(s);
    return s->impl->sync();
}

int mrlbm_get_moments(mrlbm_sim *s, int32_t field, double *host) {
    DEV_CHECK(s);
    if (field < 0 || field >= MRLBM_NUM_FIELDS || !host) return set_error(MRLBM_ERR_INVALID, "bad field or buffer");
    return s->impl->get(field, host);
}

int mrlbm_set_moments(mrlbm_sim *s, int32_t field, const double *host) {
    DEV_CHECK(s);
    if (field < 0 || field >= MRLBM_NUM_FIELDS || !host) return set_error(MRLBM_ERR_INVALID, "bad field or buffer");
    return s->impl->set(field, host);
}

int mrlbm_get_forces(mrlbm_sim *s, int64_t first, int32_t count, double *out4) {
    DEV_CHECK(s);
    if (!out4 && count > 0) return set_error(MRLBM_ERR_INVALID, "null buffer");
    return s->impl->get_forces(first, count, out4);
}

int mrlbm_get_node_types(mrlbm_sim *s, uint32_t *out) {
    SIM_CHECK(s);
    const mrlbm::Geometry &g = s->impl->g;
    for (int64_t i = 0; i < g.n; ++i) out[i] = (uint32_t(g.ref_type[i]) << 20) | (g.ref_index[i] & 0xFFFFFu);
    return MRLBM_OK;
}

int mrlbm_get_wall_table(mrlbm_sim *s, int32_t group, int64_t *node, double *nx, double *ny, double *delta,
                         uint32_t *in_mask, uint32_t *out_mask) {
    SIM_CHECK(s);
    if (group < 0 || group > 1) return set_error(MRLBM_ERR_INVALID, "bad wall group");
    const mrlbm::WallTable &w = s->impl->g.wall[group];
    for (size_t k = 0; k < w.node.size(); ++k) {
        node[k] = w.node[k];
        nx[k] = w.nx[k];
        ny[k] = w.ny[k];
        delta[k] = w.delta[k];
        in_mask[k] = w.in_mask[k];
        out_mask[k] = w.out_mask[k];
    }
    return MRLBM_OK;
}

int mrlbm_get_bcfluid_table(mrlbm_sim *s, int32_t group, int64_t *node, uint32_t *tag, uint32_t *in256,
                            uint32_t *out256) {
    SIM_CHECK(s);
    if (group < 0 || group > 1) return set_error(MRLBM_ERR_INVALID, "bad wall group");
    const mrlbm::Geometry &g = s->impl->g;
    const mrlbm::WallTable &w = g.wall[group];
    for (size_t k = 0; k < w.bcf_node.size(); ++k) {
        node[k] = w.bcf_node[k];
        tag[k] = w.bcf_tag[k];
    }
    for (int t = 0; t < 256; ++t) {
        in256[t] = g.bcf_in[t];
        out256[t] = g.bcf_out[t];
    }
    return MRLBM_OK;
}

int mrlbm_get_info(mrlbm_sim *s, mrlbm_info *info) {
    SIM_CHECK(s);
    const mrlbm::Geometry &g = s->impl->g;
    std::memset(info, 0, sizeof(*info));
    info->num_nodes = g.n;
    info->nb_inner = int32_t(g.wall[0].node.size());
    info->nb_outer = int32_t(g.wall[1].node.size());
    info->nb_fluid_inner = int32_t(g.wall[0].bcf_node.size());
    info->nb_fluid_outer = int32_t(g.wall[1].bcf_node.size());
    info->num_donors = g.num_donors;
    info->active_tiles = g.active_tiles;
    info->total_tiles = g.total_tiles;
    info->steps_done = s->impl->steps;
    info->precision = s->impl->d.precision;
    info->bytes_per_value = s->impl->bytes_per_value;
    return MRLBM_OK;
}

int mrlbm_time_steps(mrlbm_sim *s, int32_t nsteps, uint32_t flags, double *ms_total, double *ms_kernels) {
    DEV_CHECK(s);
    if (!ms_total) return set_error(MRLBM_ERR_INVALID, "null output");
    return s->impl->time_steps(nsteps, flags, ms_total, ms_kernels);
}

int mrlbm_checkpoint_write(mrlbm_sim *s, const char *path) {
    DEV_CHECK(s);
    if (!path) return set_error(MRLBM_ERR_INVALID, "null path");
    const mrlbm::Geometry &g = s->impl->g;
    // payload: the moments exactly as stored (rho - rho0, u, m, or the scaled deviations of the
    // 16-bit forms), so a restart continues bit for bit
    std::vector<unsigned char> payload(size_t(10) * g.n * s->impl->bytes_per_value);
    if (int rc = s->impl->get_raw(payload.data())) return rc;
    CheckpointHeader h{};
    std::memcpy(h.magic, "MRLBMCK1", 8);
    h.version = 2;
    h.precision = uint32_t(s->impl->d.precision);
    h.nx = g.nx;
    h.ny = g.ny;
    h.nz = g.nz;
    h.fields = 10;
    h.steps = s->impl->steps;
    h.payload_bytes = payload.size();
    h.checksum = fnv1a(payload.data(), h.payload_bytes);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(MRLBM_ERR_IO, std::string("cannot open ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, fp) == 1 &&
              std::fwrite(payload.data(), 1, h.payload_bytes, fp) == h.payload_bytes;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? MRLBM_OK : set_error(MRLBM_ERR_IO, "checkpoint write failed");
}

int mrlbm_checkpoint_read(mrlbm_sim *s, const char *path) {
    DEV_CHECK(s);
    if (!path) return set_error(MRLBM_ERR_INVALID, "null path");
    const mrlbm::Geometry &g = s->impl->g;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(MRLBM_ERR_IO, std::string("cannot open ") + path);
    CheckpointHeader h{};
    std::vector<unsigned char> payload(size_t(10) * g.n * s->impl->bytes_per_value);
    const bool head = std::fread(&h, sizeof(h), 1, fp) == 1;
    bool ok = head && std::memcmp(h.magic, "MRLBMCK1", 8) == 0 && h.version == 2 &&
              h.precision == uint32_t(s->impl->d.precision) && h.nx == g.nx && h.ny == g.ny && h.nz == g.nz &&
              h.fields == 10 && h.payload_bytes == payload.size();
    if (ok) ok = std::fread(payload.data(), 1, h.payload_bytes, fp) == h.payload_bytes;
    if (ok) {
        char extra;
        ok = std::fread(&extra, 1, 1, fp) == 0;  // no trailing bytes
    }
    std::fclose(fp);
    if (!ok) return set_error(MRLBM_ERR_CORRUPT, "checkpoint header, precision or size mismatch");
    if (fnv1a(payload.data(), h.payload_bytes) != h.checksum)
        return set_error(MRLBM_ERR_CORRUPT, "checkpoint checksum mismatch");
    if (int rc = s->impl->set_raw(payload.data())) return rc;
    return s->impl->set_steps(h.steps);
}

int mrlbm_destroy(mrlbm_sim *s) {
    delete s;
    return MRLBM_OK;
}

}  // extern "C"
