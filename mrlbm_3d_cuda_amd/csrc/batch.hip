// batch.hip - many cases of one precision stepped together: one fused launch over the work
// items of every case and one wall launch over the wall nodes of every case per time step,
// captured in HIP graphs, with a launch configuration chosen by a timing tuner.
//
// The reference runs one compile-time case per process (src/solve_mrlbm.cu:3-24); a sweep of
// uneven cases there is a sequence of runs. Here each case keeps its own buffers, tables, force
// ring and step counter (the same objects a standalone simulation uses), and only the launches
// are shared.
#include <algorithm>
#include <utility>
#include <vector>

#include "sim.h"

using mrlbm::set_error;

namespace mrlbm {

class BatchBase {
  public:
    virtual ~BatchBase() = default;
    virtual int init() = 0;
    virtual int step(int n, uint32_t flags) = 0;
    virtual int sync() = 0;
    virtual int time_steps(int n, uint32_t flags, double *ms_total) = 0;
    virtual int configure(const LaunchConfig &lc) = 0;
    virtual int autotune(int trial_steps, mrlbm_tune_result *out) = 0;
    std::vector<std::unique_ptr<mrlbm_sim>> cases;
    LaunchConfig launch;
    int64_t steps = 0;
};

template <class R, class S>
class Batch final : public BatchBase {
  public:
    ~Batch() override { release(); }

    Sim<R, S> &sim(size_t c) { return static_cast<Sim<R, S> &>(*cases[c]->impl); }

    int setup() {
        HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        for (auto &c : cases) c->impl->batch_stream = stream_;
        if (const char *e = std::getenv("MRLBM_GRAPH_STEPS")) graph_steps_ = std::max(0, std::atoi(e) / 2 * 2);
        // wall arguments of both parities and block maps, without / with forces
        const size_t nc = cases.size();
        for (int p = 0; p < 2; ++p)
            for (int f = 0; f < 2; ++f) {
                std::vector<WallArgs<R, S>> wa(nc);
                for (size_t c = 0; c < nc; ++c) wa[c] = sim(c).wall_args(p, f && sim(c).forces_on(MRLBM_STEP_FORCES));
                if (int rc = upload(d_wargs_[p][f], wa)) return rc;
            }
        for (int f = 0; f < 2; ++f) {
            const uint32_t flags = f ? MRLBM_STEP_FORCES : 0u;
            std::vector<int2> blocks;
            for (size_t c = 0; c < nc; ++c) {
                Sim<R, S> &s = sim(c);
                if (!s.has_wall_launch(flags)) continue;
                const int nb = wall_blocks(s.nwall(), s.forces_on(flags) ? s.nbcf() : 0);
                for (int b = 0; b < nb; ++b) blocks.push_back(make_int2(int(c), b));
            }
            nwblocks_[f] = int(blocks.size());
            std::vector<int2> order(blocks.size());  // dispatch slot -> block (XCD-contiguous)
            for (size_t b = 0; b < blocks.size(); ++b) order[b] = blocks[size_t(xcd_contiguous(int(b), int(blocks.size())))];
            int rc = upload(d_wblocks_[f], order);
            if (rc) return rc;
        }
        // the first case's launch configuration, with the batch's own placement from the environment
        LaunchConfig lc = sim(0).launch_config();
        // boundary-path tiles in whole z chunks until the tuner says otherwise (the shorter
        // chunks of a single cache-resident case made 64 annuli slower: 20.9 -> 18.4-20.0 cases/s)
        if (!std::getenv("MRLBM_LZ_SPECIAL")) lc.lz_special = 0;
        if (const char *e = std::getenv("MRLBM_BATCH_BLOCKS_PER_CU")) lc.blocks_per_cu = std::max(0, std::atoi(e));
        if (const char *e = std::getenv("MRLBM_BATCH_PACKING")) lc.packing = std::atoi(e) == 1 ? 1 : 0;
        return configure(lc);
    }

    // Applies one launch configuration to every case and merges their work items into one list
    // per XCD group. Packing 0: group g of the launch takes group g of every case that is cut into
    // XCD slabs, and the whole item list of case c otherwise goes to group c mod 8. Packing 1:
    // every case stays whole (one item list) on the XCD with the least work so far, largest case
    // first, so that a case's ring loads share one L2. Boundary-tile items lead each group.
    int configure(const LaunchConfig &lc) override {
        if (!batch_tile_supported(lc.bx, lc.by)) return set_error(MRLBM_ERR_INVALID, "unsupported batch tile");
        HIP_TRY(hipStreamSynchronize(stream_));
        if (prepare_fused_batch<R, S>(lc.bx, lc.by, lc.variant) != 0)
            return set_error(MRLBM_ERR_UNSUPPORTED, "batched fused kernel cannot get its LDS");
        LaunchConfig per_case = lc;
        if (lc.packing == 1) per_case.xcd_remap = false;
        for (size_t c = 0; c < cases.size(); ++c)
            if (int rc = sim(c).configure_launch(per_case)) return rc;
        const int groups = lc.xcd_remap ? 8 : 1;
        // destination group of every (case, case group)
        std::vector<std::vector<int>> dst(cases.size());
        if (lc.packing == 1 && groups > 1) {
            // work of a case: nodes of its items (boundary tiles count 1.5 times)
            std::vector<std::pair<double, size_t>> work;
            for (size_t c = 0; c < cases.size(); ++c) {
                const Sim<R, S> &sc = sim(c);
                double w = 0.0;
                for (const auto &l : sc.item_lists())
                    for (const uint2 &it : l)
                        w += double(int(it.y >> 16) - int(it.y & 0xFFFFu)) * lc.bx * lc.by * ((it.x >> 31) ? 1.5 : 1.0);
                work.emplace_back(w, c);
            }
            std::stable_sort(work.begin(), work.end(), [](const auto &p, const auto &q) { return p.first > q.first; });
            std::vector<double> load(groups, 0.0);
            for (const auto &w : work) {
                const int gmin = int(std::min_element(load.begin(), load.end()) - load.begin());
                load[gmin] += w.first;
                dst[w.second].assign(sim(w.second).item_lists().size(), gmin);
            }
        } else {
            for (size_t c = 0; c < cases.size(); ++c) {
                const size_t ng = sim(c).item_lists().size();
                dst[c].resize(ng);
                for (size_t gi = 0; gi < ng; ++gi) dst[c][gi] = ng == 1 ? int(c % groups) : int(gi % groups);
            }
        }
        std::vector<std::vector<uint4>> lists(groups);
        for (int pass = 0; pass < 2; ++pass)
            for (size_t c = 0; c < cases.size(); ++c) {
                const auto &cl = sim(c).item_lists();
                for (size_t gi = 0; gi < cl.size(); ++gi) {
                    for (const uint2 &it : cl[gi]) {
                        const bool sp = (it.x >> 31) != 0;
                        if ((it.y & 0xFFFFu) >= (it.y >> 16) || sp != (pass == 0)) continue;
                        lists[dst[c][gi]].push_back(make_uint4(uint32_t(c), it.x, it.y, 0u));
                    }
                }
            }
        size_t len = 0;
        for (auto &l : lists) len = std::max(len, l.size());
        std::vector<uint4> items;
        for (auto &l : lists) {
            l.resize(len, make_uint4(0u, 0u, 0u, 0u));
            items.insert(items.end(), l.begin(), l.end());
        }
        if (d_items_) HIP_TRY(hipFree(d_items_));
        d_items_ = nullptr;
        if (int rc = upload(d_items_, items)) return rc;
        // fused arguments of both parities (they carry the tiling)
        for (int p = 0; p < 2; ++p) {
            std::vector<FusedArgs<R, S>> fa(cases.size());
            for (size_t c = 0; c < cases.size(); ++c) fa[c] = sim(c).fused_args(p);
            if (d_fargs_[p]) HIP_TRY(hipFree(d_fargs_[p]));
            d_fargs_[p] = nullptr;
            if (int rc = upload(d_fargs_[p], fa)) return rc;
        }
        nitems_ = int(items.size());
        groups_ = groups;
        // persistent launch: blocks_per_cu blocks on every CU, a multiple of the groups, each group
        // counted down by its own counters (both parities start at zero)
        pblocks_ = 0;
        if (lc.blocks_per_cu > 0) {
            int dev = 0, ncu = 0;
            HIP_TRY(hipGetDevice(&dev));
            HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            pblocks_ = (std::max(1, ncu * lc.blocks_per_cu) + groups - 1) / groups * groups;
        }
        if (!d_counters_) HIP_TRY(hipMalloc(&d_counters_, 2 * 8 * sizeof(unsigned)));
        HIP_TRY(hipMemset(d_counters_, 0, 2 * 8 * sizeof(unsigned)));
        launch = lc;
        drop_graphs();
        return MRLBM_OK;
    }

    int init() override {
        HIP_TRY(hipStreamSynchronize(stream_));  // the cases' own streams take over
        for (size_t c = 0; c < cases.size(); ++c)
            if (int rc = sim(c).init()) return rc;
        steps = 0;
        return MRLBM_OK;
    }

    // A case whose NaN/Inf watchdog has tripped stops the whole batch (MRLBM_ERR_NONFINITE).
    int check_finite() const {
        for (const auto &c : cases)
            if (c->impl->nonfinite())
                return set_error(MRLBM_ERR_NONFINITE, "non-finite moments in a batched case: the watchdog tripped");
        return MRLBM_OK;
    }

    int step(int nsteps, uint32_t flags) override {
        if (int rc = check_finite()) return rc;
        int done = 0;
        if (graph_steps_ > 0) {
            while (nsteps - done >= graph_steps_) {
                hipGraphExec_t exec = nullptr;
                if (int rc = graph_for(flags, exec)) return rc;
                HIP_TRY(hipGraphLaunch(exec, stream_));
                for (int s = 0; s < graph_steps_; ++s) advance(flags);
                done += graph_steps_;
            }
        }
        for (; done < nsteps; ++done) {
            queue_step(flags, cur_);
            advance(flags);
        }
        // the last step's force records are filed here (the fused launch of a step files the
        // previous step's)
        if (flags & MRLBM_STEP_FORCES) launch_record_batch<R, S>(d_fargs_[cur_], int(cases.size()), stream_);
        HIP_TRY(hipGetLastError());
        return MRLBM_OK;
    }

    int sync() override {
        HIP_TRY(hipStreamSynchronize(stream_));
        return check_finite();
    }

    int time_steps(int n, uint32_t flags, double *ms_total) override {
        if (n <= 0) return set_error(MRLBM_ERR_INVALID, "nsteps must be positive");
        if (graph_steps_ > 0 && n >= graph_steps_) {
            hipGraphExec_t exec = nullptr;
            if (int rc = graph_for(flags, exec)) return rc;  // captured outside the timed region
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        HIP_TRY(hipStreamSynchronize(stream_));
        HIP_TRY(hipEventRecord(e0, stream_));
        int rc = step(n, flags);
        HIP_TRY(hipEventRecord(e1, stream_));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_total = ms;
        return rc;
    }

    // Times trial_steps steps (with forces) of every candidate configuration and keeps the
    // fastest, in two stages: tile 32x8, 32x4 or 16x8 (where its LDS fits) x ring or
    // register-window march x z chunks of 4, 8, 16, 32 planes; then, around the fastest of those,
    // the assignment of the cases to the CUs -- one block per work item or a persistent launch with
    // 1, 2 or 3 blocks per CU, cases cut into XCD slabs (z or y) or kept whole on the least loaded
    // XCD, and (one block per item) boundary-path tiles in whole or half z chunks. Trials run a few eager steps, which differ by a few percent from run to run and from
    // graph replays (on 64 annuli a z chunk of 16 won the eager trials and one of 8 the
    // replays), so the four fastest configurations of both stages are timed again as the run
    // steps -- two replays of the step graph (or twice the trial steps where graphs are off) --
    // twice each in turn, and the best of those times decides. The cases' states are scrambled by the
    // trial steps: the tuner re-initialises them at the end.
    int autotune(int trial_steps, mrlbm_tune_result *out) override {
        if (trial_steps <= 0) return set_error(MRLBM_ERR_INVALID, "trial steps must be positive");
        const LaunchConfig def = launch;  // the configuration in use is timed first
        int dev = 0, max_lds = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        auto same = [](const LaunchConfig &p, const LaunchConfig &q) {
            return p.bx == q.bx && p.by == q.by && p.variant == q.variant && p.lz == q.lz &&
                   p.lz_special == q.lz_special && p.blocks_per_cu == q.blocks_per_cu && p.packing == q.packing &&
                   p.slab_axis == q.slab_axis && p.xcd_remap == q.xcd_remap;
        };
        if (int rc = init()) return rc;
        const int graph_steps = graph_steps_;
        graph_steps_ = 0;  // trials run eagerly: no graph capture per candidate
        double best = 1e300, at_default = 0.0;
        LaunchConfig best_lc = def;
        int tried = 0;
        std::vector<std::pair<double, LaunchConfig>> timed;
        auto time_config = [&](const LaunchConfig &lc, int nsteps, double &ms) -> int {
            if (int rc = configure(lc)) return rc;
            // two eager steps, and one untimed replay (capture, instantiation, upload) of the
            // step graph where graphs are on
            int rc = step(2 + graph_steps_, MRLBM_STEP_FORCES);
            if (!rc) rc = time_steps(nsteps, MRLBM_STEP_FORCES, &ms);
            ms /= nsteps;
            return rc;
        };
        auto trial = [&](const LaunchConfig &lc) -> int {
            double ms = 0.0;
            if (int rc = time_config(lc, trial_steps, ms)) return rc;
            ++tried;
            if (tried == 1) at_default = ms;
            timed.emplace_back(ms, lc);
            if (ms < best) {
                best = ms;
                best_lc = lc;
            }
            return MRLBM_OK;
        };
        std::vector<LaunchConfig> cands{def};
        const int tiles[3][2] = {{32, 8}, {32, 4}, {16, 8}};
        for (const auto &t : tiles)
            for (int variant : {1, 3, 4})
                for (int lz : {4, 8, 16, 32}) {
                    LaunchConfig lc = def;
                    lc.xcd_remap = true;
                    lc.bx = t[0];
                    lc.by = t[1];
                    lc.variant = variant;
                    lc.lz = lz;
                    lc.lz_special = 0;
                    if (fused_lds_bytes(lc.bx, lc.by, sizeof(R), lc.variant) > size_t(max_lds)) continue;
                    if (!same(lc, def)) cands.push_back(lc);
                }
        for (const LaunchConfig &lc : cands)
            if (int rc = trial(lc)) {
                graph_steps_ = graph_steps;
                return rc;
            }
        // second stage: the cases' assignment to the CUs around the fastest configuration
        const LaunchConfig base = best_lc;
        for (int packing : {0, 1})
            for (int axis : {2, 1})
                for (int bpc : {0, 1, 2, 3})
                    for (int lzs : {0, base.lz / 2}) {
                        if (packing == 1 && axis == 1) continue;  // whole cases are not cut into slabs
                        if (lzs > 0 && bpc > 0) continue;         // half chunks: with one block per item
                        LaunchConfig lc = base;
                        lc.packing = packing;
                        lc.slab_axis = axis;
                        lc.blocks_per_cu = bpc;
                        lc.lz_special = lzs;
                        if (same(lc, base) || (lzs > 0 && lzs == base.lz)) continue;
                        if (int rc = trial(lc)) {
                            graph_steps_ = graph_steps;
                            return rc;
                        }
                    }
        // third stage: the four fastest again, as the run steps, in turn, twice
        graph_steps_ = graph_steps;
        std::stable_sort(timed.begin(), timed.end(),
                         [](const auto &p, const auto &q) { return p.first < q.first; });
        timed.resize(std::min<size_t>(timed.size(), 4));
        std::vector<double> again(timed.size(), 1e300);
        const int n3 = graph_steps_ > 0 ? 2 * graph_steps_ : 2 * trial_steps;
        for (int pass = 0; pass < 2; ++pass)
            for (size_t k = 0; k < timed.size(); ++k) {
                double ms = 0.0;
                if (int rc = time_config(timed[k].second, n3, ms)) return rc;
                again[k] = std::min(again[k], ms);
            }
        for (size_t k = 0; k < timed.size(); ++k)
            if (k == 0 || again[k] < best) {
                best = again[k];
                best_lc = timed[k].second;
            }
        graph_steps_ = graph_steps;
        if (int rc = configure(best_lc)) return rc;
        if (int rc = init()) return rc;
        if (out) {
            out->bx = best_lc.bx;
            out->by = best_lc.by;
            out->variant = best_lc.variant;
            out->lz = best_lc.lz;
            out->candidates = tried;
            out->blocks_per_cu = best_lc.blocks_per_cu;
            out->ms_per_step = best;
            out->ms_per_step_default = at_default;
            out->packing = best_lc.packing;
            out->slab_axis = best_lc.slab_axis;
            out->lz_special = best_lc.lz_special;
        }
        return MRLBM_OK;
    }

  private:
    void advance(uint32_t flags) {
        cur_ ^= 1;
        ++steps;
        for (size_t c = 0; c < cases.size(); ++c) sim(c).advance(flags);
    }

    // One step from buffer parity `cur` (the caller advances the parity).
    void queue_step(uint32_t flags, int cur) {
        const BatchQueue q{pblocks_, d_counters_, cur};
        launch_fused_batch<R, S>(d_fargs_[cur], int(cases.size()), d_items_, nitems_, groups_, launch.bx, launch.by,
                                 launch.variant, q, stream_);
        const int f = (flags & MRLBM_STEP_FORCES) ? 1 : 0;
        if (nwblocks_[f] > 0)
            launch_wall_batch<R, S>(d_wargs_[cur][f], d_wblocks_[f], nwblocks_[f], stream_);
        if (flags & MRLBM_STEP_STATS)  // wall-pressure and plane statistics, launches per case
            for (size_t c = 0; c < cases.size(); ++c) {
                if (sim(c).stats_on(flags)) launch_pressure<R, S>(sim(c).pressure_args(cur), stream_);
                if (sim(c).planes_on(flags)) launch_planes<R, S>(sim(c).plane_args(cur), stream_);
            }
    }

    int graph_for(uint32_t flags, hipGraphExec_t &exec) {
        const int key = int(flags & (MRLBM_STEP_FORCES | MRLBM_STEP_STATS)) * 2 + cur_;
        if (!graphs_[key]) {
            HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            for (int s = 0; s < graph_steps_; ++s) queue_step(flags, cur_ ^ (s & 1));
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamEndCapture(stream_, &graph));
            hipGraphExec_t e = nullptr;
            const hipError_t err = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_TRY(err);
            graphs_[key] = e;
        }
        exec = graphs_[key];
        return MRLBM_OK;
    }

    void drop_graphs() {
        for (auto &g : graphs_) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }

    template <class T>
    int upload(T *&dptr, const std::vector<T> &h) {
        const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
        HIP_TRY(hipMalloc(&dptr, bytes));
        if (!h.empty()) HIP_TRY(hipMemcpy(dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return MRLBM_OK;
    }

    void release() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        drop_graphs();
        void *ptrs[] = {d_fargs_[0], d_fargs_[1], d_wargs_[0][0], d_wargs_[0][1], d_wargs_[1][0], d_wargs_[1][1],
                        d_wblocks_[0], d_wblocks_[1], d_items_, d_counters_};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        for (auto &c : cases) c->impl->batch_stream = nullptr;
        if (stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
    }

    hipStream_t stream_ = nullptr;
    int cur_ = 0;
    int graph_steps_ = 16;
    FusedArgs<R, S> *d_fargs_[2] = {nullptr, nullptr};
    WallArgs<R, S> *d_wargs_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int2 *d_wblocks_[2] = {nullptr, nullptr};
    int nwblocks_[2] = {0, 0};
    uint4 *d_items_ = nullptr;
    int nitems_ = 0, groups_ = 1;
    int pblocks_ = 0;                  // persistent fused launch: resident blocks (0: one per item)
    unsigned *d_counters_ = nullptr;   // its work-queue counters, 2 parities x 8 groups
    hipGraphExec_t graphs_[8] = {};
};

// Cases admitted from the front of the list until their device bytes exceed the budget
// (0: the device's free memory less 1 GiB).
inline int plan_batch(const mrlbm_case_desc *descs, int n, uint64_t budget, int32_t &admitted, uint64_t &bytes) {
    admitted = 0;
    bytes = 0;
    for (int i = 0; i < n; ++i) {
        if (descs[i].precision != descs[0].precision)
            return set_error(MRLBM_ERR_INVALID, "all cases of a batch share one precision");
        const uint64_t b = case_device_bytes(descs[i]);
        if (bytes + b > budget) break;
        bytes += b;
        ++admitted;
    }
    return MRLBM_OK;
}

}  // namespace mrlbm

struct mrlbm_batch {
    std::unique_ptr<mrlbm::BatchBase> impl;
};

#define BATCH_CHECK(b)                                                                 \
    do {                                                                               \
        if (!(b) || !(b)->impl) return set_error(MRLBM_ERR_INVALID, "null batch");    \
    } while (0)

extern "C" {

int mrlbm_batch_plan(const mrlbm_case_desc *descs, int32_t n, uint64_t hbm_budget_bytes, int32_t *admitted,
                     uint64_t *bytes) {
    if (!descs || n <= 0 || !admitted || !bytes) return set_error(MRLBM_ERR_INVALID, "bad batch plan arguments");
    if (hbm_budget_bytes == 0) return set_error(MRLBM_ERR_INVALID, "a plan needs an explicit budget");
    return mrlbm::plan_batch(descs, n, hbm_budget_bytes, *admitted, *bytes);
}

int mrlbm_batch_create(const mrlbm_case_desc *descs, int32_t n, uint64_t hbm_budget_bytes, mrlbm_batch **out,
                       int32_t *admitted) {
    if (!descs || n <= 0 || !out) return set_error(MRLBM_ERR_INVALID, "bad batch arguments");
    *out = nullptr;
    uint64_t budget = hbm_budget_bytes;
    if (budget == 0) {
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        budget = free_b > (size_t(1) << 30) ? uint64_t(free_b - (size_t(1) << 30)) : 0;
    }
    int32_t adm = 0;
    uint64_t bytes = 0;
    if (int rc = mrlbm::plan_batch(descs, n, budget, adm, bytes)) return rc;
    if (admitted) *admitted = adm;
    if (adm == 0) return set_error(MRLBM_ERR_INVALID, "no case fits the memory budget");
    std::unique_ptr<mrlbm::BatchBase> impl;
    switch (descs[0].precision) {
    case MRLBM_FP64: impl.reset(new mrlbm::Batch<double, double>); break;
    case MRLBM_FP32: impl.reset(new mrlbm::Batch<float, float>); break;
    case MRLBM_FP32_BF16: impl.reset(new mrlbm::Batch<float, __hip_bfloat16>); break;
    case MRLBM_FP32_FP16: impl.reset(new mrlbm::Batch<float, _Float16>); break;
    default: return set_error(MRLBM_ERR_INVALID, "unknown precision");
    }
    for (int i = 0; i < adm; ++i) {
        std::unique_ptr<mrlbm::SimBase> sim;
        if (int rc = mrlbm::make_sim(descs[i], true, sim)) return rc;
        auto h = std::make_unique<mrlbm_sim>();
        h->impl = std::move(sim);
        h->in_batch = true;
        impl->cases.push_back(std::move(h));
    }
    int rc = 0;
    switch (descs[0].precision) {
    case MRLBM_FP64: rc = static_cast<mrlbm::Batch<double, double> &>(*impl).setup(); break;
    case MRLBM_FP32: rc = static_cast<mrlbm::Batch<float, float> &>(*impl).setup(); break;
    case MRLBM_FP32_BF16: rc = static_cast<mrlbm::Batch<float, __hip_bfloat16> &>(*impl).setup(); break;
    default: rc = static_cast<mrlbm::Batch<float, _Float16> &>(*impl).setup(); break;
    }
    if (rc) return rc;
    auto b = new mrlbm_batch;
    b->impl = std::move(impl);
    *out = b;
    return MRLBM_OK;
}

int mrlbm_batch_init(mrlbm_batch *b) {
    BATCH_CHECK(b);
    return b->impl->init();
}

int mrlbm_batch_step(mrlbm_batch *b, int32_t nsteps, uint32_t flags) {
    BATCH_CHECK(b);
    if (nsteps < 0) return set_error(MRLBM_ERR_INVALID, "negative step count");
    return b->impl->step(nsteps, flags);
}

int mrlbm_batch_sync(mrlbm_batch *b) {
    BATCH_CHECK(b);
    return b->impl->sync();
}

int mrlbm_batch_num_cases(mrlbm_batch *b, int32_t *n) {
    BATCH_CHECK(b);
    if (!n) return set_error(MRLBM_ERR_INVALID, "null output");
    *n = int32_t(b->impl->cases.size());
    return MRLBM_OK;
}

int mrlbm_batch_case(mrlbm_batch *b, int32_t i, mrlbm_sim **out) {
    BATCH_CHECK(b);
    if (!out || i < 0 || size_t(i) >= b->impl->cases.size()) return set_error(MRLBM_ERR_INVALID, "bad case index");
    *out = b->impl->cases[size_t(i)].get();
    return MRLBM_OK;
}

int mrlbm_batch_time_steps(mrlbm_batch *b, int32_t nsteps, uint32_t flags, double *ms_total) {
    BATCH_CHECK(b);
    if (!ms_total) return set_error(MRLBM_ERR_INVALID, "null output");
    return b->impl->time_steps(nsteps, flags, ms_total);
}

int mrlbm_batch_autotune(mrlbm_batch *b, int32_t trial_steps, mrlbm_tune_result *out) {
    BATCH_CHECK(b);
    return b->impl->autotune(trial_steps, out);
}

int mrlbm_batch_get_launch(mrlbm_batch *b, mrlbm_tune_result *out) {
    BATCH_CHECK(b);
    if (!out) return set_error(MRLBM_ERR_INVALID, "null output");
    std::memset(out, 0, sizeof(*out));
    out->bx = b->impl->launch.bx;
    out->by = b->impl->launch.by;
    out->variant = b->impl->launch.variant;
    out->lz = b->impl->launch.lz;
    out->blocks_per_cu = b->impl->launch.blocks_per_cu;
    out->packing = b->impl->launch.packing;
    out->slab_axis = b->impl->launch.slab_axis;
    out->lz_special = b->impl->launch.lz_special;
    return MRLBM_OK;
}

int mrlbm_batch_destroy(mrlbm_batch *b) {
    delete b;
    return MRLBM_OK;
}

}  // extern "C"
