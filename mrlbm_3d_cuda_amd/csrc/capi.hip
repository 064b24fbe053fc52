// capi.hip - the extern "C" entry points of include/mrlbm.h for single simulations.
//
// Errors are returned as status codes with a thread-local message instead of exit().
#include "sim.h"

using mrlbm::set_error;
using mrlbm::FORCE_RING_STEPS;


namespace {

uint64_t fnv1a(const void *data, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Version 4: 16-bit payloads hold u_x - u_ref (u_ref in the header), and the BF16 form's velocity
// is FP16 bits times 16.
constexpr uint32_t CHECKPOINT_VERSION = 4;

struct CheckpointHeader {
    char magic[8];  // "MRLBMCK1"
    uint32_t version;
    uint32_t precision;
    int32_t nx, ny, nz;
    int32_t fields;
    int64_t steps;
    uint64_t payload_bytes;
    uint64_t checksum;     // FNV-1a over the payload
    double velocity_ref;   // u_ref of the 16-bit forms' stored u_x - u_ref (0 otherwise)
};

}  // namespace

extern "C" {

int32_t mrlbm_abi_version(void) { return MRLBM_ABI_VERSION; }

const char *mrlbm_last_error(void) { return mrlbm::g_last_error.c_str(); }

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
    if (int rc = mrlbm::make_sim(*desc, true, impl)) return rc;
    auto s = new mrlbm_sim;
    s->impl = std::move(impl);
    *out = s;
    return MRLBM_OK;
}

int mrlbm_create_geometry_only(const mrlbm_case_desc *desc, mrlbm_sim **out) {
    if (!desc || !out) return set_error(MRLBM_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<mrlbm::SimBase> impl;
    if (int rc = mrlbm::make_sim(*desc, false, impl)) return rc;
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

// calls that change a case's state or lifetime belong to its batch
#define OWN_CHECK(s)                                                                               \
    do {                                                                                           \
        DEV_CHECK(s);                                                                              \
        if ((s)->in_batch) return set_error(MRLBM_ERR_INVALID, "case is stepped by its batch");    \
    } while (0)

int mrlbm_init(mrlbm_sim *s) {
    OWN_CHECK(s);
    return s->impl->init();
}

int mrlbm_step(mrlbm_sim *s, int32_t nsteps, uint32_t flags) {
    OWN_CHECK(s);
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
    if (!out) return set_error(MRLBM_ERR_INVALID, "null output");
    const mrlbm::Geometry &g = s->impl->g;
    for (int64_t i = 0; i < g.n; ++i) out[i] = (uint32_t(g.ref_type[i]) << 20) | (g.ref_index[i] & 0xFFFFFu);
    return MRLBM_OK;
}

int mrlbm_get_wall_table(mrlbm_sim *s, int32_t group, int64_t *node, double *nx, double *ny, double *delta,
                         uint32_t *in_mask, uint32_t *out_mask) {
    SIM_CHECK(s);
    if (group < 0 || group > 1) return set_error(MRLBM_ERR_INVALID, "bad wall group");
    const mrlbm::WallTable &w = s->impl->g.wall[group];
    if (!w.node.empty() && (!node || !nx || !ny || !delta || !in_mask || !out_mask))
        return set_error(MRLBM_ERR_INVALID, "null output");
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
    if ((!w.bcf_node.empty() && (!node || !tag)) || !in256 || !out256)
        return set_error(MRLBM_ERR_INVALID, "null output");
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

int mrlbm_set_geometry(mrlbm_sim *s, const uint32_t *node_types, int32_t ngroups, const mrlbm_wall_desc *walls,
                       const uint32_t *bcf_in256, const uint32_t *bcf_out256) {
    SIM_CHECK(s);
    if (s->in_batch) return set_error(MRLBM_ERR_INVALID, "case is stepped by its batch");
    mrlbm::SimBase &sim = *s->impl;
    mrlbm::Geometry g = sim.g;
    std::string err;
    if (int rc = mrlbm::geometry_from_tables(sim.d, node_types, ngroups, walls, bcf_in256, bcf_out256, g, err))
        return set_error(rc, err);
    if (!sim.device_ready) {
        sim.g = std::move(g);
        return MRLBM_OK;
    }
    mrlbm::Geometry old = std::move(sim.g);
    sim.g = std::move(g);
    if (int rc = sim.apply_geometry()) {
        // the tables of the new geometry failed (a wall stencil outside the grid): restore the old
        const std::string msg = mrlbm::g_last_error;
        sim.g = std::move(old);
        (void)sim.apply_geometry();
        return set_error(rc, msg);
    }
    return MRLBM_OK;
}

int mrlbm_get_info(mrlbm_sim *s, mrlbm_info *info) {
    SIM_CHECK(s);
    if (!info) return set_error(MRLBM_ERR_INVALID, "null output");
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
    OWN_CHECK(s);
    if (!ms_total) return set_error(MRLBM_ERR_INVALID, "null output");
    return s->impl->time_steps(nsteps, flags, ms_total, ms_kernels);
}

int mrlbm_time_replays(mrlbm_sim *s, int32_t nsteps, uint32_t flags, double *ms_total, double *ms_step_median,
                       int32_t *replays) {
    OWN_CHECK(s);
    if (!ms_total) return set_error(MRLBM_ERR_INVALID, "null output");
    int nrep = 0;
    const int rc = s->impl->time_replays(nsteps, flags, ms_total, ms_step_median, &nrep);
    if (replays) *replays = nrep;
    return rc;
}

int mrlbm_time_memory_twin(mrlbm_sim *s, int32_t launches, double *ms_per_launch) {
    OWN_CHECK(s);
    if (!ms_per_launch) return set_error(MRLBM_ERR_INVALID, "null output");
    return s->impl->time_memory_twin(launches, ms_per_launch);
}

int mrlbm_checkpoint_write(mrlbm_sim *s, const char *path) {
    DEV_CHECK(s);
    if (!path) return set_error(MRLBM_ERR_INVALID, "null path");
    const mrlbm::Geometry &g = s->impl->g;
    // payload: the moments exactly as stored (rho - rho0, u, m, or the deviations of the 16-bit
    // forms), so a restart continues bit for bit
    std::vector<unsigned char> payload(size_t(10) * g.n * s->impl->bytes_per_value);
    if (int rc = s->impl->get_raw(payload.data())) return rc;
    CheckpointHeader h{};
    std::memcpy(h.magic, "MRLBMCK1", 8);
    h.version = CHECKPOINT_VERSION;
    h.precision = uint32_t(s->impl->d.precision);
    h.nx = g.nx;
    h.ny = g.ny;
    h.nz = g.nz;
    h.fields = 10;
    h.steps = s->impl->steps;
    h.payload_bytes = payload.size();
    h.checksum = fnv1a(payload.data(), h.payload_bytes);
    h.velocity_ref = s->impl->velocity_reference();
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(MRLBM_ERR_IO, std::string("cannot open ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, fp) == 1 &&
              std::fwrite(payload.data(), 1, h.payload_bytes, fp) == h.payload_bytes;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? MRLBM_OK : set_error(MRLBM_ERR_IO, "checkpoint write failed");
}

int mrlbm_checkpoint_read(mrlbm_sim *s, const char *path) {
    OWN_CHECK(s);
    if (!path) return set_error(MRLBM_ERR_INVALID, "null path");
    const mrlbm::Geometry &g = s->impl->g;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(MRLBM_ERR_IO, std::string("cannot open ") + path);
    CheckpointHeader h{};
    std::vector<unsigned char> payload(size_t(10) * g.n * s->impl->bytes_per_value);
    const bool head = std::fread(&h, sizeof(h), 1, fp) == 1;
    bool ok = head && std::memcmp(h.magic, "MRLBMCK1", 8) == 0 && h.version == CHECKPOINT_VERSION &&
              h.precision == uint32_t(s->impl->d.precision) && h.nx == g.nx && h.ny == g.ny && h.nz == g.nz &&
              h.fields == 10 && h.payload_bytes == payload.size() &&
              h.velocity_ref == s->impl->velocity_reference();
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

// Restart file of the reference (write_checkpoint, src/restart/restart_io.cu:3-74): int iteration;
// the 10 moment arrays of `real` over the 8 x 8 x 4 block-tiled layout (IDX_BLOCK,
// src/core/index.cuh:11-28; src/common/types.cuh:10-12), padded to whole tiles; then the six
// ghost-plane arrays (X_WEST, X_EAST, Y_SOUTH, Y_NORTH, Z_BACK, Z_FRONT, 9 populations per face
// node). The ghost planes hold the collided populations of the stored moments, which this solver
// reconstructs itself, so they are size-checked and skipped.
int mrlbm_checkpoint_import_reference(mrlbm_sim *s, const char *path, int32_t real_bytes) {
    OWN_CHECK(s);
    if (!path || (real_bytes != 4 && real_bytes != 8)) return set_error(MRLBM_ERR_INVALID, "bad path or real size");
    const mrlbm::Geometry &g = s->impl->g;
    const int64_t BX = 8, BY = 8, BZ = 4;
    const int64_t GX = (g.nx + BX - 1) / BX, GY = (g.ny + BY - 1) / BY, GZ = (g.nz + BZ - 1) / BZ;
    const int64_t blocks = GX * GY * GZ, T = blocks * BX * BY * BZ;
    const int64_t halo = 2 * 9 * blocks * (BY * BZ + BX * BZ + BX * BY);
    const uint64_t expect = sizeof(int32_t) + uint64_t(real_bytes) * uint64_t(10 * T + halo);
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(MRLBM_ERR_IO, std::string("cannot open ") + path);
    std::vector<unsigned char> raw;
    int32_t iter = 0;
    bool ok = std::fseek(fp, 0, SEEK_END) == 0;
    const long size = ok ? std::ftell(fp) : -1;
    ok = ok && size >= 0 && uint64_t(size) == expect && std::fseek(fp, 0, SEEK_SET) == 0;
    if (ok) ok = std::fread(&iter, sizeof(iter), 1, fp) == 1 && iter >= 0;
    if (ok) {
        raw.resize(size_t(10 * T) * size_t(real_bytes));
        ok = std::fread(raw.data(), 1, raw.size(), fp) == raw.size();
    }
    std::fclose(fp);
    if (!ok) return set_error(MRLBM_ERR_CORRUPT, "reference restart file: size or iteration mismatch for this grid");
    std::vector<double> all(size_t(10) * size_t(g.n));
    for (int f = 0; f < 10; ++f)
        for (int64_t z = 0; z < g.nz; ++z)
            for (int64_t y = 0; y < g.ny; ++y)
                for (int64_t x = 0; x < g.nx; ++x) {
                    const int64_t t = (x % BX) + BX * ((y % BY) + BY * (z % BZ));
                    const int64_t b = (x / BX) + GX * ((y / BY) + GY * (z / BZ));
                    const size_t src = size_t(f) * size_t(T) + size_t(t + BX * BY * BZ * b);
                    double v;
                    if (real_bytes == 8) {
                        std::memcpy(&v, raw.data() + src * 8, 8);
                    } else {
                        float fv;
                        std::memcpy(&fv, raw.data() + src * 4, 4);
                        v = double(fv);
                    }
                    all[size_t(f) * size_t(g.n) + size_t(x + g.nx * (y + g.ny * z))] = v;
                }
    if (int rc = s->impl->set_all(all)) return rc;
    return s->impl->set_steps(iter);
}

int mrlbm_stats_reset(mrlbm_sim *s) {
    DEV_CHECK(s);
    return s->impl->stats_reset();
}

int mrlbm_get_surface_pressure(mrlbm_sim *s, double *ps_mean, int64_t *samples) {
    DEV_CHECK(s);
    if (!ps_mean) return set_error(MRLBM_ERR_INVALID, "null output");
    return s->impl->get_surface_pressure(ps_mean, samples);
}

// Velocity divisor of the cylindrical components: the caller's, or the case's VEL_NORM
// (src/cases/annulus/constants.cuh:75, src/cases/cylinder/constants.cuh:56).
static int vti_check(mrlbm_sim *s, int32_t layout, double vel_norm, float *norm) {
    if (layout != MRLBM_VTI_CARTESIAN && layout != MRLBM_VTI_ANNULUS) return set_error(MRLBM_ERR_INVALID, "bad VTI layout");
    const mrlbm_case_desc &d = s->impl->d;
    const double v = vel_norm != 0.0 ? vel_norm : (d.kind == MRLBM_CASE_ANNULUS ? d.u_wall_inner[1] : d.u_max);
    if (layout == MRLBM_VTI_ANNULUS && !(std::isfinite(v) && v != 0.0))
        return set_error(MRLBM_ERR_INVALID, "the annulus layout needs a finite nonzero velocity norm");
    *norm = float(v);
    return MRLBM_OK;
}

int mrlbm_vti_points(mrlbm_sim *s, int32_t layout, double vel_norm, float *out4) {
    DEV_CHECK(s);
    if (!out4) return set_error(MRLBM_ERR_INVALID, "null output");
    float norm = 0.0f;
    if (int rc = vti_check(s, layout, vel_norm, &norm)) return rc;
    return s->impl->vti_points(layout, norm, out4);
}

int mrlbm_write_vti(mrlbm_sim *s, const char *dir, int32_t timestep, int32_t layout, double vel_norm) {
    DEV_CHECK(s);
    if (!dir || !*dir || timestep < 0) return set_error(MRLBM_ERR_INVALID, "bad directory or timestep");
    float norm = 0.0f;
    if (int rc = vti_check(s, layout, vel_norm, &norm)) return rc;
    const mrlbm::Geometry &g = s->impl->g;
    std::vector<float> pts;
    try {
        pts.resize(size_t(4) * size_t(g.n));
    } catch (const std::bad_alloc &) {
        return set_error(MRLBM_ERR_NOMEM, "host buffer of the VTI points");
    }
    if (int rc = s->impl->vti_points(layout, norm, pts.data())) return rc;
    static const char *const cartesian[4] = {"rho", "ux", "uy", "uz"};
    static const char *const annulus[4] = {"rho", "uz", "u_theta", "u_radial"};
    return mrlbm::write_vti_file(std::string(dir) + "/" + mrlbm::vti_file_name(timestep), g.nx, g.ny, g.nz,
                                 layout == MRLBM_VTI_CARTESIAN ? cartesian : annulus, pts.data());
}

int mrlbm_write_master_pvd(const char *dir, int32_t max_iter, int32_t save_every) {
    if (!dir || !*dir || max_iter < 0 || save_every <= 0)
        return set_error(MRLBM_ERR_INVALID, "bad directory, step count or save interval");
    return mrlbm::write_pvd_file(std::string(dir) + "/master.pvd", max_iter, save_every);
}

// Text statistics files: "<iter> <Fx> <Fy> <Fz>" to forces.dat and "<iter> <mass>" to
// mass_flux.dat, 16 fixed decimals, one line per step record (write_forces / write_mass_flux,
// src/statistics/stat_header.cuh:16-51).
int mrlbm_write_forces(mrlbm_sim *s, const char *dir, int64_t first, int32_t count) {
    DEV_CHECK(s);
    if (!dir || !*dir || count < 0) return set_error(MRLBM_ERR_INVALID, "bad directory or count");
    std::vector<double> f(size_t(4) * size_t(count));
    if (count > 0)
        if (int rc = s->impl->get_forces(first, count, f.data())) return rc;
    const std::string base(dir);
    FILE *ff = std::fopen((base + "/forces.dat").c_str(), "w");
    if (!ff) return set_error(MRLBM_ERR_IO, "cannot write " + base + "/forces.dat");
    FILE *fm = std::fopen((base + "/mass_flux.dat").c_str(), "w");
    if (!fm) {
        std::fclose(ff);
        return set_error(MRLBM_ERR_IO, "cannot write " + base + "/mass_flux.dat");
    }
    bool ok = true;
    for (int32_t k = 0; k < count && ok; ++k) {
        const double *r = f.data() + 4 * size_t(k);
        const long long it = static_cast<long long>(first + k);
        ok = std::fprintf(ff, "%lld %.16f %.16f %.16f\n", it, r[0], r[1], r[2]) > 0 &&
             std::fprintf(fm, "%lld %.16f\n", it, r[3]) > 0;
    }
    ok = (std::fclose(ff) == 0) && ok;
    ok = (std::fclose(fm) == 0) && ok;
    return ok ? MRLBM_OK : set_error(MRLBM_ERR_IO, "write failed in " + base);
}

int mrlbm_get_plane_means(mrlbm_sim *s, int64_t first, int32_t count, double *out2) {
    DEV_CHECK(s);
    if (!out2 && count > 0) return set_error(MRLBM_ERR_INVALID, "null buffer");
    return s->impl->get_plane_means(first, count, out2);
}

// inlet_density.dat: "<iter> <inlet mean density> <running mean>", 16 fixed decimals, the running
// mean restarting at `first` (write_average_inlet_density, src/statistics/stat_header.cuh:53-72,
// with the update of src/cases/cylinder/cylinder_kernels.cu:172).
int mrlbm_write_inlet_density(mrlbm_sim *s, const char *dir, int64_t first, int32_t count) {
    DEV_CHECK(s);
    if (!dir || !*dir || count < 0) return set_error(MRLBM_ERR_INVALID, "bad directory or count");
    std::vector<double> p(size_t(2) * size_t(count));
    if (count > 0)
        if (int rc = s->impl->get_plane_means(first, count, p.data())) return rc;
    const std::string path = std::string(dir) + "/inlet_density.dat";
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return set_error(MRLBM_ERR_IO, "cannot write " + path);
    double avg = 0.0;
    bool ok = true;
    for (int32_t k = 0; k < count && ok; ++k) {
        const double rho = p[2 * size_t(k)];
        avg += (rho - avg) / double(k + 1);
        ok = std::fprintf(fp, "%lld %.16f %.16f\n", static_cast<long long>(first + k), rho, avg) > 0;
    }
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? MRLBM_OK : set_error(MRLBM_ERR_IO, "write failed: " + path);
}

// pressure.dat: header "theta z Cp", then per inner-wall node (wall-table order) its angle about
// the centre in degrees [0, 360), its z and the time-averaged wall pressure, 12 significant
// digits (write_pressure, src/cases/cylinder/cylinder_kernels.cuh:251-285).
int mrlbm_write_pressure(mrlbm_sim *s, const char *dir) {
    DEV_CHECK(s);
    if (!dir || !*dir) return set_error(MRLBM_ERR_INVALID, "bad directory");
    const mrlbm::Geometry &g = s->impl->g;
    const mrlbm::WallTable &w = g.wall[0];
    std::vector<double> ps(std::max<size_t>(w.node.size(), 1));
    int64_t samples = 0;
    if (int rc = s->impl->get_surface_pressure(ps.data(), &samples)) return rc;
    const std::string path = std::string(dir) + "/pressure.dat";
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return set_error(MRLBM_ERR_IO, "cannot write " + path);
    const double pi = 3.14159265358979323846;
    bool ok = std::fputs("theta z Cp\n", fp) >= 0;
    for (size_t k = 0; k < w.node.size() && ok; ++k) {
        const int64_t i = w.node[k];
        const unsigned x = unsigned(i % g.nx), y = unsigned((i / g.nx) % g.ny), z = unsigned(i / (int64_t(g.nx) * g.ny));
        const double dx = double(x) - w.xc, dy = double(y) - w.yc;
        double theta = std::atan2(dy, dx) * 180.0 / pi;
        if (theta < 0) theta += 360.0;
        ok = std::fprintf(fp, "%.12g %u %.12g\n", theta, z, ps[k]) > 0;
    }
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? MRLBM_OK : set_error(MRLBM_ERR_IO, "write failed: " + path);
}

int mrlbm_sim_info(mrlbm_sim *s, char *buf, int64_t cap, int64_t *len) {
    SIM_CHECK(s);
    if (cap < 0 || (cap > 0 && !buf)) return set_error(MRLBM_ERR_INVALID, "bad buffer");
    const std::string t = mrlbm::sim_info_string(*s->impl);
    if (len) *len = int64_t(t.size());
    if (cap > 0) {
        const size_t k = std::min(t.size(), size_t(cap - 1));
        std::memcpy(buf, t.data(), k);
        buf[k] = '\0';
    }
    return MRLBM_OK;
}

int mrlbm_destroy(mrlbm_sim *s) {
    if (s && s->in_batch) return set_error(MRLBM_ERR_INVALID, "case is owned by its batch");
    delete s;
    return MRLBM_OK;
}

}  // extern "C"
