// output.hip - visualisation output and the run summary.
//
// The point arrays of a VTK ImageData file are computed on the device from the stored moments
// (one pass over the moment fields, 4 float32 values per node), so a write moves 16 bytes per
// node over PCIe instead of the moments themselves. The host then writes the file in the
// reference writer's format: an XML header, an <AppendedData encoding="raw"> section of four
// arrays each prefixed by its UInt64 byte count, x fastest (src/output/saveData.cu:42-116,
// 118-230), and the master .pvd collection of those files (src/output/saveData.cu:3-40).
//
// This file is compiled without FMA contraction: the cylindrical velocity components are
// evaluated in the reference writer's float expression order (saveData.cu:180-208), each
// product and sum rounded on its own, so a file written from FP64 moments is byte-identical to
// what the reference writer makes of the same moments.
#include <cinttypes>
#include <cstdio>

#include "sim.h"

namespace mrlbm {
namespace {

template <class R, class S>
__global__ void __launch_bounds__(256) vti_points_kernel(FieldSet<const S> in, StoreScale<R> sc, R rho0, VtiArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    R M[10];
    load_node<R, S>(in, i, sc, rho0, M);  // M[0] = rho0 + stored deviation, in R
    const uint32_t n = a.n;
    const bool solid = (a.solid[i >> 5] >> (i & 31u)) & 1u;
    const float qnan = __builtin_nanf("");
    a.out[i] = float(M[0]);  // RHO_0 + rho, never masked (saveData.cu:105-106, 211-212)
    if (a.layout == MRLBM_VTI_CARTESIAN) {
        // ux, uy, uz with SOLID points as NaN (saveData.cu:107-112)
#pragma unroll
        for (int f = 1; f < 4; ++f) a.out[size_t(f) * n + i] = solid ? qnan : float(M[f]);
        return;
    }
    a.out[n + i] = solid ? qnan : float(M[3]);  // uz (saveData.cu:220-221)
    // u_theta and u_radial about (xc, yc), divided by the velocity norm (saveData.cu:180-208)
    float ut = qnan, ur = qnan;
    if (!solid) {
        const float ux = float(M[1]);
        const float uy = float(M[2]);
        const uint32_t x = i % uint32_t(a.nx);
        const uint32_t y = (i / uint32_t(a.nx)) % uint32_t(a.ny);
        const float dx = float(double(x) - a.xc);
        const float dy = float(double(y) - a.yc);
        const float r2 = dx * dx + dy * dy;
        if (r2 < 1e-12f) {
            ut = 0.0f;
            ur = 0.0f;
        } else {
            const float inv_r = 1.0f / sqrtf(r2);
            const float cos_t = dx * inv_r;
            const float sin_t = dy * inv_r;
            ut = (-ux * sin_t + uy * cos_t) / a.vel_norm;
            ur = (ux * cos_t + uy * sin_t) / a.vel_norm;
        }
    }
    a.out[2 * size_t(n) + i] = ut;
    a.out[3 * size_t(n) + i] = ur;
}

}  // namespace

template <class R, class S>
void launch_vti_points(const FieldSet<const S> &in, const StoreScale<R> &sc, R rho0, const VtiArgs &a,
                       hipStream_t stream) {
    hipLaunchKernelGGL((vti_points_kernel<R, S>), dim3((a.n + 255) / 256), dim3(256), 0, stream, in, sc, rho0, a);
}

template void launch_vti_points<double, double>(const FieldSet<const double> &, const StoreScale<double> &, double,
                                                const VtiArgs &, hipStream_t);
template void launch_vti_points<float, float>(const FieldSet<const float> &, const StoreScale<float> &, float,
                                              const VtiArgs &, hipStream_t);
template void launch_vti_points<float, __hip_bfloat16>(const FieldSet<const __hip_bfloat16> &,
                                                       const StoreScale<float> &, float, const VtiArgs &, hipStream_t);
template void launch_vti_points<float, _Float16>(const FieldSet<const _Float16> &, const StoreScale<float> &, float,
                                                 const VtiArgs &, hipStream_t);

// ---------------------------------------------------------------------------------------------
// host-side formats

namespace {

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

int io_error(const std::string &path) { return set_error(MRLBM_ERR_IO, "cannot write " + path); }

}  // namespace

std::string vti_file_name(int64_t timestep) { return "data" + std::to_string(10000000LL + timestep) + ".vti"; }

int write_vti_file(const std::string &path, int nx, int ny, int nz, const char *const names[4], const float *points) {
    const uint64_t n = uint64_t(nx) * uint64_t(ny) * uint64_t(nz);
    const uint64_t bytes = n * sizeof(float);
    char ext[96];
    std::snprintf(ext, sizeof(ext), "0 %d 0 %d 0 %d", nx - 1, ny - 1, nz - 1);
    std::string head;
    head += "<?xml version=\"1.0\"?>\n";
    head += "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    head += std::string("  <ImageData WholeExtent=\"") + ext + "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    head += std::string("    <Piece Extent=\"") + ext + "\">\n";
    head += "      <PointData>\n";
    for (int k = 0; k < 4; ++k) {
        // each appended block is its UInt64 byte count followed by the values
        const uint64_t offset = uint64_t(k) * (sizeof(uint64_t) + bytes);
        head += std::string("        <DataArray type=\"Float32\" Name=\"") + names[k] +
                "\" format=\"appended\" offset=\"" + std::to_string(offset) + "\"/>\n";
    }
    head += "      </PointData>\n    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n_";
    static const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";

    File file;
    file.f = std::fopen(path.c_str(), "wb");
    if (!file.f) return io_error(path);
    if (std::fwrite(head.data(), 1, head.size(), file.f) != head.size()) return io_error(path);
    for (int k = 0; k < 4; ++k) {
        if (std::fwrite(&bytes, sizeof(bytes), 1, file.f) != 1) return io_error(path);
        if (n && std::fwrite(points + size_t(k) * n, sizeof(float), n, file.f) != n) return io_error(path);
    }
    if (std::fwrite(tail, 1, sizeof(tail) - 1, file.f) != sizeof(tail) - 1) return io_error(path);
    if (std::fclose(file.f) != 0) {
        file.f = nullptr;
        return io_error(path);
    }
    file.f = nullptr;
    return MRLBM_OK;
}

int write_pvd_file(const std::string &path, int32_t max_iter, int32_t save_every) {
    std::string s = "<VTKFile type=\"Collection\" version=\"0.1\">\n\t<Collection>\n";
    for (int64_t it = 0; it <= max_iter; it += save_every)
        s += "\t\t<DataSet timestep=\"" + std::to_string(it / save_every) + "\" file=\"" + vti_file_name(it) +
             "\" />\n";
    s += "\t</Collection>\n</VTKFile>\n";
    File file;
    file.f = std::fopen(path.c_str(), "wb");
    if (!file.f) return io_error(path);
    if (std::fwrite(s.data(), 1, s.size(), file.f) != s.size()) return io_error(path);
    if (std::fclose(file.f) != 0) {
        file.f = nullptr;
        return io_error(path);
    }
    file.f = nullptr;
    return MRLBM_OK;
}

// Run summary text <- getSimInfoString (src/utils/file_utils.cuh:47-164): case, lattice
// parameters and the device-side layout in place of the CUDA launch parameters.
std::string sim_info_string(const SimBase &s) {
    const mrlbm_case_desc &d = s.d;
    const Geometry &g = s.g;
    static const char *kinds[] = {"periodic box", "cylinder", "annulus"};
    static const char *precs[] = {"fp64", "fp32", "fp32 compute, bf16 storage", "fp32 compute, fp16 storage"};
    const double nu = (d.tau - 0.5) / 3.0;
    const double omega = 1.0 / d.tau;
    double u_ref = d.u_max, length = 0.0;
    if (d.kind == MRLBM_CASE_CYLINDER) {
        length = 2.0 * d.r_inner;
    } else if (d.kind == MRLBM_CASE_ANNULUS) {
        u_ref = std::fabs(d.u_wall_inner[1] - d.u_wall_outer[1]);
        length = d.r_outer - d.r_inner;
    }
    std::string o;
    char line[256];
    auto row = [&](const char *label, const char *fmt, auto... v) {
        char val[160];
        std::snprintf(val, sizeof(val), fmt, v...);
        std::snprintf(line, sizeof(line), "%-25s : %s\n", label, val);
        o += line;
    };
    o += "========================= SIMULATION INFORMATION =========================\n";
    row("Case", "%s", d.kind >= 0 && d.kind <= 2 ? kinds[d.kind] : "?");
    row("Velocity set", "%s", "D3Q27");
    if (length > 0.0 && nu > 0.0) row("Re", "%.6f", u_ref * length / nu);
    row("Precision", "%s", d.precision >= 0 && d.precision <= 3 ? precs[d.precision] : "?");
    row("Grid size", "%d x %d x %d", d.nx, d.ny, d.nz);
    row("Total grid points", "%" PRId64, g.n);
    row("Grid points (million)", "%.6f", double(g.n) * 1e-6);
    row("Periodic x/y/z", "%d %d %d", d.periodic_x, d.periodic_y, d.periodic_z);
    if (d.kind == MRLBM_CASE_CYLINDER || d.kind == MRLBM_CASE_ANNULUS) {
        o += "\n------------------------------ walls ------------------------------\n";
        row(d.kind == MRLBM_CASE_CYLINDER ? "Radius" : "Inner radius", "%.6f", d.r_inner);
        if (d.kind == MRLBM_CASE_ANNULUS) row("Outer radius", "%.6f", d.r_outer);
        row("Center xc", "%.6f", d.xc);
        row("Center yc", "%.6f", d.yc);
        row("Wall nodes inner/outer", "%zu %zu", g.wall[0].node.size(), g.wall[1].node.size());
    }
    o += "\n---------------------------- lattice units ----------------------------\n";
    row("uo", "%.6f", d.u_max);
    row("Mach", "%.6f", d.u_max * std::sqrt(3.0));
    row("Viscosity", "%.6f", nu);
    row("Tau", "%.6f", d.tau);
    row("Omega", "%.6f", omega);
    row("Steps done", "%" PRId64, s.steps);
    o += "\n------------------------------- device -------------------------------\n";
    row("Bytes per moment value", "%d", s.bytes_per_value);
    row("Moment memory (MB)", "%.3f", double(case_device_bytes(d)) / (1024.0 * 1024.0));
    if (s.device_ready) {
        const LaunchConfig lc = s.current_launch();
        row("Fused tile", "%d x %d, z chunk %d", lc.bx, lc.by, lc.lz);
        row("March variant", "%d", lc.variant);
        row("LDS per block (KB)", "%.3f",
            double(fused_lds_bytes(lc.bx, lc.by, s.bytes_per_value == 8 ? 8 : 4, lc.variant)) / 1024.0);
        int dev = 0;
        hipDeviceProp_t prop{};
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            row("Device", "%s (%s)", prop.name, prop.gcnArchName);
            row("Compute units", "%d", prop.multiProcessorCount);
            row("Device memory (MB)", "%.0f", double(prop.totalGlobalMem) / (1024.0 * 1024.0));
        }
    }
    o += "==========================================================================\n";
    return o;
}

}  // namespace mrlbm
