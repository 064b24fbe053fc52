// geometry.cpp - host-side case setup (see geometry.h).
#include "geometry.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lattice.h"

namespace mrlbm {

namespace {

uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Octant b holds the seven non-zero lattice directions whose components are each 0 or carry the
// octant's sign; bit 0/1/2 of b selects +x/+y/+z (the sets of nodeClass.cuh:67-75).
bool in_octant(int q, int b) {
    const int s[3] = {(b & 1) ? 1 : -1, (b & 2) ? 1 : -1, (b & 4) ? 1 : -1};
    const int c[3] = {LATTICE.cx[q], LATTICE.cy[q], LATTICE.cz[q]};
    if (c[0] == 0 && c[1] == 0 && c[2] == 0) return false;
    for (int a = 0; a < 3; ++a)
        if (c[a] != 0 && c[a] != s[a]) return false;
    return true;
}

// Face/edge/corner type of a node from the non-periodic domain sides
// (cases/cylinder/boundaries.cuh:8-94).
uint8_t domain_type(const mrlbm_case_desc &d, int x, int y, int z) {
    const bool W = !d.periodic_x && x == 0, E = !d.periodic_x && x == d.nx - 1;
    const bool S = !d.periodic_y && y == 0, N = !d.periodic_y && y == d.ny - 1;
    const bool B = !d.periodic_z && z == 0, F = !d.periodic_z && z == d.nz - 1;
    const int nsides = int(W || E) + int(S || N) + int(B || F);
    if (nsides == 3) {
        if (N) return W ? (F ? RT_NWF : RT_NWB) : (F ? RT_NEF : RT_NEB);
        return W ? (F ? RT_SWF : RT_SWB) : (F ? RT_SEF : RT_SEB);
    }
    if (nsides == 2) {
        if (N && W) return RT_NORTH_WEST;
        if (N && E) return RT_NORTH_EAST;
        if (N && F) return RT_NORTH_FRONT;
        if (N && B) return RT_NORTH_BACK;
        if (S && W) return RT_SOUTH_WEST;
        if (S && E) return RT_SOUTH_EAST;
        if (S && F) return RT_SOUTH_FRONT;
        if (S && B) return RT_SOUTH_BACK;
        if (W && F) return RT_WEST_FRONT;
        if (W && B) return RT_WEST_BACK;
        if (E && F) return RT_EAST_FRONT;
        return RT_EAST_BACK;
    }
    if (N) return RT_NORTH;
    if (S) return RT_SOUTH;
    if (W) return RT_WEST;
    if (E) return RT_EAST;
    if (F) return RT_FRONT;
    if (B) return RT_BACK;
    return RT_BULK;
}

// Node-type grid with a one-node SOLID frame in x and y, so that the neighbour look-ups of the
// classification (load_neighbors, nodeClass.cuh:27-60) need no x/y bounds tests; z wraps when
// periodic and reads SOLID beyond the end planes otherwise.
struct TypeGrid {
    int nx, ny, nz, px, py;
    bool pz;
    std::vector<uint8_t> t;
    TypeGrid(int nx_, int ny_, int nz_, bool pz_) : nx(nx_), ny(ny_), nz(nz_), px(nx_ + 2), py(ny_ + 2), pz(pz_) {
        t.assign(size_t(px) * py * nz, RT_SOLID);
    }
    uint8_t &at(int x, int y, int z) { return t[size_t(x + 1) + size_t(px) * (size_t(y + 1) + size_t(py) * z)]; }
    uint8_t nb(int x, int y, int z, int q) const {
        int zn = z + LATTICE.cz[q];
        if (zn < 0 || zn >= nz) {
            if (!pz) return RT_SOLID;
            zn = zn < 0 ? zn + nz : zn - nz;
        }
        return t[size_t(x + 1 + LATTICE.cx[q]) + size_t(px) * (size_t(y + 1 + LATTICE.cy[q]) + size_t(py) * zn)];
    }
    // bitmask over octants that contain a neighbour of type `want`
    uint32_t octants_with(int x, int y, int z, uint8_t want) const {
        uint32_t bits = 0;
        for (int q = 1; q < Q; ++q)
            if (nb(x, y, z, q) == want)
                for (int b = 0; b < 8; ++b)
                    if (in_octant(q, b)) bits |= 1u << b;
        return bits;
    }
    bool any_nb(int x, int y, int z, uint8_t want) const {
        for (int q = 1; q < Q; ++q)
            if (nb(x, y, z, q) == want) return true;
        return false;
    }
};

int popcount8(uint32_t v) {
    int c = 0;
    for (int b = 0; b < 8; ++b) c += (v >> b) & 1u;
    return c;
}

}  // namespace

double hash_uniform(uint64_t seed, uint64_t counter) {
    return double(mix64(seed ^ mix64(counter)) >> 11) * 0x1.0p-53;
}

// Reference tile activity and the per-node info words of the fused kernel, from the node types,
// indices and wall lists of g.
void finish_geometry(const mrlbm_case_desc &d, Geometry &g) {
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    // reference tile activity: 8x8x4 tiles, active if any non-SOLID node; the kernels index the
    // mask without blockIdx.z (mlbm.cu:87), so the bz = 0 tile decides its whole column
    const int GX = (NX + 7) / 8, GY = (NY + 7) / 8, GZ = (NZ + 3) / 4;
    std::vector<uint8_t> col_active(size_t(GX) * GY, 0);
    for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x)
            for (int z = 0; z < std::min(4, NZ); ++z)
                if (g.ref_type[g.idx(x, y, z)] != RT_SOLID) col_active[size_t(x / 8) + size_t(GX) * (y / 8)] = 1;
    g.total_tiles = GX * GY * GZ;
    g.active_tiles = 0;
    for (uint8_t a : col_active) g.active_tiles += int(a) * GZ;

    // per-node info words for the fused kernel
    g.info.assign(g.n, NC_BULK);
    for (int z = 0; z < NZ; ++z)
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) {
                const int64_t i = g.idx(x, y, z);
                if (!col_active[size_t(x / 8) + size_t(GX) * (y / 8)]) {
                    g.info[i] = NC_FROZEN;
                    continue;
                }
                uint32_t code = NC_KEEP, payload = 0;
                const uint8_t t = g.ref_type[i];
                switch (t) {
                case RT_BULK: code = NC_BULK; break;
                case RT_INNER: code = NC_WALL; payload = g.ref_index[i]; break;
                case RT_OUTER: code = NC_WALL; payload = uint32_t(g.wall[0].node.size()) + g.ref_index[i]; break;
                case RT_BCSOLID_INNER:
                case RT_BCSOLID_OUTER: {
                    const uint32_t tg8 = g.ref_index[i];
                    if (tg8 == 7 || tg8 == 11 || tg8 == 13 || tg8 == 14 || tg8 == 112 || tg8 == 176 || tg8 == 208 ||
                        tg8 == 224) {
                        code = NC_BCS;
                        payload = tg8;
                    }
                    break;
                }
                case RT_BCFLUID_INNER:
                case RT_BCFLUID_OUTER: {
                    const uint32_t tg8 = g.ref_index[i];
                    if (tg8 == 17 || tg8 == 34 || tg8 == 68 || tg8 == 136) {
                        code = NC_BCF;
                        payload = tg8;
                    }
                    break;
                }
                default:
                    // domain faces have closed forms only in the cylinder case; the annulus
                    // module leaves them unchanged (annulus_kernels.cuh:54-57)
                    if (d.kind == MRLBM_CASE_CYLINDER) {
                        switch (t) {
                        case RT_WEST: code = NC_WEST; break;
                        case RT_EAST: code = NC_EAST; break;
                        case RT_NORTH: code = NC_NORTH; break;
                        case RT_SOUTH: code = NC_SOUTH; break;
                        case RT_NORTH_WEST: code = NC_NORTH_WEST; break;
                        case RT_SOUTH_WEST: code = NC_SOUTH_WEST; break;
                        case RT_NORTH_EAST: code = NC_NORTH_EAST; break;
                        case RT_SOUTH_EAST: code = NC_SOUTH_EAST; break;
                        case RT_FRONT: code = NC_FRONT; break;
                        case RT_BACK: code = NC_BACK; break;
                        case RT_NORTH_FRONT: code = NC_NORTH_FRONT; break;
                        case RT_NORTH_BACK: code = NC_NORTH_BACK; break;
                        case RT_SOUTH_FRONT: code = NC_SOUTH_FRONT; break;
                        case RT_SOUTH_BACK: code = NC_SOUTH_BACK; break;
                        case RT_WEST_FRONT: code = NC_WEST_FRONT; break;
                        case RT_WEST_BACK: code = NC_WEST_BACK; break;
                        case RT_EAST_FRONT: code = NC_EAST_FRONT; break;
                        case RT_EAST_BACK: code = NC_EAST_BACK; break;
                        case RT_NWF: code = NC_WEST_CORNER; payload = 0; break;
                        case RT_NWB: code = NC_WEST_CORNER; payload = 1; break;
                        case RT_SWF: code = NC_WEST_CORNER; payload = 2; break;
                        case RT_SWB: code = NC_WEST_CORNER; payload = 3; break;
                        case RT_NEF:
                        case RT_NEB:
                        case RT_SEF:
                        case RT_SEB: code = NC_EAST_CORNER; break;
                        default: break;
                        }
                    }
                    break;
                }
                g.info[i] = code | (payload << INFO_PAYLOAD_SHIFT);
            }
}

int build_geometry(const mrlbm_case_desc &d, Geometry &g, std::string &err) {
    if (d.nx < 3 || d.ny < 3 || d.nz < 1 || int64_t(d.nx) * d.ny * d.nz > (int64_t(1) << 31)) {
        err = "grid size out of range";
        return MRLBM_ERR_INVALID;
    }
    if (d.kind != MRLBM_CASE_PERIODIC_BOX && d.kind != MRLBM_CASE_CYLINDER && d.kind != MRLBM_CASE_ANNULUS) {
        err = "unknown case kind";
        return MRLBM_ERR_INVALID;
    }
    // check_tau / check_mach (solver/initializeLBM.cuh:8-26)
    if (!(d.tau > 0.5 && d.tau < 2.5)) {
        err = "tau out of the stable range (0.5, 2.5)";
        return MRLBM_ERR_TAU;
    }
    if (!(d.rho0 > 0.0)) {
        err = "rho0 must be positive";
        return MRLBM_ERR_INVALID;
    }
    if (d.u_max / std::sqrt(1.0 / 3.0) > 0.33) {
        err = "Mach number above 0.33";
        return MRLBM_ERR_MACH;
    }
    const bool walls = d.kind != MRLBM_CASE_PERIODIC_BOX;
    const bool ann = d.kind == MRLBM_CASE_ANNULUS;
    if (walls && (!(d.r_inner > 0.0) || (ann && !(d.r_outer > d.r_inner)))) {
        err = "bad wall radii";
        return MRLBM_ERR_GEOMETRY;
    }
    g.nx = d.nx;
    g.ny = d.ny;
    g.nz = d.nz;
    g.n = int64_t(d.nx) * d.ny * d.nz;
    const int NX = d.nx, NY = d.ny, NZ = d.nz;

    TypeGrid tg(NX, NY, NZ, d.periodic_z != 0);
    for (int z = 0; z < NZ; ++z)
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) tg.at(x, y, z) = walls ? domain_type(d, x, y, z) : RT_BULK;
    g.ref_index.assign(g.n, 0);
    g.ngroups = walls ? (ann ? 2 : 1) : 0;

    if (walls) {
        const double r_in = d.r_inner, r_out = d.r_outer, r_mid = 0.5 * (r_in + r_out);
        auto radius = [&](int x, int y) {
            const double dx = double(x) - d.xc, dy = double(y) - d.yc;
            return std::sqrt(dx * dx + dy * dy);
        };
        // solid marking: every z plane is the same disc / ring
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) {
                const double r = radius(x, y);
                if (r <= r_in || (ann && r >= r_out))
                    for (int z = 0; z < NZ; ++z) tg.at(x, y, z) = RT_SOLID;
            }
        // classification window: cylinder bounding box +4 nodes, annulus whole domain
        int xa = 0, xb = NX, ya = 0, yb = NY;
        if (!ann) {
            xa = std::max(0, int(std::ceil(d.xc - r_in)) - 4);
            xb = std::min(NX, int(std::floor(d.xc + r_in)) + 5);
            ya = std::max(0, int(std::ceil(d.yc - r_in)) - 4);
            yb = std::min(NY, int(std::floor(d.yc + r_in)) + 5);
        }
        // without z periodicity the wall and BCFLUID scans skip the end planes
        // (nodeClass.cuh:171-172, 220-221)
        const int za = d.periodic_z ? 0 : 1, zb = d.periodic_z ? NZ : NZ - 1;
        // wall nodes: SOLID with BULK neighbours in more than two octants. Converting a node does
        // not change any other node's test (only SOLID centres and BULK neighbours count).
        for (int z = za; z < zb; ++z)
            for (int y = ya; y < yb; ++y)
                for (int x = xa; x < xb; ++x) {
                    if (tg.at(x, y, z) != RT_SOLID) continue;
                    if (popcount8(tg.octants_with(x, y, z, RT_BULK)) > 2)
                        tg.at(x, y, z) = (ann && radius(x, y) >= r_mid) ? RT_OUTER : RT_INNER;
                }
        // BCFLUID nodes: BULK with a SOLID neighbour, tagged by the octants holding SOLID
        std::vector<uint32_t> tag(g.n, 0);
        for (int z = za; z < zb; ++z)
            for (int y = ya; y < yb; ++y)
                for (int x = xa; x < xb; ++x) {
                    if (tg.at(x, y, z) != RT_BULK || !tg.any_nb(x, y, z, RT_SOLID)) continue;
                    tag[g.idx(x, y, z)] = tg.octants_with(x, y, z, RT_SOLID);
                    tg.at(x, y, z) = (ann && radius(x, y) >= r_mid) ? RT_BCFLUID_OUTER : RT_BCFLUID_INNER;
                }
        // BCSOLID nodes (end planes without z periodicity): BACK/FRONT nodes with BULK neighbours
        // in fewer than four octants, tagged by those octants (classify_bcsolid_nodes_triangular,
        // cylinder nodeClass.cuh:254-287, annulus nodeClass.cuh:309-357). BACK/FRONT are not BULK,
        // so converting one leaves the other tests unchanged.
        if (!d.periodic_z)
            for (int z = 0; z < NZ; z += std::max(1, NZ - 1))
                for (int y = ya; y < yb; ++y)
                    for (int x = xa; x < xb; ++x) {
                        const uint8_t t = tg.at(x, y, z);
                        if (t != RT_BACK && t != RT_FRONT) continue;
                        const uint32_t bits = tg.octants_with(x, y, z, RT_BULK);
                        if (popcount8(bits) >= 4) continue;
                        tag[g.idx(x, y, z)] = bits;
                        tg.at(x, y, z) = (ann && radius(x, y) >= r_mid) ? RT_BCSOLID_OUTER : RT_BCSOLID_INNER;
                    }
        g.wall[0].radius = r_in;
        g.wall[1].radius = r_out;
        for (int gi = 0; gi < 2; ++gi) {
            g.wall[gi].xc = d.xc;
            g.wall[gi].yc = d.yc;
        }
        g.wall[1].fluid_dir = -1.0;
        for (int k = 0; k < 3; ++k) {
            g.wall[0].uw[k] = d.u_wall_inner[k];
            g.wall[1].uw[k] = d.u_wall_outer[k];
        }
        // lists in x-fastest order (the reference's z, y, x scan order)
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y)
                for (int x = 0; x < NX; ++x) {
                    const uint8_t t = tg.at(x, y, z);
                    const int64_t i = g.idx(x, y, z);
                    if (t == RT_INNER || t == RT_OUTER) {
                        WallTable &w = g.wall[t == RT_OUTER];
                        g.ref_index[i] = uint32_t(w.node.size());
                        w.node.push_back(i);
                    } else if (t == RT_BCFLUID_INNER || t == RT_BCFLUID_OUTER) {
                        WallTable &w = g.wall[t == RT_BCFLUID_OUTER];
                        g.ref_index[i] = tag[i];
                        w.bcf_node.push_back(i);
                        w.bcf_tag.push_back(tag[i]);
                    } else if (t == RT_BCSOLID_INNER || t == RT_BCSOLID_OUTER) {
                        g.ref_index[i] = tag[i];
                    }
                }
        for (int gi = 0; gi < g.ngroups; ++gi) {
            WallTable &w = g.wall[gi];
            if (w.node.empty()) {
                err = "empty wall list";
                return MRLBM_ERR_GEOMETRY;
            }
            if (w.node.size() >= (size_t(1) << 20)) {
                err = "wall list index overflows the 20-bit node-type field";
                return MRLBM_ERR_GEOMETRY;
            }
            for (int64_t i : w.node) {
                const int x = int(i % NX), y = int((i / NX) % NY), z = int(i / (int64_t(NX) * NY));
                // unit normal from the centre, wall point and depth delta
                // (compute_unit_vectors_boundary_nodes, cylinder_helpers.cuh:10-47)
                const double dx = double(x) - d.xc, dy = double(y) - d.yc;
                const double inv_r = 1.0 / std::sqrt(dx * dx + dy * dy);
                const double nxv = dx * inv_r, nyv = dy * inv_r;
                const double xw = d.xc + w.radius * nxv, yw = d.yc + w.radius * nyv;
                w.nx.push_back(nxv);
                w.ny.push_back(nyv);
                w.delta.push_back(std::fabs((xw - double(x)) * nxv + (yw - double(y)) * nyv));
                // incoming: opposite directions of octants free of SOLID neighbours, plus rest
                // (find_incomings_outgoings, cylinder_helpers.cuh:134-228)
                uint32_t in = 1u;
                const uint32_t solid_oct = tg.octants_with(x, y, z, RT_SOLID);
                for (int b = 0; b < 8; ++b) {
                    if (solid_oct & (1u << b)) continue;
                    for (int q = 1; q < Q; ++q)
                        if (in_octant(q, b)) in |= 1u << LATTICE.opp[q];
                }
                uint32_t out = 1u;
                for (int q = 0; q < Q; ++q)
                    if (in & (1u << LATTICE.opp[q])) out |= 1u << q;
                w.in_mask.push_back(in);
                w.out_mask.push_back(out);
            }
        }
        // per-tag BCFLUID masks from the first node of each list with that tag; groups in order,
        // so the outer list overrides tags the inner list set (setup_bcfluid_masks,
        // cylinder_helpers.cuh:230-299, called inner then outer by annulus_kernels.cu:37-38)
        for (int gi = 0; gi < g.ngroups; ++gi) {
            uint8_t seen[256] = {0};
            const WallTable &w = g.wall[gi];
            for (size_t k = 0; k < w.bcf_node.size(); ++k) {
                const uint32_t t = w.bcf_tag[k];
                if (seen[t]) continue;
                seen[t] = 1;
                const int64_t i = w.bcf_node[k];
                const int x = int(i % NX), y = int((i / NX) % NY), z = int(i / (int64_t(NX) * NY));
                uint32_t in = (1u << Q) - 1u;
                for (int q = 0; q < Q; ++q)
                    if (tg.nb(x, y, z, q) == RT_SOLID) in &= ~(1u << LATTICE.opp[q]);
                uint32_t out = 0;
                for (int q = 0; q < Q; ++q)
                    if (in & (1u << LATTICE.opp[q])) out |= 1u << q;
                g.bcf_in[t] = in;
                g.bcf_out[t] = out;
            }
        }
        // inlet profile per (y, z)
        g.inlet.assign(size_t(NY) * NZ, 0.0);
        const double H = double(NY - 1);
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y) {
                double u = d.u_max;
                if (d.inlet_profile == MRLBM_INLET_PARABOLIC) u = 4.0 * d.u_max * double(y) * (H - double(y)) / (H * H);
                const double r = hash_uniform(d.seed, uint64_t(3 * g.n) + uint64_t(y) + uint64_t(NY) * uint64_t(z));
                g.inlet[size_t(y) + size_t(NY) * z] = u * (1.0 + d.inlet_noise * (2.0 * r - 1.0));
            }
    }

    // copy the type grid out
    g.ref_type.assign(g.n, RT_BULK);
    for (int z = 0; z < NZ; ++z)
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) g.ref_type[g.idx(x, y, z)] = tg.at(x, y, z);

    finish_geometry(d, g);
    return MRLBM_OK;
}

int geometry_from_tables(const mrlbm_case_desc &d, const uint32_t *types, int ngroups, const mrlbm_wall_desc *walls,
                         const uint32_t *bcf_in256, const uint32_t *bcf_out256, Geometry &g, std::string &err) {
    if (!types || ngroups < 0 || ngroups > 2 || (ngroups > 0 && !walls)) {
        err = "node types missing or bad wall group count";
        return MRLBM_ERR_INVALID;
    }
    const int64_t n = g.n;
    std::vector<uint8_t> rt(n);
    std::vector<uint32_t> ri(n);
    int64_t count[2] = {0, 0};
    bool has_bcf = false;
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t t = types[i] >> 20, k = types[i] & 0xFFFFFu;
        if (t > RT_SEB) {
            err = "node " + std::to_string(i) + ": unknown node type " + std::to_string(t);
            return MRLBM_ERR_GEOMETRY;
        }
        rt[i] = uint8_t(t);
        ri[i] = k;
        if (t == RT_INNER || t == RT_OUTER) {
            const int gi = t == RT_OUTER;
            if (gi >= ngroups) {
                err = "wall node of a group without a wall table";
                return MRLBM_ERR_GEOMETRY;
            }
            ++count[gi];
        }
        if ((t == RT_BCFLUID_INNER || t == RT_BCFLUID_OUTER || t == RT_BCSOLID_INNER || t == RT_BCSOLID_OUTER) &&
            k > 255) {
            err = "BCFLUID / BCSOLID tag above 255";
            return MRLBM_ERR_GEOMETRY;
        }
        if (t == RT_BCFLUID_INNER || t == RT_BCFLUID_OUTER) {
            has_bcf = true;
            if ((t == RT_BCFLUID_OUTER) >= ngroups) {
                err = "BCFLUID node of a group without a wall table";
                return MRLBM_ERR_GEOMETRY;
            }
        }
    }
    if (has_bcf && (!bcf_in256 || !bcf_out256)) {
        err = "BCFLUID nodes need the per-tag mask tables";
        return MRLBM_ERR_INVALID;
    }
    WallTable wt[2];
    for (int gi = 0; gi < ngroups; ++gi) {
        const mrlbm_wall_desc &w = walls[gi];
        if (w.nb < 0 || int64_t(w.nb) != count[gi] ||
            (w.nb > 0 && (!w.node || !w.nx || !w.ny || !w.delta || !w.in_mask || !w.out_mask)) ||
            !(w.radius > 0.0) || (w.fluid_side != 1.0 && w.fluid_side != -1.0)) {
            err = "wall table " + std::to_string(gi) + ": entry count, arrays, radius or fluid side do not match";
            return MRLBM_ERR_GEOMETRY;
        }
        WallTable &o = wt[gi];
        for (int32_t k = 0; k < w.nb; ++k) {
            const int64_t i = w.node[k];
            if (i < 0 || i >= n || rt[i] != RT_INNER + gi || ri[i] != uint32_t(k)) {
                err = "wall table " + std::to_string(gi) + " entry " + std::to_string(k) +
                      ": its node's type is not (wall << 20 | entry)";
                return MRLBM_ERR_GEOMETRY;
            }
            o.node.push_back(i);
            o.nx.push_back(w.nx[k]);
            o.ny.push_back(w.ny[k]);
            o.delta.push_back(w.delta[k]);
            o.in_mask.push_back(w.in_mask[k]);
            o.out_mask.push_back(w.out_mask[k]);
        }
        o.radius = w.radius;
        o.xc = w.xc;
        o.yc = w.yc;
        o.fluid_dir = w.fluid_side;
        for (int c = 0; c < 3; ++c) o.uw[c] = w.u_wall[c];
    }
    // BCFLUID lists in x-fastest order
    for (int64_t i = 0; i < n; ++i)
        if (rt[i] == RT_BCFLUID_INNER || rt[i] == RT_BCFLUID_OUTER) {
            WallTable &o = wt[rt[i] == RT_BCFLUID_OUTER];
            o.bcf_node.push_back(i);
            o.bcf_tag.push_back(ri[i]);
        }
    g.ref_type = std::move(rt);
    g.ref_index = std::move(ri);
    g.ngroups = ngroups;
    for (int gi = 0; gi < 2; ++gi) g.wall[gi] = std::move(wt[gi]);
    for (int t = 0; t < 256; ++t) {
        g.bcf_in[t] = bcf_in256 ? bcf_in256[t] : 0u;
        g.bcf_out[t] = bcf_out256 ? bcf_out256[t] : 0u;
    }
    finish_geometry(d, g);
    return MRLBM_OK;
}

}  // namespace mrlbm
