// geometry.h - host-side case setup: node classification, wall and BCFLUID tables.
//
// Runtime form of the reference's case modules (cases/cylinder/nodeClass.cuh,
// cases/cylinder/cylinder_helpers.cuh, cases/annulus/nodeClass.cuh, annulus_helpers.cuh,
// solver/initializeLBM.cuh:58-80, simulation.cu:174-214). Everything is computed once, in FP64,
// on x-fastest arrays (index = x + nx * (y + ny * z)).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mrlbm.h"

namespace mrlbm {

// Reference node types (common/nodeTypeMap.cuh:38-80); kept for mrlbm_get_node_types.
enum RefType : uint8_t {
    RT_BULK = 0, RT_SOLID, RT_INNER, RT_OUTER, RT_BCFLUID_INNER, RT_BCFLUID_OUTER,
    RT_BCSOLID_INNER, RT_BCSOLID_OUTER, RT_NORTH, RT_SOUTH, RT_EAST, RT_WEST, RT_FRONT, RT_BACK,
    RT_NORTH_EAST, RT_NORTH_WEST, RT_NORTH_FRONT, RT_NORTH_BACK, RT_SOUTH_EAST, RT_SOUTH_WEST,
    RT_SOUTH_FRONT, RT_SOUTH_BACK, RT_EAST_FRONT, RT_EAST_BACK, RT_WEST_FRONT, RT_WEST_BACK,
    RT_NWF, RT_NWB, RT_SWF, RT_SWB, RT_NEF, RT_NEB, RT_SEF, RT_SEB
};

// What the fused kernel does at a node (low 5 bits of the per-node info word).
enum NodeCode : uint32_t {
    NC_BULK = 0,    // moments of the pulled populations
    NC_KEEP = 1,    // moments unchanged by streaming, then collided (SOLID in an active tile,
                    // BCFLUID tags without a closed form, faces without a closed form)
    NC_FROZEN = 2,  // never updated (every node of an inactive reference tile)
    NC_WALL = 3,    // curved-wall node: incoming rotated moments, finished by the wall kernel
    NC_BCF = 4,     // BCFLUID closed form (tags 17, 34, 68, 136)
    NC_WEST = 5,    // inlet
    NC_EAST = 6,    // convective outlet
    NC_NORTH = 7,
    NC_SOUTH = 8,
    NC_NORTH_WEST = 9,
    NC_SOUTH_WEST = 10,
    NC_NORTH_EAST = 11,
    NC_SOUTH_EAST = 12,
    // end planes without z periodicity (cylinder faces; BCSOLID in both curved-wall cases)
    NC_FRONT = 13,
    NC_BACK = 14,
    NC_NORTH_FRONT = 15,
    NC_NORTH_BACK = 16,
    NC_SOUTH_FRONT = 17,
    NC_SOUTH_BACK = 18,
    NC_WEST_FRONT = 19,
    NC_WEST_BACK = 20,
    NC_EAST_FRONT = 21,
    NC_EAST_BACK = 22,
    NC_WEST_CORNER = 23,  // payload: 0 NWF, 1 NWB, 2 SWF, 3 SWB
    NC_EAST_CORNER = 24,
    NC_BCS = 25,  // BCSOLID closed form, payload: tag (7, 11, 13, 14, 112, 176, 208, 224)
};
constexpr uint32_t INFO_CODE_MASK = 0x1Fu;
constexpr uint32_t INFO_DONOR = 0x20u;  // node's pre-collision moments are read by a wall stencil
constexpr int INFO_PAYLOAD_SHIFT = 6;   // wall index (NC_WALL) or tag (NC_BCF)

struct WallTable {
    std::vector<int64_t> node;
    std::vector<double> nx, ny, delta;
    std::vector<uint32_t> in_mask, out_mask;
    std::vector<int64_t> bcf_node;
    std::vector<uint32_t> bcf_tag;
    double radius = 0.0;
    double xc = 0.0, yc = 0.0;  // centre of the wall circle
    double fluid_dir = 1.0;  // +1 inner wall (fluid outside), -1 outer wall (curvedLBM.cu:72-76)
    double uw[3] = {0.0, 0.0, 0.0};
};

struct Geometry {
    int nx = 0, ny = 0, nz = 0;
    int64_t n = 0;
    std::vector<uint8_t> ref_type;
    std::vector<uint32_t> ref_index;  // BCFLUID tag or wall list index (type << 20 | index)
    std::vector<uint32_t> info;       // NodeCode | flags | payload, per node
    WallTable wall[2];
    int ngroups = 0;
    uint32_t bcf_in[256] = {0}, bcf_out[256] = {0};
    std::vector<double> inlet;        // prescribed inlet u_x per (y, z)
    int active_tiles = 0, total_tiles = 0;
    int num_donors = 0;
    int64_t idx(int x, int y, int z) const { return x + int64_t(nx) * (y + int64_t(ny) * z); }
};

// Counter-based uniform [0, 1) used for the initial state and the inlet perturbation:
// splitmix64(seed ^ splitmix64(counter)) >> 11, times 2^-53.
double hash_uniform(uint64_t seed, uint64_t counter);

// Validates the descriptor and builds the geometry. Returns an mrlbm_status; err gets a message.
int build_geometry(const mrlbm_case_desc &d, Geometry &g, std::string &err);

// Reference tile activity and the fused kernel's per-node info words from g's node types,
// indices and wall lists.
void finish_geometry(const mrlbm_case_desc &d, Geometry &g);

// Replaces g's node types, wall and BCFLUID tables with caller-supplied ones in the reference's
// encoding (mrlbm_set_geometry); the grid and the inlet profile of g are kept. On error g is
// unchanged.
int geometry_from_tables(const mrlbm_case_desc &d, const uint32_t *types, int ngroups, const mrlbm_wall_desc *walls,
                         const uint32_t *bcf_in256, const uint32_t *bcf_out256, Geometry &g, std::string &err);

}  // namespace mrlbm
