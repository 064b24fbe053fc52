/*
 * mrlbm.h - C ABI of the MI355X-native D3Q27 moment-representation regularized LBM solver.
 *
 * Every entry point is plain C: pointers, sizes and status codes, no C++ or torch types.
 * The library owns all device memory; caller buffers are read or written during the call
 * and never retained. All functions return an int status (MRLBM_OK == 0); on failure
 * mrlbm_last_error() returns a thread-local description of the last error.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   mrlbm_case_desc          <- compile-time case constants: MRLBM_3D/src/config.h:4-13,
 *                               src/cases/cylinder/constants.cuh:8-54,
 *                               src/cases/annulus/constants.cuh:8-81
 *   mrlbm_create             <- setup_case + initialize_simulation (allocation, node types,
 *                               case_module.initialize, constant upload, tau/Mach guards):
 *                               src/solve_mrlbm.cu:3-24, src/simulation.cu:13-35,
 *                               src/solver/initializeLBM.cuh:8-80
 *   mrlbm_init               <- gpu_initialize_Moments_GhostInterface (K9):
 *                               src/solver/initializeLBM.cu:39-121
 *   mrlbm_step               <- one iteration of the main loop: streaming, boundary_treatment,
 *                               post_streaming, collision, post_collision:
 *                               src/main.cu:21-35, src/simulation.cu:81-113
 *   mrlbm_get_forces         <- d_TotalFx/Fy/Fz/m read-back of compute_force_mass_kernel (K5):
 *                               src/statistics/forces.cu:25-133,
 *                               src/cases/cylinder/cylinder_kernels.cuh:578-588
 *   mrlbm_get_moments/_set   <- copyMomentsDeviceToHost / HostToDevice:
 *                               src/solve_mrlbm.cuh:82-147
 *   mrlbm_get_node_types     <- nodeVar::nodeType (type << 20 | index):
 *                               src/common/nodeTypeMap.cuh:82-98
 *   mrlbm_get_wall_table     <- boundaryVar (boundaryList, unit normals, delta, masks):
 *                               src/cases/cylinder/cylinder_structs.h:5-21
 *   mrlbm_get_bcfluid_table  <- bcfluidList + h_incomingMask_bcfluid/h_outgoingMask_bcfluid:
 *                               src/cases/cylinder/cylinder_helpers.cuh:230-299
 *   mrlbm_checkpoint_write/read <- write_checkpoint / read_checkpoint:
 *                               src/restart/restart_io.cu:3-92
 *   mrlbm_destroy            <- finalize_simulation + free_case: src/simulation.cu:161-172
 *   mrlbm_last_error         <- checkCudaErrors / exit(EXIT_FAILURE): src/utils/cudaHelpers.cuh:10-28
 */
#ifndef MRLBM_H
#define MRLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRLBM_ABI_VERSION 2

/* ---- status codes (replace exit(EXIT_FAILURE) in the reference) ---- */
enum mrlbm_status {
    MRLBM_OK = 0,
    MRLBM_ERR_INVALID = 1,     /* bad argument or descriptor field */
    MRLBM_ERR_TAU = 2,         /* tau outside (0.5, 2.5): check_tau, initializeLBM.cuh:8-15 */
    MRLBM_ERR_MACH = 3,        /* Ma = u/cs > 0.33: check_mach, initializeLBM.cuh:17-26 */
    MRLBM_ERR_GEOMETRY = 4,    /* wall or tag overflow, empty wall list */
    MRLBM_ERR_DEVICE = 5,      /* HIP runtime error */
    MRLBM_ERR_NOMEM = 6,       /* allocation failed */
    MRLBM_ERR_IO = 7,          /* checkpoint file missing or unreadable */
    MRLBM_ERR_CORRUPT = 8,     /* checkpoint header, size or checksum mismatch */
    MRLBM_ERR_NONFINITE = 9,   /* NaN/Inf watchdog tripped (see mrlbm_step) */
    MRLBM_ERR_UNSUPPORTED = 10 /* feature not available for this case/precision */
};

/* ---- case kinds: the reference's CASE_ID (config.h:4-8) as a runtime value ---- */
enum mrlbm_case_kind {
    MRLBM_CASE_PERIODIC_BOX = 0, /* all nodes BULK, periodic in x, y, z */
    MRLBM_CASE_CYLINDER = 1,     /* channel with one circular cylinder (setup_cylinder_case) */
    MRLBM_CASE_ANNULUS = 2       /* two concentric walls (setup_annulus_case) */
};

/* ---- precision: storage type of the 10 moments / arithmetic type of the update ---- */
enum mrlbm_precision {
    MRLBM_FP64 = 0,      /* store f64, compute f64 */
    MRLBM_FP32 = 1,      /* store f32, compute f32 (reference default: types.cuh:5) */
    MRLBM_FP32_BF16 = 2, /* store bf16 deviations, compute f32 */
    MRLBM_FP32_FP16 = 3  /* store fp16 scaled deviations, compute f32 */
};

enum mrlbm_inlet_profile {
    MRLBM_INLET_UNIFORM = 0,  /* u_x = u_max (reference: boundaries.cuh:217) */
    MRLBM_INLET_PARABOLIC = 1 /* u_x(y) = 4 u_max y (H - y) / H^2, H = ny - 1 */
};

/* ---- moment fields, in the argument order of pop_reconstruction (reconstruction.cuh:5-10) ---- */
enum mrlbm_field {
    MRLBM_F_RHO = 0, /* stored as rho - rho0 (mlbm.cu:154) */
    MRLBM_F_UX = 1,
    MRLBM_F_UY = 2,
    MRLBM_F_UZ = 3,
    MRLBM_F_MXX = 4,
    MRLBM_F_MYY = 5,
    MRLBM_F_MZZ = 6,
    MRLBM_F_MXY = 7,
    MRLBM_F_MXZ = 8,
    MRLBM_F_MYZ = 9,
    MRLBM_NUM_FIELDS = 10
};

/* compat_flags bits */
#define MRLBM_COMPAT_INIT_M_ZERO 0x1u /* initial m = 0, as the reference's K9 computes (it reads
                                          the constant tables before they are uploaded) */
#define MRLBM_COMPAT_MASS_STRONG 0x2u /* curved walls close with MassBC::Strong,
                                          numerical_solution_strong_rotated (numerical_solution.cu:124-279);
                                          default MassBC::Equilibrium (constants.cuh:17) */
#define MRLBM_COMPAT_BCF_MASS_STRONG 0x4u /* BCFLUID density from the Strong quadratic
                                              (BCF_MASS_CONSERV, boundaries.cuh:894-904) */

/* mrlbm_step flags */
#define MRLBM_STEP_FORCES 0x1u /* accumulate per-step Fx, Fy, Fz, mass flux (K5 window) */
#define MRLBM_STEP_STATS 0x2u  /* add this step's inner-wall pressure to the FP64 statistics */

/*
 * Case descriptor. Lengths are lattice units; x is the fastest index of every host array
 * (index = x + nx * (y + ny * z)).
 */
typedef struct mrlbm_case_desc {
    int32_t kind;      /* enum mrlbm_case_kind */
    int32_t precision; /* enum mrlbm_precision */
    int32_t nx, ny, nz;
    int32_t periodic_x, periodic_y, periodic_z; /* X/Y/Z_PERIODIC (constants.cuh:8-10) */
    double tau;                                  /* relaxation time; omega = 1 / tau */
    double rho0;                                 /* RHO_0 */
    double u_max;          /* inlet velocity scale U_MAX (cylinder); used for the Mach guard */
    int32_t inlet_profile; /* enum mrlbm_inlet_profile */
    uint32_t compat_flags; /* MRLBM_COMPAT_* */
    double inlet_noise;    /* relative seeded perturbation of the inlet u_x */
    double xc, yc;         /* wall centre (XC, YC) */
    double r_inner;        /* cylinder radius (D/2) or annulus inner radius R_IN */
    double r_outer;        /* annulus outer radius R_OUT (unused for the cylinder) */
    double u_wall_inner[3]; /* wall-frame velocity (U'_x normal, U'_y tangential, U'_z) */
    double u_wall_outer[3];
    double init_noise; /* amplitude a of the uniform initial velocity noise U(-a, a) */
    uint64_t seed;     /* seed of the counter-based initial-state hash */
} mrlbm_case_desc;

/* Summary of the geometry built by mrlbm_create. */
typedef struct mrlbm_info {
    int64_t num_nodes;         /* nx * ny * nz */
    int32_t nb_inner, nb_outer; /* wall nodes of the inner / outer wall list */
    int32_t nb_fluid_inner, nb_fluid_outer; /* BCFLUID nodes */
    int32_t num_donors;        /* nodes read by the curved-wall interpolation stencils */
    int32_t active_tiles, total_tiles; /* reference 8x8x4 tile activity (simulation.cu:174-214) */
    int64_t steps_done;
    int32_t precision;
    int32_t bytes_per_value; /* storage bytes per moment value */
} mrlbm_info;

typedef struct mrlbm_sim mrlbm_sim;

/* Fill desc with the defaults of the given case kind (64x64x16 box / C2 channel / C1 annulus). */
int mrlbm_case_desc_default(int32_t kind, mrlbm_case_desc *desc);

/* Validate desc, classify the geometry, build wall tables, allocate device memory. */
int mrlbm_create(const mrlbm_case_desc *desc, mrlbm_sim **out);

/*
 * Build the geometry of desc (node types, wall and BCFLUID tables, tile activity) without
 * touching a device: only the mrlbm_get_node_types / _wall_table / _bcfluid_table / _info
 * queries and mrlbm_destroy are valid on the result. No reference counterpart: it exposes the
 * host half of case_module.initialize (cases/cylinder/cylinder_kernels.cu:5-49) for callers
 * without a GPU.
 */
int mrlbm_create_geometry_only(const mrlbm_case_desc *desc, mrlbm_sim **out);

/* Seeded initial state: rho = rho0, u = U(-a, a) per component, m = u u (or 0, compat). */
int mrlbm_init(mrlbm_sim *sim);

/*
 * Advance nsteps time steps (asynchronous on the simulation's stream).
 * NaN/Inf watchdog (no reference counterpart: the reference has no state check): the fused pass
 * raises a per-simulation flag when the density it computes at any node is not finite. Once the
 * flag is up, mrlbm_step queues nothing and returns MRLBM_ERR_NONFINITE, and mrlbm_sync and
 * mrlbm_get_forces (which still copy their data) return it too. Setting a new state clears it:
 * mrlbm_init, mrlbm_set_moments, mrlbm_checkpoint_read, mrlbm_checkpoint_import_reference.
 */
int mrlbm_step(mrlbm_sim *sim, int32_t nsteps, uint32_t flags);

/* Wait for all work queued on the simulation's stream (MRLBM_ERR_NONFINITE: see mrlbm_step). */
int mrlbm_sync(mrlbm_sim *sim);

/* Copy one moment field to / from a host array of nx*ny*nz doubles (x fastest). */
int mrlbm_get_moments(mrlbm_sim *sim, int32_t field, double *host);
int mrlbm_set_moments(mrlbm_sim *sim, int32_t field, const double *host);

/*
 * Per-step force record (Fx, Fy, Fz, mass flux) in FP64, for steps taken with
 * MRLBM_STEP_FORCES. Records are kept in a device ring of the last `capacity` steps;
 * `first` counts from the first step of the simulation.
 */
int mrlbm_get_forces(mrlbm_sim *sim, int64_t first, int32_t count, double *out4);

/* Node types in the reference encoding (type << 20 | index), x-fastest. */
int mrlbm_get_node_types(mrlbm_sim *sim, uint32_t *out);

/*
 * Wall table of wall group g (0 inner, 1 outer): global x-fastest node index, unit normal,
 * delta, incoming and outgoing masks. Arrays must hold the count reported by mrlbm_get_info.
 */
int mrlbm_get_wall_table(mrlbm_sim *sim, int32_t group, int64_t *node, double *nx, double *ny,
                         double *delta, uint32_t *in_mask, uint32_t *out_mask);

/* BCFLUID list of group g (x-fastest index and tag) and the 256-entry per-tag mask tables. */
int mrlbm_get_bcfluid_table(mrlbm_sim *sim, int32_t group, int64_t *node, uint32_t *tag,
                            uint32_t *in_mask256, uint32_t *out_mask256);

int mrlbm_get_info(mrlbm_sim *sim, mrlbm_info *info);

/*
 * Caller-supplied geometry <- the case module's initialize hook (struct Case, src/common/
 * globalStructs.cuh:46-62, filled at src/cases/cylinder/cylinder_case.cu:14-26): the node types
 * of initialize_nodeType / boundary_definitions (src/solver/initializeLBM.cuh:58-80,
 * src/cases/cylinder/boundaries.cuh:8-94) in the type << 20 | index encoding
 * (src/common/nodeTypeMap.cuh:82-98), the boundaryVar wall lists of cylinder_initialize
 * (src/cases/cylinder/cylinder_structs.h:5-21, cylinder_kernels.cu:5-49) and the BCFLUID mask
 * tables (setup_bcfluid_masks, src/cases/cylinder/cylinder_helpers.cuh:230-299).
 *
 * One wall group (0 inner, 1 outer): nb wall nodes; node[k] is the x-fastest index of entry k,
 * whose node type must be (NODE_INNER + group) << 20 | k; unit normal (nx, ny) and depth delta of
 * the wall point (xc, yc) + radius * n on the group's circle; incoming and outgoing masks (bit q);
 * fluid_side +1 when the fluid lies outside the circle (cylinder, annulus inner wall), -1 inside
 * (annulus outer wall); wall-frame velocity (normal, tangential, axial). The two groups may be two
 * bodies with their own centres (the wall-pressure statistics follow group 0).
 */
typedef struct mrlbm_wall_desc {
    int32_t nb;
    int32_t reserved;
    const int64_t *node;
    const double *nx;
    const double *ny;
    const double *delta;
    const uint32_t *in_mask;
    const uint32_t *out_mask;
    double xc, yc;
    double radius;
    double fluid_side;
    double u_wall[3];
} mrlbm_wall_desc;

/*
 * Replace the geometry of a simulation (from mrlbm_create or mrlbm_create_geometry_only) with
 * node_types (nx*ny*nz values, x fastest), ngroups wall tables and, when BCFLUID nodes are
 * present, the 256-entry per-tag BCFLUID masks. BCFLUID lists follow the node types in x-fastest
 * order (tag = index field). Node codes and closed forms are those of mrlbm_create for the case
 * kind (faces get closed forms in the cylinder kind only). The grid, descriptor and inlet profile
 * are kept; a device simulation is re-initialised as by mrlbm_init. Inconsistent tables give
 * MRLBM_ERR_GEOMETRY and leave the simulation unchanged; refused for cases owned by a batch.
 */
int mrlbm_set_geometry(mrlbm_sim *sim, const uint32_t *node_types, int32_t ngroups,
                       const mrlbm_wall_desc *walls, const uint32_t *bcf_in_mask256,
                       const uint32_t *bcf_out_mask256);

/*
 * Time `nsteps` steps on device with HIP events (inputs resident in HBM). Writes the wall
 * time in ms to ms_total and, if non-null, the mean duration of one launch of each kernel
 * (fused step, curved wall) measured by per-launch events to ms_kernels[0..1].
 */
int mrlbm_time_steps(mrlbm_sim *sim, int32_t nsteps, uint32_t flags, double *ms_total,
                     double *ms_kernels);

/*
 * Time nsteps steps as mrlbm_step runs them, with the step graph captured and instantiated before
 * the first event (so the timed region holds replays only): nsteps / 16 graph replays of 16 steps,
 * an HIP event after each, then the remaining steps queued one by one. Writes the whole region's
 * time to ms_total, the median over the replays of one replay's time divided by its 16 steps to
 * ms_step_median (if non-null; 0 when nsteps < 16) and the number of replays to replays (if
 * non-null). Replaces the reference's wall-clock MLUPS timer (calculate_mlups,
 * src/utils/file_utils.cuh:211-221), which also counts initialisation and output.
 */
int mrlbm_time_replays(mrlbm_sim *sim, int32_t nsteps, uint32_t flags, double *ms_total,
                       double *ms_step_median, int32_t *replays);

/*
 * The fused pass's attainable memory floor on this device (no reference counterpart): a
 * memory-only twin of the fused kernel -- the same work items, per plane the same loads of each
 * tile and its one-node ring and the same stores of every node, without arithmetic, LDS or
 * barriers -- launched `launches` times after one untimed launch. Writes the mean duration of one
 * launch (ms). The simulation's state is unchanged (the twin copies it into the buffer the next
 * step overwrites).
 */
int mrlbm_time_memory_twin(mrlbm_sim *sim, int32_t launches, double *ms_per_launch);

/*
 * Checkpoint: a header (magic, version, precision, grid, step count, FNV-1a checksum) and the
 * moments exactly as stored on the device, field-major without padding, so a restart continues
 * bit for bit. Reading rejects a file of another precision or grid, trailing bytes or a bad
 * checksum with MRLBM_ERR_CORRUPT, and a missing file with MRLBM_ERR_IO.
 */
int mrlbm_checkpoint_write(mrlbm_sim *sim, const char *path);
int mrlbm_checkpoint_read(mrlbm_sim *sim, const char *path);

/*
 * Import a restart file written by the reference <- read_checkpoint (src/restart/restart_io.cu:
 * 19-36): the iteration count and the 10 moment arrays of `real` (real_bytes 4 or 8) in the
 * reference's 8 x 8 x 4 block-tiled order; the trailing ghost-plane arrays are size-checked and
 * skipped (they are the collided populations of the stored moments). The file must match this
 * case's grid exactly (MRLBM_ERR_CORRUPT otherwise); the step count becomes the iteration.
 */
int mrlbm_checkpoint_import_reference(mrlbm_sim *sim, const char *path, int32_t real_bytes);

int mrlbm_destroy(mrlbm_sim *sim);

/*
 * Time-averaged wall pressure of the inner wall (the cylinder, or the annulus's inner wall) <-
 * compute_surface_pressure + surface_pressure_extrapolation (K8,
 * src/cases/cylinder/cylinder_kernels.cu:238-283, src/boundary/extrapolation_utils.cuh:109-154)
 * and write_pressure (src/cases/cylinder/cylinder_kernels.cuh:251-285). Every step queued with
 * MRLBM_STEP_STATS adds each inner-wall node's extrapolated wall pressure to an FP64 sum;
 * mrlbm_get_surface_pressure returns the means (nb_inner values, wall-table order) and the
 * number of steps averaged; mrlbm_stats_reset starts a new window. The statistics are not part
 * of checkpoints.
 */
int mrlbm_stats_reset(mrlbm_sim *sim);
int mrlbm_get_surface_pressure(mrlbm_sim *sim, double *ps_mean, int64_t *samples);

/*
 * Text statistics files in the reference's formats:
 *   mrlbm_write_forces   <- write_forces / write_mass_flux (src/statistics/stat_header.cuh:16-51):
 *                           <dir>/forces.dat "iter Fx Fy Fz" and <dir>/mass_flux.dat "iter m",
 *                           16 fixed decimals, for the force records first .. first + count - 1
 *   mrlbm_write_pressure <- write_pressure (src/cases/cylinder/cylinder_kernels.cuh:251-285):
 *                           <dir>/pressure.dat "theta z Cp" header, then per inner-wall node its
 *                           angle in degrees [0, 360), z and the time-averaged wall pressure
 * Files are overwritten; errors return MRLBM_ERR_IO.
 */
int mrlbm_write_forces(mrlbm_sim *sim, const char *dir, int64_t first, int32_t count);
int mrlbm_write_pressure(mrlbm_sim *sim, const char *dir);

/*
 * Inlet / outlet plane means of the cylinder channel <- inlet_avg_density (K6,
 * src/cases/cylinder/cylinder_kernels.cu:285-338) and outlet_avg_ux (K7, src/solver/mlbm.cu:18-81):
 * every step queued with MRLBM_STEP_STATS files the mean density over the plane x = 0 and the
 * mean u_x over x = nx - 2 of the state after the step (FP64 sums) in a 65,536-step ring beside
 * the force records. mrlbm_get_plane_means returns count pairs for steps first .. first + count - 1
 * (records of steps queued without the flag are not written); MRLBM_ERR_UNSUPPORTED for other
 * case kinds. mrlbm_write_inlet_density writes <dir>/inlet_density.dat "iter rho_inlet
 * rho_inlet_average" (write_average_inlet_density, src/statistics/stat_header.cuh:53-72) with the
 * running mean taken over the records written.
 */
int mrlbm_get_plane_means(mrlbm_sim *sim, int64_t first, int32_t count, double *out2);
int mrlbm_write_inlet_density(mrlbm_sim *sim, const char *dir, int64_t first, int32_t count);

/*
 * Visualisation output <- write_vti_3d / write_vti_3d_annulus / write_master_pvd
 * (src/output/saveData.cu:3-230, called from src/cases/cylinder/cylinder_kernels.cu:133 and
 * src/simulation.cu:8,133). The point arrays are computed on the device from the current moments:
 *   MRLBM_VTI_CARTESIAN: rho (rho0 + stored deviation), ux, uy, uz
 *   MRLBM_VTI_ANNULUS:   rho, uz, u_theta, u_radial -- the in-plane velocity projected on the
 *                        polar frame about (xc, yc) and divided by vel_norm (0 at the axis)
 * as float32, x fastest, with the velocity arrays NaN at SOLID nodes. vel_norm == 0 selects the
 * case's VEL_NORM (annulus: u_wall_inner[1]; otherwise u_max).
 * mrlbm_vti_points returns the four arrays (4 * num_nodes floats, field-major);
 * mrlbm_write_vti writes <dir>/data<10000000 + timestep>.vti (VTK ImageData, appended raw data
 * with UInt64 size headers); mrlbm_write_master_pvd writes <dir>/master.pvd listing the files of
 * steps 0, save_every, ... <= max_iter. File errors return MRLBM_ERR_IO (the reference skips
 * the file silently).
 */
enum mrlbm_vti_layout {
    MRLBM_VTI_CARTESIAN = 0,
    MRLBM_VTI_ANNULUS = 1
};
int mrlbm_vti_points(mrlbm_sim *sim, int32_t layout, double vel_norm, float *out4);
int mrlbm_write_vti(mrlbm_sim *sim, const char *dir, int32_t timestep, int32_t layout, double vel_norm);
int mrlbm_write_master_pvd(const char *dir, int32_t max_iter, int32_t save_every);

/*
 * Run summary text <- getSimInfoString / writeSimInfo (src/utils/file_utils.cuh:47-185): case,
 * lattice parameters and the device layout. Writes at most cap - 1 characters and a NUL to buf
 * (buf may be NULL when cap is 0) and the full length to *len.
 */
int mrlbm_sim_info(mrlbm_sim *sim, char *buf, int64_t cap, int64_t *len);

/*
 * Batches: many cases of one precision stepped together -- one fused launch over the work items
 * of every case and one wall launch over the wall nodes of every case per step, replayed from
 * captured HIP graphs. No counterpart in the reference, which compiles one case per run
 * (src/solve_mrlbm.cu:3-24, src/main.cu:7-40); each case keeps the semantics of mrlbm_create.
 *
 * mrlbm_batch_create admits cases from the front of descs while their device bytes stay within
 * hbm_budget_bytes (0: free device memory less 1 GiB) and reports the count in *admitted.
 * mrlbm_batch_plan does the same admission on the host only (explicit budget).
 * mrlbm_batch_case returns a handle owned by the batch: get/set moments, forces, tables, info and
 * checkpoint_write work on it; init, step, time_steps, checkpoint_read and destroy are refused.
 * mrlbm_batch_autotune times trial_steps steps of each launch configuration -- tile, march
 * variant and z chunk, then for the fastest of those the cases' assignment to the CUs (blocks
 * per CU of a persistent launch, cases cut into XCD slabs or kept whole, slab axis) -- times the
 * four fastest again as the run steps (graph replays), keeps the fastest and re-initialises
 * every case.
 */
typedef struct mrlbm_batch mrlbm_batch;

typedef struct mrlbm_tune_result {
    int32_t bx, by;         /* fused tile */
    int32_t variant;        /* 1 ring march, 3 / 4 register-window march with 1 / 2 LDS planes */
    int32_t lz;             /* z chunk */
    int32_t candidates;     /* configurations timed */
    int32_t blocks_per_cu;  /* persistent fused launch: blocks resident per CU, each pulling work
                               items from its XCD's queue (0: one block per work item) */
    double ms_per_step;         /* chosen configuration */
    double ms_per_step_default; /* default configuration */
    int32_t packing;        /* cases on the XCDs: 0 every case of 32 planes or more cut into one z
                               slab per XCD, 1 whole cases on the least loaded XCD, largest first */
    int32_t slab_axis;      /* XCD slabs of a cut case: 2 z slabs, 1 y slabs (rows of tiles) */
    int32_t lz_special;     /* z chunk of the tiles on the boundary path (0: lz) */
} mrlbm_tune_result;

int mrlbm_batch_plan(const mrlbm_case_desc *descs, int32_t n, uint64_t hbm_budget_bytes, int32_t *admitted,
                     uint64_t *bytes);
int mrlbm_batch_create(const mrlbm_case_desc *descs, int32_t n, uint64_t hbm_budget_bytes,
                       mrlbm_batch **out, int32_t *admitted);
int mrlbm_batch_init(mrlbm_batch *batch);
int mrlbm_batch_step(mrlbm_batch *batch, int32_t nsteps, uint32_t flags);
int mrlbm_batch_sync(mrlbm_batch *batch);
int mrlbm_batch_num_cases(mrlbm_batch *batch, int32_t *n);
int mrlbm_batch_case(mrlbm_batch *batch, int32_t index, mrlbm_sim **out);
int mrlbm_batch_time_steps(mrlbm_batch *batch, int32_t nsteps, uint32_t flags, double *ms_total);
int mrlbm_batch_autotune(mrlbm_batch *batch, int32_t trial_steps, mrlbm_tune_result *out);
int mrlbm_batch_get_launch(mrlbm_batch *batch, mrlbm_tune_result *out);
int mrlbm_batch_destroy(mrlbm_batch *batch);

const char *mrlbm_last_error(void);
int32_t mrlbm_abi_version(void);

/*
 * Read-back that allocates nothing. No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight monitoring loops, harnesses that
 * forbid hipMalloc), where mrlbm_get_moments allocates a device staging array on every call.
 * Copies field `field` of the nodes first .. first + count - 1 (x-fastest index; field 0 is
 * rho - rho0) as doubles to host, with the values mrlbm_get_moments returns, through a fixed
 * staging array of MRLBM_RANGE_MAX_VALUES doubles (16 MiB) per GPU that the library holds until
 * the process ends. count above MRLBM_RANGE_MAX_VALUES returns MRLBM_ERR_INVALID (read longer
 * spans in several calls). Calls that overlap in time on one GPU share that array and race:
 * serialise them.
 */
#define MRLBM_RANGE_MAX_VALUES (1 << 21)
int mrlbm_get_moments_range(mrlbm_sim *sim, int32_t field, int64_t first, int64_t count, double *host);

#ifdef __cplusplus
}
#endif

#endif /* MRLBM_H */
