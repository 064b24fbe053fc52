/*
 * mrlbm_oracle.h - C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a plain C++ (OpenMP) restatement of the reference's
 * per-node update (sakthi-lbm/MRLBM_3D_CUDA, two-pass order: streaming + node BC, curved wall,
 * incoming force, collision, outgoing force). It exists to check the HIP library and to time a
 * CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library never links, loads or calls it.
 */
#ifndef MRLBM_ORACLE_H
#define MRLBM_ORACLE_H

#include <stdint.h>

#include "../include/mrlbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

/* real_bytes: 8 runs the restatement in double, 4 in float (CPU-baseline timing only). */
int oracle_create(const mrlbm_case_desc *desc, int32_t real_bytes, oracle_sim **out);
int oracle_init(oracle_sim *s);
int oracle_step(oracle_sim *s, int32_t nsteps, uint32_t flags);
int oracle_get_moments(oracle_sim *s, int32_t field, double *host);
int oracle_set_moments(oracle_sim *s, int32_t field, const double *host);
int oracle_get_forces(oracle_sim *s, int64_t first, int32_t count, double *out4);
int oracle_get_plane_means(oracle_sim *s, int64_t first, int32_t count, double *out2);
int oracle_get_node_types(oracle_sim *s, uint32_t *out);
int oracle_get_wall_table(oracle_sim *s, int32_t group, int64_t *node, double *nx, double *ny,
                          double *delta, uint32_t *in_mask, uint32_t *out_mask);
int oracle_get_bcfluid_table(oracle_sim *s, int32_t group, int64_t *node, uint32_t *tag,
                             uint32_t *in_mask256, uint32_t *out_mask256);
int oracle_get_info(oracle_sim *s, mrlbm_info *info);
int oracle_destroy(oracle_sim *s);
int32_t oracle_num_threads(void);

/* Per-node building blocks, exposed for closed-form tests (double precision). */
void oracle_pop_reconstruction(const double *m10, double *f27); /* m10[0] = rho (not rho-1) */
void oracle_eval_moments(const double *f27, double *m10);       /* m10[0] = rho */
void oracle_collide(double omega, double *m10);
/* node BC of the given reference node type on 27 pulled pops; old10 holds the node's old
 * moments (m10[0] = rho) and, for the outlet types, old10_xm1 the x-1 neighbour's. */
int oracle_node_bc(uint32_t type, uint32_t tag, uint32_t in_mask, double omega, double u_in,
                   const double *f27, const double *old10, const double *old10_xm1,
                   double *out10);
void oracle_incoming_rotated(double nx, double ny, uint32_t in_mask, const double *f27,
                             double *rho, double *m6);
void oracle_extrapolation_coeffs(double delta, double *c3);
/* instantaneous wall pressure of every inner-wall node (nb_inner values) */
int oracle_surface_pressure(oracle_sim *s, double *out);
/* xy8 = (xw, yw, x1, y1, x2, y2, x3, y3); the quadratic wall-pressure extrapolation */
double oracle_surface_pressure_extrapolation(double xc, double yc, const double *xy8, const double *rho3);
/* VTK output (src/output/saveData.cu) as the reference's FP64 build evaluates it: the point arrays
 * of write_vti_3d (layout 0: rho, ux, uy, uz) or write_vti_3d_annulus (layout 1: rho, uz,
 * u_theta, u_radial) from FP64 moments (rho_dev = rho - rho0; solid[i] != 0 marks SOLID nodes),
 * the files themselves, and master.pvd. */
void oracle_vti_points(int32_t layout, int32_t nx, int32_t ny, int32_t nz, double rho0, double xc, double yc,
                       double vel_norm, const double *rho_dev, const double *ux, const double *uy,
                       const double *uz, const uint8_t *solid, float *out4);
int oracle_write_vti(const char *path, int32_t layout, int32_t nx, int32_t ny, int32_t nz, const float *pts4);
int oracle_write_master_pvd(const char *path, int32_t max_iter, int32_t save_every);
void oracle_curved_solution_strong(double nx, double ny, uint32_t in_mask, uint32_t out_mask, double omega,
                                   const double *prime7, double rho_i, double mxy_i, double mxz_i, double *out10);
/* BCFLUID closed form of one node (tags 17/34/68/136), Equilibrium or Strong density */
int oracle_bcfluid(uint32_t tag, uint32_t in_mask, double omega, int32_t strong, const double *f27, double *out10);
void oracle_curved_solution(double nx, double ny, uint32_t in_mask, const double *prime7,
                            double rho_i, double mxy_i, double mxz_i, double *out10);

#ifdef __cplusplus
}
#endif

#endif
