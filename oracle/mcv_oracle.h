/*
 * mcv_oracle -- CPU restatement of the MCV_SW time step.  TEST INFRASTRUCTURE.
 *
 * This library is the checker for the HIP product path.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product never links it, and nothing here runs on a GPU.
 *
 * It restates the reference's FP64 algorithm with the reference's evaluation
 * order, compiled with strict IEEE semantics (no contraction), so that it
 * reproduces the reference Fortran bit for bit (pinned against the reference
 * build's vectors in tests/golden).
 *
 * Array layout everywhere: the reference's column-major (i, j, panel) with the
 * halo included, i and j in [1-H, nPV+H], panel in [1, 6] (stat_mod.F90:7-45).
 * 2x2 matrices are stored as 4 consecutive doubles per point in Fortran order
 * (m11, m21, m12, m22), as mesh_mod.F90:20-23 allocates them (2,2,i,j,p).
 */
#ifndef MCV_ORACLE_H
#define MCV_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh geometry of mesh_mod.F90:132-163 with projection_mod.F90:8-322.
 * Outputs are (nh*nh*6) doubles each, matrices 4x that. */
int orc_mesh_build(int dof, int n, double *xP, double *yP, double *lonP,
                   double *latP, double *sqrtG, double *G, double *IG,
                   double *A, double *IA, double *f);

/* ghost_mod.f90:21-79 (initGhost): coef and 1-based iref, (nPV, H) each. */
int orc_ghost_build(int dof, int n, const double *xP, const double *yP,
                    double *coef, int32_t *iref);

/* One time step (RK4 or RK3_TVD, time_scheme_mod.F90:13-142) applied in place
 * to the state arrays.  scheme: 4 = RK4, 3 = RK3_TVD.  nsteps steps. */
int orc_run(int dof, int n, int scheme, double dt, double dx,
            const double *sqrtG, const double *G, const double *IG,
            const double *A, const double *IA, const double *f,
            const double *phi_s, const double *coef, const int32_t *iref,
            double *phi, double *phiG, double *u, double *v,
            double *contraU, double *contraV, double *zonal, double *merid,
            int nsteps);

/* spatial_operators_mod.F90:12-133 alone: tendencies on (nPV, nPV, 6). */
int orc_tendency(int dof, int n, double dx, const double *sqrtG,
                 const double *IG, const double *f, const double *phi_s,
                 const double *phi, const double *phiG, const double *u,
                 const double *v, const double *contraU, const double *contraV,
                 double *t_phiG, double *t_u, double *t_v);

/* time_scheme_mod.F90:125-142 alone (correct_bdy_ghost), in place. */
int orc_boundary(int dof, int n, double dx, const double *G, const double *IG,
                 const double *A, const double *IA, const double *coef,
                 const int32_t *iref, double *phi, double *u, double *v,
                 double *contraU, double *contraV, double *zonal,
                 double *merid);

/* diag_mod.F90:10-67: total mass and energy. */
int orc_diag(int dof, int n, const double *sqrtG, const double *phi_s,
             const double *phi, const double *phiG, const double *u,
             const double *v, const double *contraU, const double *contraV,
             double *mass, double *energy);

#ifdef __cplusplus
}
#endif
#endif
