/* BSSN in the rescaled reference-metric formulation (bundles bssn_cartesian,
 * bssn_spherical_sym2). Entry points of
 * Tutorial-Start_to_Finish-BSSNCurvilinear-Two_BHs_Collide.ipynb's C driver;
 * common declarations in nrpy_amd.h. Evolved gridfunctions in alphabetical
 * order (aDD00..22, alpha, betU0..2, cf, hDD00..22, lambdaU0..2, trK,
 * vetU0..2), auxevol RbarDD00..22, aux H. */
#ifndef NRPY_AMD_BSSN_H
#define NRPY_AMD_BSSN_H
#include "nrpy_amd.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Replaces Ricci_eval(rfm_struct*, const paramstruct*, const REAL *in_gfs, REAL *auxevol_gfs)
 * (Two_BHs_Collide.ipynb:486-540; BSSN_quantities.py:260-425): the six
 * RbarDD components into auxevol_gfs at interior points. */
int Ricci_eval(const rfm_struct *rfmstruct, const paramstruct *params, const double *in_gfs, double *auxevol_gfs,
               hipStream_t stream);
/* Ricci_eval runs as Ricci_eval_g0 (RbarDD00, 01, 02) and Ricci_eval_g1 (RbarDD11, 12, 22)
 * (groups listed in <bundle>.json "ricci_groups"). */
int Ricci_eval_g0(const rfm_struct *rfmstruct, const paramstruct *params, const double *in_gfs, double *auxevol_gfs,
                  hipStream_t stream);
int Ricci_eval_g1(const rfm_struct *rfmstruct, const paramstruct *params, const double *in_gfs, double *auxevol_gfs,
                  hipStream_t stream);

/* Replaces rhs_eval(rfm_struct*, const paramstruct*, const REAL *auxevol_gfs,
 * const REAL *in_gfs, REAL *rhs_gfs) (Two_BHs_Collide.ipynb:436-484): BSSN +
 * gauge RHSs (1+log lapse, covariant Gamma-driving shift) with FD4, upwinding
 * on beta^i, RbarDD read from auxevol_gfs. Runs as rhs_eval_g0 (aDD, trK),
 * rhs_eval_g1 (hDD, alpha, cf), rhs_eval_g2 (lambdaU, vetU, betU): one kernel per
 * output group, listed in <bundle>.json "rhs_groups". The _rk4_substep_ variants
 * fuse the RK4 update (see rhs_eval_rk4_substep in nrpy_amd.h). */
int rhs_eval(const rfm_struct *rfmstruct, const paramstruct *params, const double *auxevol_gfs,
             const double *in_gfs, double *rhs_gfs, hipStream_t stream);
int rhs_eval_g0(const rfm_struct *rfmstruct, const paramstruct *params, const double *auxevol_gfs,
                const double *in_gfs, double *rhs_gfs, hipStream_t stream);
int rhs_eval_g1(const rfm_struct *rfmstruct, const paramstruct *params, const double *auxevol_gfs,
                const double *in_gfs, double *rhs_gfs, hipStream_t stream);
int rhs_eval_rk4_substep_g0(const int S, const rfm_struct *rfmstruct, const paramstruct *params,
                            const double *auxevol_gfs, const double *in_gfs, double *y_n, double *rt, double *out,
                            const double dt, hipStream_t stream);
int rhs_eval_rk4_substep_g1(const int S, const rfm_struct *rfmstruct, const paramstruct *params,
                            const double *auxevol_gfs, const double *in_gfs, double *y_n, double *rt, double *out,
                            const double dt, hipStream_t stream);
int rhs_eval_g2(const rfm_struct *rfmstruct, const paramstruct *params, const double *auxevol_gfs,
                const double *in_gfs, double *rhs_gfs, hipStream_t stream);
int rhs_eval_rk4_substep_g2(const int S, const rfm_struct *rfmstruct, const paramstruct *params,
                            const double *auxevol_gfs, const double *in_gfs, double *y_n, double *rt, double *out,
                            const double dt, hipStream_t stream);

/* Replaces Hamiltonian_constraint(rfm_struct*, const paramstruct*, REAL *in_gfs, REAL *aux_gfs)
 * (Two_BHs_Collide.ipynb:542-556; BSSN_constraints.py:49-80). */
int Hamiltonian_constraint(const rfm_struct *rfmstruct, const paramstruct *params, const double *in_gfs,
                           double *aux_gfs, hipStream_t stream);

/* Momentum constraint M^i (BSSN/BSSN_constraints.py:87-126, rescaled by ReU) into aux_gfs
 * MU0..MU2 at interior points: momentum_constraint(params, xx, in_gfs, aux_gfs) of
 * Tutorial-Start_to_Finish-BSSNCurvilinear-Setting_up_Exact_Initial_Data.ipynb (cell 17),
 * with the reference metric read from rfm tables. */
int momentum_constraint(const rfm_struct *rfmstruct, const paramstruct *params, const double *in_gfs,
                        double *aux_gfs, hipStream_t stream);

/* Replaces enforce_detgammabar_constraint(const rfm_struct*, const paramstruct*, REAL *in_gfs)
 * (Enforce_Detgammabar_Constraint.py:60-75), at all points, in place. */
int enforce_detgammabar_constraint(const rfm_struct *rfmstruct, const paramstruct *params, double *in_gfs,
                                   hipStream_t stream);

/* Vector-Jacobian products (no counterpart in the reference; see rhs_eval_vjp in
 * nrpy_amd_scalarwave.h for the conventions). rhs_eval_vjp: cotangent of the 24 RHS
 * outputs -> cotangents of in_gfs (cot_in_gfs) and of RbarDD (cot_auxevol_gfs);
 * vjp_G holds <bundle>.json "vjp_slots" gridfunctions. Ricci_eval_vjp: cotangent of
 * the six RbarDD -> cotangent of in_gfs; vjp_G holds "ricci_vjp_slots" gridfunctions,
 * auxevol_gfs and cot_auxevol_gfs are unused (may be NULL). */
int rhs_eval_vjp(const rfm_struct *rfmstruct, const paramstruct *params, const double *auxevol_gfs,
                 const double *in_gfs, const double *cot_rhs_gfs, double *vjp_G, double *cot_in_gfs,
                 double *cot_auxevol_gfs, hipStream_t stream);
int Ricci_eval_vjp(const rfm_struct *rfmstruct, const paramstruct *params, const double *auxevol_gfs,
                   const double *in_gfs, const double *cot_rhs_gfs, double *vjp_G, double *cot_in_gfs,
                   double *cot_auxevol_gfs, hipStream_t stream);

/* For the adjoint of the BSSN RK4 step and of the H-L2 diagnostic (no counterparts in the
 * reference): the VJP of enforce_detgammabar_constraint (pointwise at all points; writes
 * only the hDD cotangents of cot_in_gfs, vjp_G: "detg_vjp_slots"), the VJP of
 * Hamiltonian_constraint (cotangent of H -> cotangent of in_gfs, "hamiltonian_vjp_slots"),
 * and the derivatives of the Brill-Lindquist initial data with respect to BH1_mass and
 * BH2_mass at all points (device xx arrays). */
int enforce_detgammabar_constraint_vjp(const rfm_struct *rfmstruct, const paramstruct *params,
                                       const double *auxevol_gfs, const double *in_gfs, const double *cot_rhs_gfs,
                                       double *vjp_G, double *cot_in_gfs, double *cot_auxevol_gfs, hipStream_t stream);
int Hamiltonian_constraint_vjp(const rfm_struct *rfmstruct, const paramstruct *params, const double *auxevol_gfs,
                               const double *in_gfs, const double *cot_rhs_gfs, double *vjp_G, double *cot_in_gfs,
                               double *cot_auxevol_gfs, hipStream_t stream);
int initial_data_dmass(const paramstruct *params, double *xx[3], double *dID_dm1, double *dID_dm2,
                       hipStream_t stream);

/* Batched evolutions (no counterpart in the reference): ninst independent grids of this
 * system (own sizes, parameters, boundary maps, rfm tables, dt), one launch per kernel
 * per RK4 substep for all of them. params: host array of ninst paramstructs; rfms, bcs:
 * host arrays of the per-instance handles; arrays: host array of 6*ninst device
 * pointers, instance k's y_n, k_odd, k_even, y_nplus1_running_total, auxevol, aux;
 * dts: host array of time steps. The batch keeps copies of params and handles: the
 * per-instance rfm/bc handles and arrays must outlive it. */
typedef struct nrpy_batch nrpy_batch;
int nrpy_batch_create(const int ninst, const paramstruct *params, const rfm_struct *const *rfms,
                      const bc_struct *const *bcs, double *const *arrays, const double *dts, nrpy_batch **out);
void nrpy_batch_free(nrpy_batch *b);
/* Rebuild the tile tables after nrpy_set_block_shape on a batched kernel id. */
int nrpy_batch_prepare(nrpy_batch *b);
/* Composed ghost fill of array role (0..5) of every instance. */
int nrpy_batch_apply_bcs(const nrpy_batch *b, const int role, const int NUM_GFS, const int8_t *gfs_parity,
                         hipStream_t stream);
/* One RK4 step of every instance (as MoL_step_forward_in_time), and its hipGraph
 * (replay with MoL_graph_launch). */
int MoL_batch_step(nrpy_batch *b, hipStream_t stream);
int MoL_batch_graph_create(nrpy_batch *b, MoL_graph **out);

/* Replaces initial_data(const paramstruct*, REAL *xx[3], REAL *in_gfs)
 * (BSSN_ID_function_string.py:9-39) for Brill-Lindquist data
 * (BrillLindquist.py:36-82); xx are device coordinate arrays. */
int initial_data(const paramstruct *params, double *xx[3], double *in_gfs, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
