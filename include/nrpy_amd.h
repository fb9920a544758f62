/* C ABI of the MI355X system bundles (libnrpy_<bundle>.so): entry points common
 * to every bundle.
 *
 * Each bundle is one generated HIP translation unit for one equation system in
 * one coordinate system (ScalarWave, curvilinear ScalarWave, BSSN, ...). Its
 * entry points are the functions the reference's Start-to-Finish C drivers
 * call, with three changes that every declaration below shares:
 *   - array arguments are DEVICE pointers (HBM), except where a comment says
 *     "host"; the caller allocates them, as RK_Allocate_Memory.h does
 *     (MoLtimestepping/C_Code_Generation.py:68-105);
 *   - compute entry points take a trailing hipStream_t and are asynchronous on it;
 *   - they return an int status (NRPY_OK = 0, negative on error) where the
 *     reference returns void and calls exit(1) on bad input.
 *
 * paramstruct is the reference's struct of C parameters
 * (NRPy_param_funcs.py:227-243: every registered parameter, members sorted by
 * name). Its members differ between bundles, so it is opaque here; the layout of
 * each bundle is listed, member by member with defaults, in
 * nrpytutorial_amd/csrc/generated/<bundle>.json ("paramstruct"), which is what
 * the Python host code builds its ctypes mirror from.
 *
 * Per-system kernels (rhs_eval, Ricci_eval, ...) are declared in
 * nrpy_amd_scalarwave.h, nrpy_amd_scalarwave_curvilinear.h, nrpy_amd_maxwell.h and
 * nrpy_amd_bssn.h. The Einstein Toolkit plugin form of the BSSN kernels (bundle
 * bssn_etk) exports only the kernels of nrpy_amd_bssn_etk.h, none of the grid,
 * boundary or time-stepping entry points below.
 */
#ifndef NRPY_AMD_H
#define NRPY_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define NRPY_OK 0
#define NRPY_ERR_NULL -1      /* a required pointer was NULL */
#define NRPY_ERR_SHAPE -2     /* inconsistent grid sizes */
#define NRPY_ERR_LAUNCH -3    /* kernel launch or HIP runtime failure */
#define NRPY_ERR_PARAM -4     /* invalid argument value */
#define NRPY_ERR_ALLOC -5     /* device or host allocation failed */
#define NRPY_ERR_BCMAP -6     /* boundary map: Cartesian round trip disagrees, or bad parity */
#define NRPY_ERR_TOO_LARGE -7 /* request exceeds a fixed capacity */

typedef struct __paramstruct__ paramstruct; /* bundle-specific, see above */
typedef struct __rfmstruct__ rfm_struct;    /* device tables of the precomputed reference metric */
typedef struct bc_struct bc_struct;         /* boundary lists, host and device copies */
typedef struct MoL_graph MoL_graph;         /* one captured RK4 step */

/* Boundary-condition records, identical to BCs_data_structs.h:2-49. */
typedef struct __ghostzone_map__ { short i0, i1, i2; } gz_map;
typedef struct __parity__ { int8_t parity[10]; } parity_condition;
typedef struct __inner_bc__ { gz_map inner_bc_dest_pt; gz_map inner_bc_src_pt; int8_t parity[10]; } inner_bc;
typedef struct __outer_bc__ { gz_map outer_bc_dest_pt; int8_t FACEi0, FACEi1, FACEi2; } outer_bc;

/* ---- grid setup (host) ---------------------------------------------------- */

/* Replaces set_Nxx_dxx_invdx_params__and__xx() emitted by
 * reference_metric.py:1330-1395. EigenCoord=1 sets up the grid of the
 * coordinate system's EigenCoord (used for boundary maps), 0 the chosen one.
 * Fills params->Nxx*, Nxx_plus_2NGHOSTS*, dxx*, invdx* and the host arrays
 * xx[d][0 .. Nxx[d]+2*NGHOSTS) with cell-centred coordinates. */
int set_Nxx_dxx_invdx_params__and__xx(const int EigenCoord, const int Nxx[3], paramstruct *params, double *xx[3]);

/* Replaces xx_to_Cart() of reference_metric.py:1397-1409 (host xx). */
void xx_to_Cart(const paramstruct *params, double *xx[3], const int i0, const int i1, const int i2, double xCart[3]);

/* Replaces find_timestep() of reference_metric.py:1412-1439:
 * CFL_FACTOR * min over interior points and directions of the proper grid
 * spacing (host xx). */
double find_timestep(const paramstruct *params, double *xx[3], const double CFL_FACTOR);

/* Replaces rfm_struct__malloc.h / rfm_struct__define.h
 * (reference_metric.py:771-886): tabulates the precomputed reference-metric
 * functions of the host xx arrays and uploads them; *out is freed with
 * rfm_struct_free. */
int rfm_struct_create(const paramstruct *params, double *xx[3], rfm_struct **out);
void rfm_struct_free(rfm_struct *rfmstruct);

/* ---- boundary conditions -------------------------------------------------- */

/* Replaces set_up__bc_gz_map_and_parity_condns.h:15-113 (host; params and xx
 * of the EigenCoord grid). bc_gz_map and bc_parity_conditions hold
 * Nxx_plus_2NGHOSTS0*1*2 entries. Returns NRPY_ERR_BCMAP where the reference
 * prints an error and exits. */
int set_up__bc_gz_map_and_parity_condns(const paramstruct *params, double *xx[3], gz_map *bc_gz_map,
                                        parity_condition *bc_parity_conditions);

/* Replaces set_bcstruct.h:32-124 (driver_bcstruct.h): builds the per-layer
 * outer/inner lists on the host from the EigenCoord grid, the composed
 * one-pass ghost map, and their device copies. bcstruct_create_host builds the
 * host lists only (no device memory). Freed with bcstruct_free
 * (bcstruct_freemem.h). */
int bcstruct_create(const paramstruct *params, double *xx[3], bc_struct **out);
int bcstruct_create_host(const paramstruct *params, double *xx[3], bc_struct **out);
void bcstruct_free(bc_struct *bc);

/* Inspection of the host lists (no counterpart in the reference, whose struct
 * is public): per-layer counts, one layer's lists, the composed map. */
int bcstruct_get_counts(const bc_struct *bc, int *num_ob_gz_pts, int *num_ib_gz_pts, int *nghost_total,
                        int *nterms_total);
int bcstruct_get_layer(const bc_struct *bc, int which_gz, outer_bc *outer_out, inner_bc *inner_out);
int bcstruct_get_composed(const bc_struct *bc, int *gdest, int *rowptr, int *src, double *coeff, int *negmask);

/* Replaces apply_bcs_curvilinear() of apply_bcs_curvilinear.h:7-64: for every
 * gridfunction, layer by layer, quadratic extrapolation at outer points then
 * parity copies at inner points. gfs_parity is a HOST array of NUM_GFS parity
 * types (0..9); gfs is the device array of NUM_GFS gridfunctions. */
int apply_bcs_curvilinear(const paramstruct *params, const bc_struct *bc, const int NUM_GFS,
                          const int8_t *gfs_parity, double *gfs, hipStream_t stream);

/* Same result in one launch: every ghost point as a signed linear combination
 * of interior points (same values as apply_bcs_curvilinear up to the rounding
 * of the expanded sums; bit-identical for data exactly representable in
 * them). */
int apply_bcs_curvilinear_composed(const paramstruct *params, const bc_struct *bc, const int NUM_GFS,
                                   const int8_t *gfs_parity, double *gfs, hipStream_t stream);

/* ---- method of lines (RK4) ----------------------------------------------- */

/* One RK4 step: MoL_step_forward_in_time() of
 * MoLtimestepping/C_Code_Generation.py:266-329 for the diagonal RK4 schedule
 * (y_n, k_odd, k_even, y_nplus1_running_total), each substep being
 * [Ricci_eval] -> RHS + RK update at interior points -> ghost fill ->
 * [det(gammabar) enforcement], as the generated RK_MoL.h sequences them.
 * auxevol_gfs may be NULL for systems without auxiliary evolved gridfunctions;
 * rfmstruct is required. */
int MoL_step_forward_in_time(const rfm_struct *rfmstruct, const paramstruct *params, const bc_struct *bcstruct,
                             double *y_n_gfs, double *k_odd_gfs, double *k_even_gfs,
                             double *y_nplus1_running_total_gfs, double *auxevol_gfs, const double dt,
                             hipStream_t stream);

/* The same step captured into a hipGraph (pointers and dt fixed at capture),
 * replayed nsteps times per launch. No counterpart in the reference; it is the
 * step_rk4 entry of a HIP-native time loop. */
int MoL_graph_create(const rfm_struct *rfmstruct, const paramstruct *params, const bc_struct *bcstruct,
                     double *y_n_gfs, double *k_odd_gfs, double *k_even_gfs, double *y_nplus1_running_total_gfs,
                     double *auxevol_gfs, const double dt, MoL_graph **out);
int MoL_graph_launch(MoL_graph *g, const int nsteps, hipStream_t stream);
void MoL_graph_free(MoL_graph *g);

/* One substep's RHS evaluation fused with its RK4 update (substep S = 0..3 of
 * the schedule above); writes interior points of `out` and updates rt
 * (y_nplus1_running_total) as the separate RHS + update loops would. */
int rhs_eval_rk4_substep(const int S, const rfm_struct *rfmstruct, const paramstruct *params,
                         const double *auxevol_gfs, const double *in_gfs, double *y_n, double *rt, double *out,
                         const double dt, hipStream_t stream);

/* ---- adjoint -------------------------------------------------------------- */

/* Transpose of the ghost fill (no counterpart in the reference, which has no
 * adjoint): adds to every interior point the cotangents of the ghost points whose
 * values were computed from it (with the fill's coefficients and parity signs), then
 * zeroes the ghost cotangents. cot_gfs is a device array of NUM_GFS gridfunctions. */
int bc_apply_T(const paramstruct *params, const bc_struct *bc, const int NUM_GFS, const int8_t *gfs_parity,
               double *cot_gfs, hipStream_t stream);

/* Device vector helpers used by the adjoint step: out = a*x + b*y (y may be NULL),
 * and out = v, over n doubles. Kernels, not memsets, so they can be graph-captured. */
int nrpy_lincomb(size_t n, double *out, double a, const double *x, double b, const double *y,
                 hipStream_t stream);
int nrpy_fill(size_t n, double *out, double v, hipStream_t stream);
/* RK4 update of the diagonal schedule at every point of n doubles (the update loop of the
 * reference's RK_MoL.h, MoLtimestepping/C_Code_Generation.py:266-329): substeps 0-2
 * turn the RHS in k into the next stage input (and accumulate rt), substep 3 updates y_n. */
int nrpy_rk4_update_all(const int S, size_t n, double *k, double *y_n, double *rt, const double dt,
                        hipStream_t stream);
/* acc = (first ? L : acc) + w and, if kbar is not NULL, kbar = a*L + b*w: one pass of the
 * RK4 step adjoint's stage bookkeeping. */
int nrpy_adjoint_combine(size_t n, double *acc, const double *L, const double *w, double *kbar, double a, double b,
                         int first, hipStream_t stream);

/* ---- diagnostics and tuning ---------------------------------------------- */

/* Sum over interior points of gfs[gf]^2 into *result (device), deterministic
 * two-pass reduction; workspace holds >= 1024 doubles (device). The L2 norm of
 * the Hamiltonian constraint (Two_BHs_Collide.ipynb output) is built on it. */
int gridfunction_sumsq(const paramstruct *params, const double *gfs, int gf, double *workspace, double *result,
                       hipStream_t stream);

/* gridfunction_sumsq without the workspace argument, allocating nothing (no counterpart
 * in the reference): for callers that cannot allocate device memory, such as code under
 * hipGraph stream capture or a serving loop that forbids hipMalloc. The partial sums live
 * in a fixed 8 KiB device array (1024 doubles, enough for any grid: the reduction
 * strides over the points) that the library keeps on each GPU it runs on until the
 * process ends. Calls that overlap in time on one GPU share that array and race: order
 * them. Same result as gridfunction_sumsq, bit for bit. */
int gridfunction_sumsq_noalloc(const paramstruct *params, const double *gfs, int gf, double *result,
                               hipStream_t stream);

/* Block shape (threads per block <= 256, a multiple of 64) for the kernel with
 * id k (ids per bundle in <bundle>.json "kernel_ids"); (0,0,0) restores the
 * default. No counterpart in the reference. */
int nrpy_set_block_shape(int k, int bx, int by, int bz);

#ifdef __cplusplus
}
#endif
#endif /* NRPY_AMD_H */
