/* Scalar wave equation in a curvilinear reference metric (bundles
 * scalarwave_sinhsphericalv2, scalarwave_sinhcylindricalv2).
 * Entry points of Tutorial-Start_to_Finish-ScalarWaveCurvilinear.ipynb's C
 * driver with rfm precompute; common declarations in nrpy_amd.h. */
#ifndef NRPY_AMD_SCALARWAVE_CURVILINEAR_H
#define NRPY_AMD_SCALARWAVE_CURVILINEAR_H
#include "nrpy_amd.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Replaces rhs_eval(rfm_struct *, const paramstruct *, const REAL *in_gfs, REAL *rhs_gfs)
 * of the curvilinear scalar-wave driver (ScalarWaveCurvilinear_RHSs.py:43-98):
 * vv_rhs = c^2 (ghat^ij uu_{,ij} - ghat^ij Gammahat^k_ij uu_{,k}). */
int rhs_eval(const rfm_struct *rfmstruct, const paramstruct *params, const double *in_gfs, double *rhs_gfs,
             hipStream_t stream);

/* Plane-wave data in the curvilinear coordinates (ScalarWave/InitialData.py:59-88). */
int initial_data(const paramstruct *params, double *xx[3], double *in_gfs, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
