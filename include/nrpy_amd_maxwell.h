/* Vacuum Maxwell System II, Cartesian coordinates (bundle maxwell_cartesian).
 * Entry points of Tutorial-Start_to_Finish-Solving_Maxwells_Equations_in_Vacuum-Cartesian.ipynb's
 * C driver; common declarations in nrpy_amd.h. Evolved gridfunctions in
 * alphabetical order: AU0..2, EU0..2, Gamma, psi; aux DivE, G. */
#ifndef NRPY_AMD_MAXWELL_H
#define NRPY_AMD_MAXWELL_H
#include "nrpy_amd.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Replaces rhs_eval(const paramstruct *, const REAL *in_gfs, REAL *rhs_gfs) of the
 * Maxwell Cartesian driver (VacuumMaxwell_Flat_Evol_Cartesian.py:84-112, System II). */
int rhs_eval(const paramstruct *params, const double *in_gfs, double *rhs_gfs, hipStream_t stream);

/* The constraints C = div E (DivE) and G = Gamma - div A at interior points
 * (VacuumMaxwell_Flat_Evol_Cartesian.py:56-58, 93-96). */
int Maxwell_constraints(const paramstruct *params, const double *in_gfs, double *aux_gfs, hipStream_t stream);

/* Toroidal dipole data (Maxwell/InitialData.py:25-113) at all points. */
int initial_data(const paramstruct *params, double *xx[3], double *in_gfs, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
