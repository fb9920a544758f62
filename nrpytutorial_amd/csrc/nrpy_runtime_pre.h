// Common definitions for every generated system bundle (one translation unit
// per bundle: ScalarWave, Maxwell, BSSN in a given coordinate system).
//
// Memory layout follows the reference: structure-of-arrays, gridfunction index
// slowest, IDX4S(g,i,j,k) = i + N0*(j + N1*(k + N2*g))
// (Tutorial-Start_to_Finish-BSSNCurvilinear-Two_BHs_Collide.ipynb:904-905).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif
#ifndef M_SQRT1_2
#define M_SQRT1_2 0.707106781186547524400844362104849039
#endif

// Status codes returned by every entry point (the reference calls exit(1)
// on these conditions; set_up__bc_gz_map_and_parity_condns.h:56-64,101-107).
#define NRPY_OK 0
#define NRPY_ERR_NULL -1        // a required pointer was NULL
#define NRPY_ERR_SHAPE -2       // inconsistent grid sizes
#define NRPY_ERR_LAUNCH -3      // kernel launch or HIP runtime failure
#define NRPY_ERR_PARAM -4       // invalid argument value
#define NRPY_ERR_ALLOC -5       // device or host allocation failed
#define NRPY_ERR_BCMAP -6       // boundary map: Cartesian disagreement or bad parity
#define NRPY_ERR_TOO_LARGE -7   // request exceeds a fixed capacity

// Linear index of a point inside one gridfunction; 32-bit is enough for
// 2^31 points per gridfunction.
#define IDX3S(i, j, k) ((i) + Nxx_plus_2NGHOSTS0 * ((j) + Nxx_plus_2NGHOSTS1 * (k)))
// Index into an array of gridfunctions; the gridfunction offset is 64-bit.
#define IDX4(g, i, j, k)                                                                       \
  ((size_t)(g) * ((size_t)Nxx_plus_2NGHOSTS0 * Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2) +    \
   (size_t)IDX3S(i, j, k))
#define IDX4S(g, i, j, k) IDX4(g, i, j, k)
#define IDX4ptS(g, idx) ((size_t)(g) * ((size_t)Nxx_plus_2NGHOSTS0 * Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2) + (size_t)(idx))

#define NRPY_HIP_OK(call) ((call) == hipSuccess)
