/* Oracle (test infrastructure only): common definitions for the C restatement
 * of the reference's generated code. Layout IDX4S as in
 * Tutorial-Start_to_Finish-BSSNCurvilinear-Two_BHs_Collide.ipynb:904-905. */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include <string.h>
#include <stdint.h>
#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif
#ifndef M_SQRT1_2
#define M_SQRT1_2 0.707106781186547524400844362104849039
#endif
#define IDX4S(g,i,j,k) ((size_t)(i) + (size_t)Nxx_plus_2NGHOSTS0*((size_t)(j) + (size_t)Nxx_plus_2NGHOSTS1*((size_t)(k) + (size_t)Nxx_plus_2NGHOSTS2*(size_t)(g))))
#define IDX3S(i,j,k) ((i) + Nxx_plus_2NGHOSTS0*((j) + Nxx_plus_2NGHOSTS1*(k)))
/* SIMD/SIMD_intrinsics.h:119-124, scalar form: the shift acts as a negative velocity. */
#define UPWIND_ALG(UpwindVecU) ((UpwindVecU) > 0.0 ? 1.0 : 0.0)
