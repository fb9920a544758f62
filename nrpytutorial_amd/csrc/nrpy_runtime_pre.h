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

// Pass-2 helper of the generated vector-Jacobian products: slot k of the pointwise
// adjoint at the point offset (o0,o1,o2) from (i0,i1,i2), or 0 outside the interior
// (only interior points carry right-hand sides).
#define VJP_G_AT(k, o0, o1, o2)                                                                \
  (((i0) + (o0) >= NGHOSTS && (i0) + (o0) < Nxx0 + NGHOSTS && (i1) + (o1) >= NGHOSTS &&       \
    (i1) + (o1) < Nxx1 + NGHOSTS && (i2) + (o2) >= NGHOSTS && (i2) + (o2) < Nxx2 + NGHOSTS)    \
       ? vjp_G[IDX4(k, (i0) + (o0), (i1) + (o1), (i2) + (o2))]                               \
       : 0.0)

// Launch geometry. Kernels run on a 1-D grid of blocks of shape
// (BX, BY, BZ) tiling the (n0, n1, n2) index box. The hardware hands block b
// to XCD b % 8, so consecutive blocks land on different L2 caches; the remap
// below gives each XCD a contiguous range of tiles (i2 slowest), so stencil
// neighbours in i1 and i2 are mostly resident in the same XCD's L2.
#define NRPY_NUM_XCD 8
#define NRPY_BLOCK_COORDS(n0, n1, n2)                                                          \
  int nrpy_bx, nrpy_by, nrpy_bz;                                                               \
  {                                                                                            \
    const int nbx_ = ((n0) + (int)blockDim.x - 1) / (int)blockDim.x;                           \
    const int nby_ = ((n1) + (int)blockDim.y - 1) / (int)blockDim.y;                           \
    const int nb_ = (int)gridDim.x;                                                            \
    const int b_ = (int)blockIdx.x;                                                            \
    const int q_ = nb_ / NRPY_NUM_XCD, r_ = nb_ % NRPY_NUM_XCD;                                \
    const int x_ = b_ % NRPY_NUM_XCD, l_ = b_ / NRPY_NUM_XCD;                                  \
    const int t_ = (x_ < r_) ? x_ * (q_ + 1) + l_ : r_ * (q_ + 1) + (x_ - r_) * q_ + l_;      \
    nrpy_bx = t_ % nbx_;                                                                       \
    nrpy_by = (t_ / nbx_) % nby_;                                                              \
    nrpy_bz = t_ / (nbx_ * nby_);                                                              \
  }

// Batched launches (nrpy_batch.h): array roles of an instance and index ranges.
#define NRPY_BATCH_ROLE_YN 0
#define NRPY_BATCH_ROLE_KODD 1
#define NRPY_BATCH_ROLE_KEVEN 2
#define NRPY_BATCH_ROLE_RT 3
#define NRPY_BATCH_ROLE_AUXEVOL 4
#define NRPY_BATCH_ROLE_AUX 5
#define NRPY_RANGE_INTERIOR 0
#define NRPY_RANGE_ALL 1

// Per-kernel block shapes, settable at run time (nrpy_set_block_shape) so an
// autotuner can pick them; index by NRPY_KERNEL_* ids defined per bundle.
#define NRPY_MAX_KERNELS 256  /* = loop.MAX_KERNEL_IDS */
static int nrpy_block_shape[NRPY_MAX_KERNELS][3];
static inline dim3 nrpy_block(int k, int dx, int dy, int dz) {
  if (k >= 0 && k < NRPY_MAX_KERNELS && nrpy_block_shape[k][0] > 0)
    return dim3(nrpy_block_shape[k][0], nrpy_block_shape[k][1], nrpy_block_shape[k][2]);
  return dim3(dx, dy, dz);
}
static inline unsigned nrpy_nblocks(dim3 b, int n0, int n1, int n2) {
  return (unsigned)(((n0 + (int)b.x - 1) / (int)b.x) * ((n1 + (int)b.y - 1) / (int)b.y) *
                    ((n2 + (int)b.z - 1) / (int)b.z));
}
// Set the block shape of kernel id k (threads per block must be <= 256 and a
// multiple of 64); (0,0,0) restores the default. Returns 0 or NRPY_ERR_PARAM.
extern "C" int nrpy_set_block_shape(int k, int bx, int by, int bz) {
  if (k < 0 || k >= NRPY_MAX_KERNELS) return NRPY_ERR_PARAM;
  if (bx == 0 && by == 0 && bz == 0) {
    nrpy_block_shape[k][0] = nrpy_block_shape[k][1] = nrpy_block_shape[k][2] = 0;
    return NRPY_OK;
  }
  if (bx < 1 || by < 1 || bz < 1) return NRPY_ERR_PARAM;
  const int t = bx * by * bz;
  if (t > 256 || t % 64 != 0) return NRPY_ERR_PARAM;
  nrpy_block_shape[k][0] = bx;
  nrpy_block_shape[k][1] = by;
  nrpy_block_shape[k][2] = bz;
  return NRPY_OK;
}
