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
// (only interior points carry right-hand sides). The load is unconditional and the
// interior test selects its value: a load under the test compiles to a branch per load,
// each waiting for its data before the next is issued. Offsets reach at most NGHOSTS (G)
// beyond the array along each axis; with the plane index held in [1, N2 - 2] every
// address stays inside slot k's own array (the lowest is N0 (N1 - G) - G >= 0 points in,
// the highest N0 (N1 - G - 1) - G + 1 > 0 points short of its end, as N1 >= 2 G + 1),
// and wherever the clamp moves the index the interior test is false.
__device__ __forceinline__ double nrpy_vjp_select(const bool inside, const double v) { return inside ? v : 0.0; }
#define VJP_G_AT(k, o0, o1, o2)                                                                \
  nrpy_vjp_select((i0) + (o0) >= NGHOSTS && (i0) + (o0) < Nxx0 + NGHOSTS && (i1) + (o1) >= NGHOSTS && \
                      (i1) + (o1) < Nxx1 + NGHOSTS && (i2) + (o2) >= NGHOSTS &&                  \
                      (i2) + (o2) < Nxx2 + NGHOSTS,                                              \
                  vjp_G[IDX4(k, (i0) + (o0), (i1) + (o1),                                        \
                             min(max((i2) + (o2), 1), Nxx_plus_2NGHOSTS2 - 2))])

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

// x-neighbours of a stencil by cross-lane moves (DPP). A row is TX lanes holding
// consecutive i0 of one (i1, i2): 16 lanes (row_shl:k / row_shr:k move a value k lanes
// within each 16-lane row) or the whole 64-lane wave (wave_shl:1 / wave_shr:1, chained).
// The row's first m lanes are live (m = TX, or the grid's width when it is narrower,
// m >= 2 HX). Lane l holds its centre value c = f(i0) and a halo value h: f(i0 - HX) in
// the first HX lanes, f(i0 + HX) in the others. nrpy_xshift<TX, HX, K>(c, h, l, m)
// returns f(i0 + K), |K| <= HX: c of lane l + K, or, where that lane is outside the live
// part of the row, h of the lane HX - |K| lanes back towards the row's middle, which
// loaded exactly that point.
template <int CTRL>
__device__ __forceinline__ double nrpy_dpp(const double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
// lane l receives v of lane l + K (K > 0) or l - |K| (K < 0) of its row
template <int TX, int K>
__device__ __forceinline__ double nrpy_lane_move(const double v) {
  if constexpr (K == 0) {
    return v;
  } else if constexpr (TX == 16) {
    return nrpy_dpp<(K > 0 ? 0x100 + K : 0x110 - K)>(v);          /* row_shl:K / row_shr:-K */
  } else {
    static_assert(TX == 64, "rows are 16 lanes or the whole wave");
    return nrpy_dpp<(K > 0 ? 0x130 : 0x138)>(nrpy_lane_move<64, (K > 0 ? K - 1 : K + 1)>(v));  /* wave_shl:1 / wave_shr:1 */
  }
}
template <int TX, int HX, int K>
__device__ __forceinline__ double nrpy_xshift(const double c, const double h, const int l, const int m) {
  static_assert(K != 0 && (K > 0 ? K : -K) <= HX && 2 * HX <= TX, "x offset out of the halo");
  // both moves run in every lane, before the select: a lane move inside a branch would
  // read the lanes that did not take the branch as 0
  const double in_row = nrpy_lane_move<TX, K>(c);
  const double halo = nrpy_lane_move<TX, (K > 0 ? K - HX : K + HX)>(h);
  return ((K > 0) ? (l + K < m) : (l + K >= 0)) ? in_row : halo;
}
// Lanes that recompute a point another lane owns store into this array instead
// (a pointer select, so the computation stays free of branches).
static __device__ double nrpy_discard[64];
#define NRPY_OWNED(p) (nrpy_own ? (p) : &nrpy_discard[threadIdx.x & 63])

// Tile-relative addressing (codegen/addressing.py). In a kernel whose block covers a
// fixed tile and whose accesses lie within NGHOSTS points of (i0, i1, i2), an index
// splits into a part common to the block (gridfunction offset and the tile's corner,
// moved back by NGHOSTS in y and z and by XB in x: 64-bit, scalar registers) and a 32-bit
// byte offset per lane, which is then never negative. A load becomes (scalar base +
// 32-bit lane offset + constant), one lane offset serving every gridfunction. The
// bundle's parameter check keeps the lane offset below 2^32 bytes. N0 and N1 are the
// allocated row and plane sizes (Nxx_plus_2NGHOSTS0/1, or the toolkit's cctk_ash).
#define NRPY_TILE_ORIGIN_N(XB, N0, N1)                                                         \
  const int nrpy_o0 = i0 - (int)threadIdx.x - (XB), nrpy_o1 = i1 - (int)threadIdx.y - NGHOSTS, \
            nrpy_o2 = i2 - (int)threadIdx.z - NGHOSTS;                                         \
  const uint32_t nrpy_off = 8u * (uint32_t)((i0 - nrpy_o0) + (N0) * ((i1 - nrpy_o1) + (N1) * (i2 - nrpy_o2))); \
  const ptrdiff_t nrpy_org = (ptrdiff_t)nrpy_o0 + (ptrdiff_t)(N0) * ((ptrdiff_t)nrpy_o1 + (ptrdiff_t)(N1) * nrpy_o2)
// index B + (point index inside the gridfunction), B the part common to the block (the
// gridfunction's offset, or 0): B is added first so that the compiler keeps the sum of
// the uniform parts in scalar registers. x offset as the load's constant offset (origin
// XB = 0): one lane offset per (y, z) stencil row, and the compiler may fuse loads of
// x-neighbours
#define NRPY_TILE_XC(B, N0, N1, i, j, k)                                                       \
  ((B) + nrpy_org + (ptrdiff_t)((i) - i0) +                                                    \
   (size_t)((nrpy_off + 8u * (uint32_t)((N0) * (((j) - i1) + (N1) * ((k) - i2)))) >> 3))
// x offset inside the lane offset (origin XB = NGHOSTS): one lane offset per stencil
// point, no loads fused
#define NRPY_TILE_XL(B, N0, N1, i, j, k)                                                       \
  ((B) + nrpy_org +                                                                            \
   (size_t)((nrpy_off + 8u * (uint32_t)(((i) - i0) + (N0) * (((j) - i1) + (N1) * ((k) - i2)))) >> 3))
#define NRPY_TILE_ORIGIN(XB) NRPY_TILE_ORIGIN_N(XB, Nxx_plus_2NGHOSTS0, Nxx_plus_2NGHOSTS1)
#define NRPY_GF_OFFSET(g) ((size_t)(g) * ((size_t)Nxx_plus_2NGHOSTS0 * Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2))
#define IDX4_TILE_XC(g, i, j, k) NRPY_TILE_XC(NRPY_GF_OFFSET(g), Nxx_plus_2NGHOSTS0, Nxx_plus_2NGHOSTS1, i, j, k)
#define IDX4_TILE_XL(g, i, j, k) NRPY_TILE_XL(NRPY_GF_OFFSET(g), Nxx_plus_2NGHOSTS0, Nxx_plus_2NGHOSTS1, i, j, k)
// the toolkit's access form: one array per gridfunction, rows and planes of cctk_ash
#define NRPY_ETK_TILE_ORIGIN(XB) NRPY_TILE_ORIGIN_N(XB, cctkGH.ash0, cctkGH.ash1)
#define CCTK_GFINDEX3D_TILE_XC(GH, i, j, k) NRPY_TILE_XC((ptrdiff_t)0, (GH).ash0, (GH).ash1, i, j, k)
#define CCTK_GFINDEX3D_TILE_XL(GH, i, j, k) NRPY_TILE_XL((ptrdiff_t)0, (GH).ash0, (GH).ash1, i, j, k)

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
