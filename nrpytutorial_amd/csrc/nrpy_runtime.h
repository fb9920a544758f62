// Hand-written runtime shared by all generated system bundles.
//
// Included by each generated bundle after it has defined NGHOSTS, paramstruct,
// nrpy_check_params() and the generated host functions
//   EigenCoord_xx_to_Cart(), EigenCoord_Cart_to_xx(), eval_parity_dot_products().
//
// Contents:
//  1. Boundary-condition maps (host). set_up__bc_gz_map_and_parity_condns and
//     set_bcstruct restate CurviBoundaryConditions/boundary_conditions/
//     set_up__bc_gz_map_and_parity_condns.h:15-113 and set_bcstruct.h:32-124:
//     every point goes to Cartesian coordinates and back through the
//     EigenCoord inverse map, the nearest grid index is found with
//     (int)(x + 0.5), and ghost points are sorted by layer and face into an
//     outer (extrapolation) list and an inner (parity copy) list.
//  2. apply_bcs_curvilinear: the reference's sequential order
//     (apply_bcs_curvilinear.h:7-64) as one launch per (layer, face
//     direction) for outer points and one per layer for inner points.
//  3. A composed ghost map: because both boundary operations are linear,
//     every ghost value is a fixed linear combination of interior values with
//     a sign per parity type. The builder expands the sequential order once on
//     the host; apply_bcs_curvilinear_composed then fills every ghost of every
//     gridfunction in a single launch that reads only interior points. The RK
//     substep kernels write interior points only, so this fill runs right
//     after them with no ordering between ghost points.
//  4. Deterministic sum-of-squares reduction over interior points.
//  5. RK4 update macro used by the fused RHS + method-of-lines kernels.
#pragma once
#include <vector>
#include <map>
#include <algorithm>
#include <mutex>

// ---------------------------------------------------------------------------
// Data structures of BCs_data_structs.h:2-49.
typedef struct __ghostzone_map__ {
  short i0, i1, i2;
} gz_map;

typedef struct __parity__ {
  int8_t parity[10];
} parity_condition;

typedef struct __inner_bc__ {
  gz_map inner_bc_dest_pt;
  gz_map inner_bc_src_pt;
  int8_t parity[10];
} inner_bc;

typedef struct __outer_bc__ {
  gz_map outer_bc_dest_pt;
  int8_t FACEi0, FACEi1, FACEi2;
} outer_bc;

static const int8_t MAXFACE = -1;
static const int8_t NUL = +0;
static const int8_t MINFACE = +1;

// Device + host boundary-condition descriptor (opaque to callers).
struct bc_struct {
  int N[3];                        // Nxx_plus_2NGHOSTS of the grid it was built for
  int num_ob_gz_pts[NGHOSTS];
  int num_ib_gz_pts[NGHOSTS];
  std::vector<outer_bc> outer[NGHOSTS];
  std::vector<inner_bc> inner[NGHOSTS];
  int ob_dir_end[NGHOSTS][3];      // outer list: end of the i0-, i1-, i2-face groups
  // device copies of the lists (all layers concatenated)
  int *d_ob_dest = nullptr;
  int *d_ob_stride = nullptr;      // FACE . (1, N0, N0*N1)
  int8_t *d_ob_face = nullptr;     // FACE, 3 per outer point (Sommerfeld radiation condition)
  int *d_ib_dest = nullptr;
  int *d_ib_src = nullptr;
  int8_t *d_ib_parity = nullptr;   // 10 per inner point
  int ob_off[NGHOSTS + 1];
  int ib_off[NGHOSTS + 1];
  // the i0-face points of each layer whose (i1, i2) is not interior (the Sommerfeld
  // condition takes the interior rows' i0-face points in one tiled launch)
  int *d_ob_xedge_dest = nullptr;
  int8_t *d_ob_xedge_face = nullptr;
  int ob_xedge_off[NGHOSTS + 1];
  // composed ghost map (CSR over ghost points)
  int nghost = 0, nterms = 0;
  int *d_gdest = nullptr;
  int *d_rowptr = nullptr;
  int *d_src = nullptr;
  double *d_coeff = nullptr;
  int *d_negmask = nullptr;        // bit t set: parity type t contributes with a minus sign
  std::vector<int> h_gdest, h_rowptr, h_src, h_negmask;
  std::vector<double> h_coeff;
  // transpose of the composed map (CSR over interior source points), for bc_apply_T
  int ntsrc = 0;
  int *d_tsrc = nullptr;
  int *d_trowptr = nullptr;
  int *d_tghost = nullptr;
  double *d_tcoeff = nullptr;
  int *d_tnegmask = nullptr;
  std::vector<int> h_tsrc, h_trowptr, h_tghost, h_tnegmask;
  std::vector<double> h_tcoeff;
  // right-hand-side scratch of the fused Sommerfeld RK4 step (nrpy_sommerfeld_scratch)
  double *d_som_rhs = nullptr;
  size_t som_rhs_n = 0;
};

struct nrpy_gf_parity_args {
  int8_t p[64];
};

// ---------------------------------------------------------------------------
// 1. Host map construction.

// Restates set_up__bc_gz_map_and_parity_condns.h:15-113. Returns NRPY_OK or
// NRPY_ERR_BCMAP where the reference would print an error and exit(1).
extern "C" int set_up__bc_gz_map_and_parity_condns(const paramstruct *params, double *xx[3], gz_map *bc_gz_map,
                                                   parity_condition *bc_parity_conditions) {
  if (!params || !xx || !xx[0] || !xx[1] || !xx[2] || !bc_gz_map || !bc_parity_conditions) return NRPY_ERR_NULL;
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;
  const double dxx0 = params->dxx0, dxx1 = params->dxx1, dxx2 = params->dxx2;
  const int Nt[3] = {Nxx_plus_2NGHOSTS0, Nxx_plus_2NGHOSTS1, Nxx_plus_2NGHOSTS2};
  const double xxmin[3] = {xx[0][0] - ((double)(0 - NGHOSTS) + (1.0 / 2.0)) * dxx0,
                           xx[1][0] - ((double)(0 - NGHOSTS) + (1.0 / 2.0)) * dxx1,
                           xx[2][0] - ((double)(0 - NGHOSTS) + (1.0 / 2.0)) * dxx2};
  for (int i2 = 0; i2 < Nxx_plus_2NGHOSTS2; i2++)
    for (int i1 = 0; i1 < Nxx_plus_2NGHOSTS1; i1++)
      for (int i0 = 0; i0 < Nxx_plus_2NGHOSTS0; i0++) {
        double xCart[3];
        EigenCoord_xx_to_Cart(params, xx, i0, i1, i2, xCart);
        double c2x[3];
        EigenCoord_Cart_to_xx(params, xCart, c2x);
        const int ii[3] = {
            (int)((c2x[0] - xxmin[0] - (1.0 / 2.0) * dxx0 + ((double)NGHOSTS) * dxx0) / dxx0 + 0.5),
            (int)((c2x[1] - xxmin[1] - (1.0 / 2.0) * dxx1 + ((double)NGHOSTS) * dxx1) / dxx1 + 0.5),
            (int)((c2x[2] - xxmin[2] - (1.0 / 2.0) * dxx2 + ((double)NGHOSTS) * dxx2) / dxx2 + 0.5)};
        for (int d = 0; d < 3; d++)
          if (ii[d] < 0 || ii[d] >= Nt[d]) return NRPY_ERR_BCMAP;
        double xCart_in[3];
        EigenCoord_xx_to_Cart(params, xx, ii[0], ii[1], ii[2], xCart_in);
        for (int d = 0; d < 3; d++)
          if (fabs(xCart[d] - xCart_in[d]) > 1e-8) return NRPY_ERR_BCMAP;
        const int idx3 = IDX3S(i0, i1, i2);
        if (ii[0] == i0 && ii[1] == i1 && ii[2] == i2) {
          bc_gz_map[idx3].i0 = -1;
          bc_gz_map[idx3].i1 = -1;
          bc_gz_map[idx3].i2 = -1;
          for (int p = 0; p < 10; p++) bc_parity_conditions[idx3].parity[p] = 1;
        } else {
          bc_gz_map[idx3].i0 = (short)ii[0];
          bc_gz_map[idx3].i1 = (short)ii[1];
          bc_gz_map[idx3].i2 = (short)ii[2];
          double par10[10];
          eval_parity_dot_products(params, par10, xx[0][i0], xx[1][i1], xx[2][i2], xx[0][ii[0]], xx[1][ii[1]],
                                   xx[2][ii[2]]);
          for (int p = 0; p < 10; p++) {
            const double v = par10[p];
            if ((v > 0 && fabs(v - 1.0) > 1e-8) || (v <= 0 && fabs(v + 1.0) > 1e-8)) return NRPY_ERR_BCMAP;
            bc_parity_conditions[idx3].parity[p] = (v < 0.0) ? -1 : +1;
          }
        }
      }
  return NRPY_OK;
}

// Restates set_bcstruct.h:32-124: layer by layer, six faces in fixed order,
// the face extents growing after each face.
static void nrpy_set_bcstruct(const paramstruct *params, const gz_map *bc_gz_map,
                              const parity_condition *bc_parity_conditions, bc_struct *bc) {
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;
  int imin[3] = {NGHOSTS, NGHOSTS, NGHOSTS};
  int imax[3] = {Nxx_plus_2NGHOSTS0 - NGHOSTS, Nxx_plus_2NGHOSTS1 - NGHOSTS, Nxx_plus_2NGHOSTS2 - NGHOSTS};
  for (int which_gz = 0; which_gz < NGHOSTS; which_gz++) {
    std::vector<outer_bc> &ob = bc->outer[which_gz];
    std::vector<inner_bc> &ib = bc->inner[which_gz];
    ob.clear();
    ib.clear();
    // One pass over the six faces; each point goes to the outer or inner list.
    auto visit = [&](int a0, int b0, int a1, int b1, int a2, int b2, int f0, int f1, int f2) {
      for (int i2 = a2; i2 < b2; i2++)
        for (int i1 = a1; i1 < b1; i1++)
          for (int i0 = a0; i0 < b0; i0++) {
            const int idx3 = IDX3S(i0, i1, i2);
            if (bc_gz_map[idx3].i0 == -1) {
              outer_bc o;
              o.outer_bc_dest_pt.i0 = (short)i0;
              o.outer_bc_dest_pt.i1 = (short)i1;
              o.outer_bc_dest_pt.i2 = (short)i2;
              o.FACEi0 = (int8_t)f0;
              o.FACEi1 = (int8_t)f1;
              o.FACEi2 = (int8_t)f2;
              ob.push_back(o);
            } else {
              inner_bc n;
              n.inner_bc_dest_pt.i0 = (short)i0;
              n.inner_bc_dest_pt.i1 = (short)i1;
              n.inner_bc_dest_pt.i2 = (short)i2;
              n.inner_bc_src_pt = bc_gz_map[idx3];
              for (int p = 0; p < 10; p++) n.parity[p] = bc_parity_conditions[idx3].parity[p];
              ib.push_back(n);
            }
          }
    };
    visit(imin[0] - 1, imin[0], imin[1], imax[1], imin[2], imax[2], MINFACE, NUL, NUL);
    imin[0]--;
    visit(imax[0], imax[0] + 1, imin[1], imax[1], imin[2], imax[2], MAXFACE, NUL, NUL);
    imax[0]++;
    bc->ob_dir_end[which_gz][0] = (int)ob.size();
    visit(imin[0], imax[0], imin[1] - 1, imin[1], imin[2], imax[2], NUL, MINFACE, NUL);
    imin[1]--;
    visit(imin[0], imax[0], imax[1], imax[1] + 1, imin[2], imax[2], NUL, MAXFACE, NUL);
    imax[1]++;
    bc->ob_dir_end[which_gz][1] = (int)ob.size();
    visit(imin[0], imax[0], imin[1], imax[1], imin[2] - 1, imin[2], NUL, NUL, MINFACE);
    imin[2]--;
    visit(imin[0], imax[0], imin[1], imax[1], imax[2], imax[2] + 1, NUL, NUL, MAXFACE);
    imax[2]++;
    bc->ob_dir_end[which_gz][2] = (int)ob.size();
    bc->num_ob_gz_pts[which_gz] = (int)ob.size();
    bc->num_ib_gz_pts[which_gz] = (int)ib.size();
  }
}

// Expand the sequential BC order into interior-only linear combinations.
struct nrpy_term {
  int src;
  int negmask;
  double coeff;
};

static int nrpy_build_composed_map(const paramstruct *params, bc_struct *bc) {
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;
  const int Ntot = Nxx_plus_2NGHOSTS0 * Nxx_plus_2NGHOSTS1 * Nxx_plus_2NGHOSTS2;
  auto interior = [&](int idx) {
    const int i0 = idx % Nxx_plus_2NGHOSTS0;
    const int i1 = (idx / Nxx_plus_2NGHOSTS0) % Nxx_plus_2NGHOSTS1;
    const int i2 = idx / (Nxx_plus_2NGHOSTS0 * Nxx_plus_2NGHOSTS1);
    return i0 >= NGHOSTS && i0 < Nxx_plus_2NGHOSTS0 - NGHOSTS && i1 >= NGHOSTS &&
           i1 < Nxx_plus_2NGHOSTS1 - NGHOSTS && i2 >= NGHOSTS && i2 < Nxx_plus_2NGHOSTS2 - NGHOSTS;
  };
  std::vector<int> slot(Ntot, -1);            // ghost idx3 -> row in 'rows'
  std::vector<std::vector<nrpy_term>> rows;
  std::vector<int> dests;
  auto expand = [&](int idx, double c, int neg, std::vector<nrpy_term> &acc) -> bool {
    if (idx < 0 || idx >= Ntot) return false;
    if (interior(idx)) {
      acc.push_back({idx, neg, c});
      return true;
    }
    if (slot[idx] < 0) return false;          // ghost read before it was set
    for (const nrpy_term &t : rows[slot[idx]]) acc.push_back({t.src, t.negmask ^ neg, t.coeff * c});
    return true;
  };
  auto finish = [&](int dest, std::vector<nrpy_term> &acc) {
    std::map<std::pair<int, int>, double> merged;
    for (const nrpy_term &t : acc) merged[{t.src, t.negmask}] += t.coeff;
    std::vector<nrpy_term> row;
    for (auto &kv : merged)
      if (kv.second != 0.0) row.push_back({kv.first.first, kv.first.second, kv.second});
    if (slot[dest] < 0) {
      slot[dest] = (int)rows.size();
      rows.push_back(row);
      dests.push_back(dest);
    } else {
      rows[slot[dest]] = row;
    }
  };
  for (int gz = 0; gz < NGHOSTS; gz++) {
    for (const outer_bc &o : bc->outer[gz]) {
      const int dest = IDX3S(o.outer_bc_dest_pt.i0, o.outer_bc_dest_pt.i1, o.outer_bc_dest_pt.i2);
      const int stride = o.FACEi0 + Nxx_plus_2NGHOSTS0 * (o.FACEi1 + Nxx_plus_2NGHOSTS1 * o.FACEi2);
      std::vector<nrpy_term> acc;
      if (!expand(dest + stride, 3.0, 0, acc) || !expand(dest + 2 * stride, -3.0, 0, acc) ||
          !expand(dest + 3 * stride, 1.0, 0, acc))
        return NRPY_ERR_BCMAP;
      finish(dest, acc);
    }
    for (const inner_bc &n : bc->inner[gz]) {
      const int dest = IDX3S(n.inner_bc_dest_pt.i0, n.inner_bc_dest_pt.i1, n.inner_bc_dest_pt.i2);
      const int src = IDX3S(n.inner_bc_src_pt.i0, n.inner_bc_src_pt.i1, n.inner_bc_src_pt.i2);
      int neg = 0;
      for (int p = 0; p < 10; p++)
        if (n.parity[p] < 0) neg |= (1 << p);
      std::vector<nrpy_term> acc;
      if (!expand(src, 1.0, neg, acc)) return NRPY_ERR_BCMAP;
      finish(dest, acc);
    }
  }
  bc->h_gdest.clear();
  bc->h_rowptr.assign(1, 0);
  bc->h_src.clear();
  bc->h_coeff.clear();
  bc->h_negmask.clear();
  // Order rows by destination so consecutive lanes write nearby addresses.
  std::vector<int> order(rows.size());
  for (size_t r = 0; r < rows.size(); r++) order[r] = (int)r;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return dests[a] < dests[b]; });
  for (int r : order) {
    bc->h_gdest.push_back(dests[r]);
    for (const nrpy_term &t : rows[r]) {
      bc->h_src.push_back(t.src);
      bc->h_coeff.push_back(t.coeff);
      bc->h_negmask.push_back(t.negmask);
    }
    bc->h_rowptr.push_back((int)bc->h_src.size());
  }
  bc->nghost = (int)bc->h_gdest.size();
  bc->nterms = (int)bc->h_src.size();
  // Transpose: for every interior source point, the ghost points whose value it enters.
  {
    std::vector<int> perm(bc->h_src.size()), termrow(bc->h_src.size());
    for (int r = 0; r < bc->nghost; r++)
      for (int k = bc->h_rowptr[r]; k < bc->h_rowptr[r + 1]; k++) termrow[k] = r;
    for (size_t k = 0; k < perm.size(); k++) perm[k] = (int)k;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return bc->h_src[a] < bc->h_src[b]; });
    bc->h_tsrc.clear();
    bc->h_trowptr.assign(1, 0);
    bc->h_tghost.clear();
    bc->h_tcoeff.clear();
    bc->h_tnegmask.clear();
    for (size_t q = 0; q < perm.size(); q++) {
      const int k = perm[q];
      if (bc->h_tsrc.empty() || bc->h_tsrc.back() != bc->h_src[k]) {
        if (!bc->h_tsrc.empty()) bc->h_trowptr.push_back((int)bc->h_tghost.size());
        bc->h_tsrc.push_back(bc->h_src[k]);
      }
      bc->h_tghost.push_back(bc->h_gdest[termrow[k]]);
      bc->h_tcoeff.push_back(bc->h_coeff[k]);
      bc->h_tnegmask.push_back(bc->h_negmask[k]);
    }
    if (!bc->h_tsrc.empty()) bc->h_trowptr.push_back((int)bc->h_tghost.size());
    bc->ntsrc = (int)bc->h_tsrc.size();
  }
  return NRPY_OK;
}

template <typename T>
static bool nrpy_upload(T **dst, const T *src, size_t n) {
  *dst = nullptr;
  if (n == 0) return true;
  if (hipMalloc((void **)dst, n * sizeof(T)) != hipSuccess) return false;
  return hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

extern "C" void bcstruct_free(bc_struct *bc) {
  if (!bc) return;
  (void)hipFree(bc->d_ob_dest);
  (void)hipFree(bc->d_ob_stride);
  (void)hipFree(bc->d_ob_face);
  (void)hipFree(bc->d_ob_xedge_dest);
  (void)hipFree(bc->d_ob_xedge_face);
  (void)hipFree(bc->d_ib_dest);
  (void)hipFree(bc->d_ib_src);
  (void)hipFree(bc->d_ib_parity);
  (void)hipFree(bc->d_gdest);
  (void)hipFree(bc->d_rowptr);
  (void)hipFree(bc->d_src);
  (void)hipFree(bc->d_coeff);
  (void)hipFree(bc->d_negmask);
  (void)hipFree(bc->d_tsrc);
  (void)hipFree(bc->d_trowptr);
  (void)hipFree(bc->d_tghost);
  (void)hipFree(bc->d_tcoeff);
  (void)hipFree(bc->d_tnegmask);
  (void)hipFree(bc->d_som_rhs);
  delete bc;
}

// Right-hand-side scratch of the fused Sommerfeld RK4 step: num_gfs gridfunctions over the
// whole grid, owned by the boundary struct (a grid's struct always asks for the same size,
// so it is allocated once and freed by bcstruct_free). The first call allocates it, which
// is not possible while the stream is being captured: then, and if the allocation fails,
// the result is nullptr and the caller takes the unfused step, which needs no scratch.
static double *nrpy_sommerfeld_scratch(const bc_struct *bc, const int num_gfs, hipStream_t stream) {
  bc_struct *b = const_cast<bc_struct *>(bc);
  // num_gfs grids, then the compact RHS at the points next to the i0 boundaries,
  // [gf][i2][i1][side] (side 0 at i0 = NGHOSTS, 1 at i0 = N0 - 1 - NGHOSTS)
  const size_t n = (size_t)num_gfs * bc->N[0] * bc->N[1] * bc->N[2] + (size_t)num_gfs * 2 * bc->N[1] * bc->N[2];
  if (b->d_som_rhs && b->som_rhs_n >= n) return b->d_som_rhs;
  if (b->d_som_rhs) return nullptr;              // never resized: a captured step may hold it
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
  if (hipMalloc((void **)&b->d_som_rhs, n * sizeof(double)) != hipSuccess) {
    b->d_som_rhs = nullptr;
    return nullptr;
  }
  b->som_rhs_n = n;
  return b->d_som_rhs;
}

// driver_bcstruct.h:1-44 as one call: maps at all points, then bcstruct, then
// the composed ghost map, uploaded to the current device. 'params' and 'xx'
// describe the EigenCoord grid (set_Nxx_dxx_invdx_params__and__xx with
// EigenCoord=1); xx are host arrays.
// Host-only part of bcstruct_create: maps, lists and composed map, no device
// memory (usable without a GPU).
extern "C" int bcstruct_create_host(const paramstruct *params, double *xx[3], bc_struct **out) {
  if (!params || !xx || !out) return NRPY_ERR_NULL;
  *out = nullptr;
  if (nrpy_check_params(params)) return NRPY_ERR_SHAPE;
  const size_t Ntot =
      (size_t)params->Nxx_plus_2NGHOSTS0 * params->Nxx_plus_2NGHOSTS1 * params->Nxx_plus_2NGHOSTS2;
  std::vector<gz_map> gzmap(Ntot);
  std::vector<parity_condition> parity(Ntot);
  int rc = set_up__bc_gz_map_and_parity_condns(params, xx, gzmap.data(), parity.data());
  if (rc) return rc;
  bc_struct *bc = new bc_struct();
  bc->N[0] = params->Nxx_plus_2NGHOSTS0;
  bc->N[1] = params->Nxx_plus_2NGHOSTS1;
  bc->N[2] = params->Nxx_plus_2NGHOSTS2;
  nrpy_set_bcstruct(params, gzmap.data(), parity.data(), bc);
  rc = nrpy_build_composed_map(params, bc);
  if (rc) {
    bcstruct_free(bc);
    return rc;
  }
  *out = bc;
  return NRPY_OK;
}

extern "C" int bcstruct_create(const paramstruct *params, double *xx[3], bc_struct **out) {
  int rc = bcstruct_create_host(params, xx, out);
  if (rc) return rc;
  bc_struct *bc = *out;
  *out = nullptr;
  std::vector<int> obd, obs, ibd, ibs, oxd;
  std::vector<int8_t> ibp, obf, oxf;
  const int N0 = bc->N[0], N1 = bc->N[1], N2 = bc->N[2];
  bc->ob_off[0] = bc->ib_off[0] = bc->ob_xedge_off[0] = 0;
  for (int gz = 0; gz < NGHOSTS; gz++) {
    for (int q = 0; q < bc->ob_dir_end[gz][0]; q++) {
      const outer_bc &o = bc->outer[gz][q];
      const int i1 = o.outer_bc_dest_pt.i1, i2 = o.outer_bc_dest_pt.i2;
      if (i1 >= NGHOSTS && i1 < N1 - NGHOSTS && i2 >= NGHOSTS && i2 < N2 - NGHOSTS) continue;
      oxd.push_back(o.outer_bc_dest_pt.i0 + N0 * (i1 + N1 * i2));
      oxf.push_back(o.FACEi0);
      oxf.push_back(o.FACEi1);
      oxf.push_back(o.FACEi2);
    }
    bc->ob_xedge_off[gz + 1] = (int)oxd.size();
    for (const outer_bc &o : bc->outer[gz]) {
      obd.push_back(o.outer_bc_dest_pt.i0 + N0 * (o.outer_bc_dest_pt.i1 + N1 * o.outer_bc_dest_pt.i2));
      obs.push_back(o.FACEi0 + N0 * (o.FACEi1 + N1 * o.FACEi2));
      obf.push_back(o.FACEi0);
      obf.push_back(o.FACEi1);
      obf.push_back(o.FACEi2);
    }
    for (const inner_bc &n : bc->inner[gz]) {
      ibd.push_back(n.inner_bc_dest_pt.i0 + N0 * (n.inner_bc_dest_pt.i1 + N1 * n.inner_bc_dest_pt.i2));
      ibs.push_back(n.inner_bc_src_pt.i0 + N0 * (n.inner_bc_src_pt.i1 + N1 * n.inner_bc_src_pt.i2));
      for (int p = 0; p < 10; p++) ibp.push_back(n.parity[p]);
    }
    bc->ob_off[gz + 1] = (int)obd.size();
    bc->ib_off[gz + 1] = (int)ibd.size();
  }
  bool ok = nrpy_upload(&bc->d_ob_dest, obd.data(), obd.size()) &&
            nrpy_upload(&bc->d_ob_stride, obs.data(), obs.size()) &&
            nrpy_upload(&bc->d_ob_face, obf.data(), obf.size()) &&
            nrpy_upload(&bc->d_ob_xedge_dest, oxd.data(), oxd.size()) &&
            nrpy_upload(&bc->d_ob_xedge_face, oxf.data(), oxf.size()) &&
            nrpy_upload(&bc->d_ib_dest, ibd.data(), ibd.size()) &&
            nrpy_upload(&bc->d_ib_src, ibs.data(), ibs.size()) &&
            nrpy_upload(&bc->d_ib_parity, ibp.data(), ibp.size()) &&
            nrpy_upload(&bc->d_gdest, bc->h_gdest.data(), bc->h_gdest.size()) &&
            nrpy_upload(&bc->d_rowptr, bc->h_rowptr.data(), bc->h_rowptr.size()) &&
            nrpy_upload(&bc->d_src, bc->h_src.data(), bc->h_src.size()) &&
            nrpy_upload(&bc->d_coeff, bc->h_coeff.data(), bc->h_coeff.size()) &&
            nrpy_upload(&bc->d_negmask, bc->h_negmask.data(), bc->h_negmask.size()) &&
            nrpy_upload(&bc->d_tsrc, bc->h_tsrc.data(), bc->h_tsrc.size()) &&
            nrpy_upload(&bc->d_trowptr, bc->h_trowptr.data(), bc->h_trowptr.size()) &&
            nrpy_upload(&bc->d_tghost, bc->h_tghost.data(), bc->h_tghost.size()) &&
            nrpy_upload(&bc->d_tcoeff, bc->h_tcoeff.data(), bc->h_tcoeff.size()) &&
            nrpy_upload(&bc->d_tnegmask, bc->h_tnegmask.data(), bc->h_tnegmask.size());
  if (!ok) {
    bcstruct_free(bc);
    return NRPY_ERR_ALLOC;
  }
  *out = bc;
  return NRPY_OK;
}

// Host-side inspection of the boundary lists (for bit-exact comparisons).
extern "C" int bcstruct_get_counts(const bc_struct *bc, int *num_ob_gz_pts, int *num_ib_gz_pts, int *nghost_total,
                                   int *nterms_total) {
  if (!bc) return NRPY_ERR_NULL;
  for (int gz = 0; gz < NGHOSTS; gz++) {
    if (num_ob_gz_pts) num_ob_gz_pts[gz] = bc->num_ob_gz_pts[gz];
    if (num_ib_gz_pts) num_ib_gz_pts[gz] = bc->num_ib_gz_pts[gz];
  }
  if (nghost_total) *nghost_total = bc->nghost;
  if (nterms_total) *nterms_total = bc->nterms;
  return NRPY_OK;
}

extern "C" int bcstruct_get_layer(const bc_struct *bc, int which_gz, outer_bc *outer_out, inner_bc *inner_out) {
  if (!bc) return NRPY_ERR_NULL;
  if (which_gz < 0 || which_gz >= NGHOSTS) return NRPY_ERR_PARAM;
  if (outer_out)
    memcpy(outer_out, bc->outer[which_gz].data(), sizeof(outer_bc) * bc->outer[which_gz].size());
  if (inner_out)
    memcpy(inner_out, bc->inner[which_gz].data(), sizeof(inner_bc) * bc->inner[which_gz].size());
  return NRPY_OK;
}

extern "C" int bcstruct_get_composed(const bc_struct *bc, int *gdest, int *rowptr, int *src, double *coeff,
                                     int *negmask) {
  if (!bc) return NRPY_ERR_NULL;
  if (gdest) memcpy(gdest, bc->h_gdest.data(), sizeof(int) * bc->h_gdest.size());
  if (rowptr) memcpy(rowptr, bc->h_rowptr.data(), sizeof(int) * bc->h_rowptr.size());
  if (src) memcpy(src, bc->h_src.data(), sizeof(int) * bc->h_src.size());
  if (coeff) memcpy(coeff, bc->h_coeff.data(), sizeof(double) * bc->h_coeff.size());
  if (negmask) memcpy(negmask, bc->h_negmask.data(), sizeof(int) * bc->h_negmask.size());
  return NRPY_OK;
}

// ---------------------------------------------------------------------------
// 2. Sequential-order boundary application on the device.

__global__ void nrpy_bc_outer_kernel(int npts, const int *__restrict__ dest, const int *__restrict__ stride,
                                     int num_gfs, size_t Ntot, double *__restrict__ gfs) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)npts * num_gfs) return;
  const int gf = (int)(t / npts);
  const int p = (int)(t % npts);
  double *g = gfs + (size_t)gf * Ntot;
  const int d = dest[p], s = stride[p];
  g[d] = +3.0 * g[d + s] - 3.0 * g[d + 2 * s] + 1.0 * g[d + 3 * s];
}

__global__ void nrpy_bc_inner_kernel(int npts, const int *__restrict__ dest, const int *__restrict__ src,
                                     const int8_t *__restrict__ parity, int num_gfs, nrpy_gf_parity_args gfp,
                                     size_t Ntot, double *__restrict__ gfs) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)npts * num_gfs) return;
  const int gf = (int)(t / npts);
  const int p = (int)(t % npts);
  double *g = gfs + (size_t)gf * Ntot;
  g[dest[p]] = (double)parity[10 * p + gfp.p[gf]] * g[src[p]];
}

static int nrpy_bc_check(const paramstruct *params, const bc_struct *bc, int NUM_GFS, const int8_t *gfs_parity,
                         const double *gfs, nrpy_gf_parity_args *gfp) {
  if (!params || !bc || !gfs_parity || !gfs) return NRPY_ERR_NULL;
  if (NUM_GFS < 1 || NUM_GFS > 64) return NRPY_ERR_PARAM;
  if (bc->N[0] != params->Nxx_plus_2NGHOSTS0 || bc->N[1] != params->Nxx_plus_2NGHOSTS1 ||
      bc->N[2] != params->Nxx_plus_2NGHOSTS2)
    return NRPY_ERR_SHAPE;
  for (int g = 0; g < NUM_GFS; g++) {
    if (gfs_parity[g] < 0 || gfs_parity[g] > 9) return NRPY_ERR_PARAM;
    gfp->p[g] = gfs_parity[g];
  }
  return NRPY_OK;
}

// apply_bcs_curvilinear.h:56-64 with the reference's order of updates.
// gfs_parity is a host array of NUM_GFS parity types (0..9).
extern "C" int apply_bcs_curvilinear(const paramstruct *params, const bc_struct *bc, const int NUM_GFS,
                                     const int8_t *gfs_parity, double *gfs, hipStream_t stream) {
  nrpy_gf_parity_args gfp;
  memset(&gfp, 0, sizeof(gfp));
  int rc = nrpy_bc_check(params, bc, NUM_GFS, gfs_parity, gfs, &gfp);
  if (rc) return rc;
  const size_t Ntot = (size_t)bc->N[0] * bc->N[1] * bc->N[2];
  const int B = 256;
  for (int gz = 0; gz < NGHOSTS; gz++) {
    int start = 0;
    for (int dir = 0; dir < 3; dir++) {
      const int end = bc->ob_dir_end[gz][dir];
      const int n = end - start;
      if (n > 0) {
        const size_t work = (size_t)n * NUM_GFS;
        hipLaunchKernelGGL(nrpy_bc_outer_kernel, dim3((unsigned)((work + B - 1) / B)), dim3(B), 0, stream, n,
                           bc->d_ob_dest + bc->ob_off[gz] + start, bc->d_ob_stride + bc->ob_off[gz] + start,
                           NUM_GFS, Ntot, gfs);
      }
      start = end;
    }
    const int ni = bc->num_ib_gz_pts[gz];
    if (ni > 0) {
      const size_t work = (size_t)ni * NUM_GFS;
      hipLaunchKernelGGL(nrpy_bc_inner_kernel, dim3((unsigned)((work + B - 1) / B)), dim3(B), 0, stream, ni,
                         bc->d_ib_dest + bc->ib_off[gz], bc->d_ib_src + bc->ib_off[gz],
                         bc->d_ib_parity + 10 * (size_t)bc->ib_off[gz], NUM_GFS, gfp, Ntot, gfs);
    }
  }
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// 3. Composed ghost fill: one launch, reads interior points only.

// Sum over the entries [b, e) of a map row of c_k g[idx_k], in order (acc += c v, the
// multiply-add contracted), c_k = -coef_k where mask_k has the gridfunction's parity bit,
// else coef_k. The first four entries load unconditionally (at an index held inside the
// row) and each term is added only if its entry exists, so their index loads, then their
// value loads, are in flight together: a loop waits twice per entry (most rows hold three).
// Longer rows go on in a loop.
__device__ __forceinline__ double nrpy_map_row_sum(const int b, const int e, const int *__restrict__ idx,
                                                   const double *__restrict__ coef, const int *__restrict__ mask,
                                                   const int bit, const double *g) {
  double acc = 0.0;
  if (e > b) {
    double c[4], v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int k = min(b + j, e - 1);
      c[j] = (mask[k] & bit) ? -coef[k] : coef[k];
      v[j] = g[idx[k]];
    }
    // an empty statement that uses every loaded value on every path, after all the loads:
    // otherwise the compiler moves the loads of entries past a row's end under branches
    asm volatile("" : : "v"(c[0]), "v"(v[0]), "v"(c[1]), "v"(v[1]), "v"(c[2]), "v"(v[2]), "v"(c[3]), "v"(v[3]));
#pragma unroll
    for (int j = 0; j < 4; j++) acc = b + j < e ? fma(c[j], v[j], acc) : acc;   /* acc += c * v, contracted */
  }
  for (int k = b + 4; k < e; k++) {
    const double c = (mask[k] & bit) ? -coef[k] : coef[k];
    acc += c * g[idx[k]];
  }
  return acc;
}

__global__ void nrpy_bc_composed_kernel(int nghost, const int *__restrict__ gdest, const int *__restrict__ rowptr,
                                        const int *__restrict__ src, const double *__restrict__ coeff,
                                        const int *__restrict__ negmask, int num_gfs, nrpy_gf_parity_args gfp,
                                        size_t Ntot, double *__restrict__ gfs) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nghost * num_gfs) return;
  const int gf = (int)(t / nghost);
  const int r = (int)(t % nghost);
  double *g = gfs + (size_t)gf * Ntot;
  const int bit = 1 << gfp.p[gf];
  g[gdest[r]] = nrpy_map_row_sum(rowptr[r], rowptr[r + 1], src, coeff, negmask, bit, g);
}

extern "C" int apply_bcs_curvilinear_composed(const paramstruct *params, const bc_struct *bc, const int NUM_GFS,
                                              const int8_t *gfs_parity, double *gfs, hipStream_t stream) {
  nrpy_gf_parity_args gfp;
  memset(&gfp, 0, sizeof(gfp));
  int rc = nrpy_bc_check(params, bc, NUM_GFS, gfs_parity, gfs, &gfp);
  if (rc) return rc;
  if (bc->nghost == 0) return NRPY_OK;
  const size_t Ntot = (size_t)bc->N[0] * bc->N[1] * bc->N[2];
  const int B = 256;
  const size_t work = (size_t)bc->nghost * NUM_GFS;
  hipLaunchKernelGGL(nrpy_bc_composed_kernel, dim3((unsigned)((work + B - 1) / B)), dim3(B), 0, stream,
                     bc->nghost, bc->d_gdest, bc->d_rowptr, bc->d_src, bc->d_coeff, bc->d_negmask, NUM_GFS, gfp,
                     Ntot, gfs);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// 3b. Transpose of the ghost fill, for adjoints. The ghost fill is linear:
// ghost = sum_k c_k s_k src_k (s_k the parity sign of the gridfunction's type), so
// its transpose adds, at every interior source point, sum_k c_k s_k cot(ghost_k),
// and then the ghost cotangents are cleared. Sources are interior and ghosts are
// not, so the gather reads and writes disjoint points; a second launch clears.

__global__ void nrpy_bc_transpose_kernel(int ntsrc, const int *__restrict__ tsrc, const int *__restrict__ trowptr,
                                         const int *__restrict__ tghost, const double *__restrict__ tcoeff,
                                         const int *__restrict__ tnegmask, int num_gfs, nrpy_gf_parity_args gfp,
                                         size_t Ntot, double *__restrict__ gfs) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)ntsrc * num_gfs) return;
  const int gf = (int)(t / ntsrc);
  const int r = (int)(t % ntsrc);
  double *g = gfs + (size_t)gf * Ntot;
  const int bit = 1 << gfp.p[gf];
  g[tsrc[r]] += nrpy_map_row_sum(trowptr[r], trowptr[r + 1], tghost, tcoeff, tnegmask, bit, g);
}

__global__ void nrpy_clear_ghosts_kernel(int nghost, const int *__restrict__ gdest, int num_gfs, size_t Ntot,
                                         double *__restrict__ gfs) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nghost * num_gfs) return;
  gfs[(size_t)(t / nghost) * Ntot + gdest[t % nghost]] = 0.0;
}

// Transpose of apply_bcs_curvilinear(_composed): folds the cotangents held at ghost
// points onto the interior points the ghost values were computed from, then zeroes
// the ghost cotangents. No counterpart in the reference (it has no adjoint).
extern "C" int bc_apply_T(const paramstruct *params, const bc_struct *bc, const int NUM_GFS,
                          const int8_t *gfs_parity, double *cot_gfs, hipStream_t stream) {
  nrpy_gf_parity_args gfp;
  memset(&gfp, 0, sizeof(gfp));
  int rc = nrpy_bc_check(params, bc, NUM_GFS, gfs_parity, cot_gfs, &gfp);
  if (rc) return rc;
  if (bc->nghost == 0) return NRPY_OK;
  const size_t Ntot = (size_t)bc->N[0] * bc->N[1] * bc->N[2];
  const int B = 256;
  if (bc->ntsrc > 0) {
    const size_t work = (size_t)bc->ntsrc * NUM_GFS;
    hipLaunchKernelGGL(nrpy_bc_transpose_kernel, dim3((unsigned)((work + B - 1) / B)), dim3(B), 0, stream,
                       bc->ntsrc, bc->d_tsrc, bc->d_trowptr, bc->d_tghost, bc->d_tcoeff, bc->d_tnegmask, NUM_GFS,
                       gfp, Ntot, cot_gfs);
  }
  const size_t work = (size_t)bc->nghost * NUM_GFS;
  hipLaunchKernelGGL(nrpy_clear_ghosts_kernel, dim3((unsigned)((work + B - 1) / B)), dim3(B), 0, stream,
                     bc->nghost, bc->d_gdest, NUM_GFS, Ntot, cot_gfs);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

__global__ void nrpy_lincomb_kernel(size_t n, double *__restrict__ out, double a, const double *__restrict__ x,
                                    double b, const double *__restrict__ y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = y ? a * x[i] + b * y[i] : a * x[i];
}

// RK4 update of the diagonal schedule at every point (interior and ghost) of
// NUM_EVOL_GFS gridfunctions, the loop of the reference's generated RK_MoL.h
// (MoLtimestepping/C_Code_Generation.py:266-329). k holds the RHS on entry and
// the next stage's input on exit (substeps 0-2); substep 3 updates y_n.
__global__ void nrpy_rk4_update_all_kernel(int S, size_t n, double *k, double *y_n, double *rt, double dt) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double ki = k[i];
    if (S == 0) {
      rt[i] = ki * dt * (1.0 / 6.0);
      k[i] = y_n[i] + ki * dt * (1.0 / 2.0);
    } else if (S == 1) {
      rt[i] = rt[i] + ki * dt * (1.0 / 3.0);
      k[i] = y_n[i] + ki * dt * (1.0 / 2.0);
    } else if (S == 2) {
      rt[i] = rt[i] + ki * dt * (1.0 / 3.0);
      k[i] = y_n[i] + ki * dt;
    } else {
      y_n[i] = y_n[i] + rt[i] + ki * dt * (1.0 / 6.0);
    }
  }
}

extern "C" int nrpy_rk4_update_all(const int S, size_t n, double *k, double *y_n, double *rt, const double dt,
                                   hipStream_t stream) {
  if (!k || !y_n || !rt) return NRPY_ERR_NULL;
  if (S < 0 || S > 3) return NRPY_ERR_PARAM;
  if (n == 0) return NRPY_OK;
  const int B = 256;
  size_t nb = (n + B - 1) / B;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(nrpy_rk4_update_all_kernel, dim3((unsigned)nb), dim3(B), 0, stream, S, n, k, y_n, rt, dt);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

__global__ void nrpy_fill_kernel(size_t n, double *__restrict__ out, double v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = v;
}

// out[i] = v for i < n (a kernel, not a memset, so it is safe inside captured graphs).
extern "C" int nrpy_fill(size_t n, double *out, double v, hipStream_t stream) {
  if (!out) return NRPY_ERR_NULL;
  if (n == 0) return NRPY_OK;
  const int B = 256;
  size_t nb = (n + B - 1) / B;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(nrpy_fill_kernel, dim3((unsigned)nb), dim3(B), 0, stream, n, out, v);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

__global__ void nrpy_adjoint_combine_kernel(size_t n, double *__restrict__ acc, const double *__restrict__ L,
                                            const double *__restrict__ w, double *__restrict__ kbar, double a,
                                            double b, int first) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double wi = w[i], Li = L[i];
    acc[i] = (first ? Li : acc[i]) + wi;
    if (kbar) kbar[i] = a * Li + b * wi;
  }
}

// One pass of the RK4 step adjoint: acc = (first ? L : acc) + w and, if kbar is not
// NULL, kbar = a*L + b*w (the next stage's RHS cotangent).
extern "C" int nrpy_adjoint_combine(size_t n, double *acc, const double *L, const double *w, double *kbar, double a,
                                    double b, int first, hipStream_t stream) {
  if (!acc || !L || !w) return NRPY_ERR_NULL;
  if (n == 0) return NRPY_OK;
  const int B = 256;
  size_t nb = (n + B - 1) / B;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(nrpy_adjoint_combine_kernel, dim3((unsigned)nb), dim3(B), 0, stream, n, acc, L, w, kbar, a, b,
                     first);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

// out[i] = a*x[i] + b*y[i] for i < n (y may be NULL: out = a*x). out may alias x or y.
extern "C" int nrpy_lincomb(size_t n, double *out, double a, const double *x, double b, const double *y,
                            hipStream_t stream) {
  if (!out || !x) return NRPY_ERR_NULL;
  if (n == 0) return NRPY_OK;
  const int B = 256;
  size_t nb = (n + B - 1) / B;
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(nrpy_lincomb_kernel, dim3((unsigned)nb), dim3(B), 0, stream, n, out, a, x, b, y);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// 4. Deterministic reduction: sum over interior points of f^2 for one
// gridfunction. Two passes (per-block partials, then one block), so the
// result does not depend on scheduling.

#define NRPY_RED_BLOCK 256
#define NRPY_RED_MAXBLOCKS 1024

__global__ void __launch_bounds__(NRPY_RED_BLOCK)
nrpy_sumsq_partial_kernel(const paramstruct params, const double *__restrict__ f, double *__restrict__ partials) {
  const int Nxx_plus_2NGHOSTS0 = params.Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params.Nxx_plus_2NGHOSTS1;
  const long n0 = params.Nxx0, n1 = params.Nxx1, n2 = params.Nxx2;
  const long n = n0 * n1 * n2;
  double acc = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int i0 = (int)(t % n0) + NGHOSTS;
    const int i1 = (int)((t / n0) % n1) + NGHOSTS;
    const int i2 = (int)(t / (n0 * n1)) + NGHOSTS;
    const double v = f[IDX3S(i0, i1, i2)];
    acc += v * v;
  }
  __shared__ double sh[NRPY_RED_BLOCK];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = NRPY_RED_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(NRPY_RED_BLOCK)
nrpy_sum_final_kernel(int nparts, const double *__restrict__ partials, double *__restrict__ result) {
  __shared__ double sh[NRPY_RED_BLOCK];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += NRPY_RED_BLOCK) acc += partials[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = NRPY_RED_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) result[0] = sh[0];
}

// Sum of squares of gridfunction 'gf' of 'gfs' over interior points, written
// to result_dev[0]. 'workspace' is a caller-provided device buffer of at least
// NRPY_RED_MAXBLOCKS doubles.
extern "C" int gridfunction_sumsq(const paramstruct *params, const double *gfs, int gf, double *workspace,
                                  double *result_dev, hipStream_t stream) {
  if (!params || !gfs || !workspace || !result_dev) return NRPY_ERR_NULL;
  if (nrpy_check_params(params)) return NRPY_ERR_SHAPE;
  if (gf < 0) return NRPY_ERR_PARAM;
  const size_t Ntot =
      (size_t)params->Nxx_plus_2NGHOSTS0 * params->Nxx_plus_2NGHOSTS1 * params->Nxx_plus_2NGHOSTS2;
  const long n = (long)params->Nxx0 * params->Nxx1 * params->Nxx2;
  int nb = (int)std::min<long>((n + NRPY_RED_BLOCK - 1) / NRPY_RED_BLOCK, NRPY_RED_MAXBLOCKS);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(nrpy_sumsq_partial_kernel, dim3(nb), dim3(NRPY_RED_BLOCK), 0, stream, *params,
                     gfs + (size_t)gf * Ntot, workspace);
  hipLaunchKernelGGL(nrpy_sum_final_kernel, dim3(1), dim3(NRPY_RED_BLOCK), 0, stream, nb, workspace, result_dev);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

// The same sum with no workspace argument and no allocation, for callers that
// cannot allocate device memory (code under stream capture, serving loops): the
// per-block partial sums live in 'scratch', a file-scope device array of
// NRPY_RED_MAXBLOCKS doubles (8 KiB) that each GPU running this library holds
// until the process ends. The partial kernel writes every entry the final kernel
// reads, so nothing needs clearing between calls. Calls on one GPU share the
// array: calls that overlap in time on that GPU (different streams, or host
// threads) race and must be ordered by the caller.
static __device__ __align__(256) double scratch[NRPY_RED_MAXBLOCKS];

#define NRPY_MAX_DEVICES 64
static double *nrpy_scratch_dev[NRPY_MAX_DEVICES];  // device address of 'scratch', per device
static std::mutex nrpy_scratch_lock;

// Device address of the current device's copy of 'scratch', looked up once per
// device with that device current (the host-side name is only a stand-in).
static int nrpy_scratch_address(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return NRPY_ERR_LAUNCH;
  if (dev < 0 || dev >= NRPY_MAX_DEVICES) return NRPY_ERR_TOO_LARGE;
  std::lock_guard<std::mutex> guard(nrpy_scratch_lock);
  if (!nrpy_scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NRPY_ERR_LAUNCH;
    nrpy_scratch_dev[dev] = (double *)p;
  }
  *out = nrpy_scratch_dev[dev];
  return NRPY_OK;
}

extern "C" int gridfunction_sumsq_noalloc(const paramstruct *params, const double *gfs, int gf, double *result_dev,
                                          hipStream_t stream) {
  if (!params || !gfs || !result_dev) return NRPY_ERR_NULL;
  if (nrpy_check_params(params)) return NRPY_ERR_SHAPE;
  if (gf < 0) return NRPY_ERR_PARAM;
  const size_t Ntot =
      (size_t)params->Nxx_plus_2NGHOSTS0 * params->Nxx_plus_2NGHOSTS1 * params->Nxx_plus_2NGHOSTS2;
  const long n = (long)params->Nxx0 * params->Nxx1 * params->Nxx2;
  // one partial per block; the grid-stride loop bounds the block count for any grid
  const long want = (n + NRPY_RED_BLOCK - 1) / NRPY_RED_BLOCK;
  const int nb = (int)std::max<long>(1, std::min<long>(want, NRPY_RED_MAXBLOCKS));
  if (nb > NRPY_RED_MAXBLOCKS) return NRPY_ERR_TOO_LARGE;
  double *ptr = nullptr;
  const int rc = nrpy_scratch_address(&ptr);
  if (rc != NRPY_OK) return rc;
  hipLaunchKernelGGL(nrpy_sumsq_partial_kernel, dim3(nb), dim3(NRPY_RED_BLOCK), 0, stream, *params,
                     gfs + (size_t)gf * Ntot, ptr);
  hipLaunchKernelGGL(nrpy_sum_final_kernel, dim3(1), dim3(NRPY_RED_BLOCK), 0, stream, nb, (const double *)ptr,
                     result_dev);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// 5. RK4 update, diagonal Butcher-table schedule
// (MoLtimestepping/C_Code_Generation.py:266-329, RK4 table at
// RK_Butcher_Table_Dictionary.py:86-91). 'r' is the RHS at the current point;
// the arithmetic order matches the reference's generated update loop.
//
// A fused kernel first loads y_n (and the running total) of every output
// gridfunction with NRPY_RK4_PRELOAD(S, gf, k), then stores all of them with
// NRPY_RK4_STORE(S, k, rhs). rk_out is rk_yn in the last substep, so the
// arrays cannot be __restrict__; with the loads interleaved with the stores,
// the compiler would have to complete each load before the previous store
// (one memory round trip per output). Loading everything first is the same
// computation: each thread reads and writes only its own point.
#define NRPY_RK4_PRELOAD(S, gf, k)                                                        \
  const size_t nrpy_ix_##k = IDX4(gf, i0, i1, i2);                                        \
  const double nrpy_y_##k = rk_yn[nrpy_ix_##k];                                           \
  const double nrpy_r_##k = ((S) == 0) ? 0.0 : rk_rt[nrpy_ix_##k];
#define NRPY_RK4_STORE(S, k, rhsval)                                                      \
  do {                                                                                    \
    const double _r = (rhsval);                                                           \
    if ((S) == 0) {                                                                       \
      rk_rt[nrpy_ix_##k] = _r * dt * (1.0 / 6.0);                                           \
      rk_out[nrpy_ix_##k] = nrpy_y_##k + _r * dt * (1.0 / 2.0);                             \
    } else if ((S) == 1) {                                                                \
      rk_rt[nrpy_ix_##k] = nrpy_r_##k + _r * dt * (1.0 / 3.0);                              \
      rk_out[nrpy_ix_##k] = nrpy_y_##k + _r * dt * (1.0 / 2.0);                             \
    } else if ((S) == 2) {                                                                \
      rk_rt[nrpy_ix_##k] = nrpy_r_##k + _r * dt * (1.0 / 3.0);                              \
      rk_out[nrpy_ix_##k] = nrpy_y_##k + _r * dt;                                           \
    } else {                                                                              \
      rk_out[nrpy_ix_##k] = nrpy_y_##k + nrpy_r_##k + _r * dt * (1.0 / 6.0);                \
    }                                                                                     \
  } while (0)
// The same for kernels in which some lanes recompute points other lanes own (NRPY_OWNED).
#define NRPY_RK4_STORE_OWNED(S, k, rhsval)                                                   \
  do {                                                                                    \
    const double _r = (rhsval);                                                           \
    if ((S) == 0) {                                                                       \
      *NRPY_OWNED(&rk_rt[nrpy_ix_##k]) = _r * dt * (1.0 / 6.0);                                           \
      *NRPY_OWNED(&rk_out[nrpy_ix_##k]) = nrpy_y_##k + _r * dt * (1.0 / 2.0);                             \
    } else if ((S) == 1) {                                                                \
      *NRPY_OWNED(&rk_rt[nrpy_ix_##k]) = nrpy_r_##k + _r * dt * (1.0 / 3.0);                              \
      *NRPY_OWNED(&rk_out[nrpy_ix_##k]) = nrpy_y_##k + _r * dt * (1.0 / 2.0);                             \
    } else if ((S) == 2) {                                                                \
      *NRPY_OWNED(&rk_rt[nrpy_ix_##k]) = nrpy_r_##k + _r * dt * (1.0 / 3.0);                              \
      *NRPY_OWNED(&rk_out[nrpy_ix_##k]) = nrpy_y_##k + _r * dt;                                           \
    } else {                                                                              \
      *NRPY_OWNED(&rk_out[nrpy_ix_##k]) = nrpy_y_##k + nrpy_r_##k + _r * dt * (1.0 / 6.0);                \
    }                                                                                     \
  } while (0)
