// Batched evolutions: many independent grids (uneven sizes, own parameters and
// boundary maps) advanced together, one launch per kernel per RK4 substep.
//
// The array roles (NRPY_BATCH_ROLE_*) and index ranges (NRPY_RANGE_*) are defined in
// nrpy_runtime_pre.h. Included by a generated bundle after it has defined nrpy_batch_instance:
//   typedef struct { paramstruct params; rfm_struct rfm; double *arr[NRPY_BATCH_ROLES];
//                    double dt; long Ntot; } nrpy_batch_instance;
// and the table nrpy_batched_kernels[] of {kernel id, range} pairs.
//
// Every launch covers the tiles of all instances: a block finds its instance by
// binary search in a prefix table of tile counts (NRPY_BATCH_FIND) and its tile
// inside that instance's index box (NRPY_BATCH_TILE). The prefix tables depend
// on the block shape and on the index range (interior or all points); they are
// built on the host for the current shapes by nrpy_batch_prepare, before any
// graph capture.
#pragma once

#define NRPY_BATCH_FIND()                                                                      \
  int nrpy_k;                                                                                  \
  {                                                                                            \
    int lo_ = 0, hi_ = ninst - 1;                                                              \
    const int b_ = (int)blockIdx.x;                                                            \
    while (lo_ < hi_) {                                                                        \
      const int mid_ = (lo_ + hi_ + 1) >> 1;                                                   \
      if (tile_prefix[mid_] <= b_) lo_ = mid_; else hi_ = mid_ - 1;                            \
    }                                                                                          \
    nrpy_k = lo_;                                                                              \
  }

#define NRPY_BATCH_TILE(n0, n1, n2)                                                            \
  int nrpy_bx, nrpy_by, nrpy_bz;                                                               \
  {                                                                                            \
    const int nbx_ = ((n0) + (int)blockDim.x - 1) / (int)blockDim.x;                           \
    const int nby_ = ((n1) + (int)blockDim.y - 1) / (int)blockDim.y;                           \
    const int t_ = (int)blockIdx.x - tile_prefix[nrpy_k];                                      \
    nrpy_bx = t_ % nbx_;                                                                       \
    nrpy_by = (t_ / nbx_) % nby_;                                                              \
    nrpy_bz = t_ / (nbx_ * nby_);                                                              \
  }

struct nrpy_batch {
  int ninst = 0;
  std::vector<nrpy_batch_instance> h_inst;
  nrpy_batch_instance *d_inst = nullptr;
  // prefix tables: key (bx, by, bz, range) -> device array of ninst + 1 tile offsets
  std::map<std::vector<int>, int *> prefix;
  std::map<std::vector<int>, int> ntiles;
  // concatenated composed ghost maps: row r belongs to instance row_inst[r]
  int nrows = 0;
  int *d_row_inst = nullptr, *d_row_dest = nullptr, *d_rowptr = nullptr, *d_src = nullptr, *d_negmask = nullptr;
  double *d_coeff = nullptr;
};

static void nrpy_batch_range(const nrpy_batch_instance &I, int range, int n[3]) {
  if (range == NRPY_RANGE_INTERIOR) {
    n[0] = I.params.Nxx0; n[1] = I.params.Nxx1; n[2] = I.params.Nxx2;
  } else {
    n[0] = I.params.Nxx_plus_2NGHOSTS0; n[1] = I.params.Nxx_plus_2NGHOSTS1; n[2] = I.params.Nxx_plus_2NGHOSTS2;
  }
}

static int nrpy_batch_make_prefix(nrpy_batch *b, dim3 block, int range) {
  std::vector<int> key = {(int)block.x, (int)block.y, (int)block.z, range};
  if (b->prefix.count(key)) return NRPY_OK;
  std::vector<int> pre(b->ninst + 1, 0);
  for (int k = 0; k < b->ninst; k++) {
    int n[3];
    nrpy_batch_range(b->h_inst[k], range, n);
    const long t = (long)nrpy_nblocks(block, n[0], n[1], n[2]);
    if ((long)pre[k] + t > 2147483647L) return NRPY_ERR_TOO_LARGE;
    pre[k + 1] = pre[k] + (int)t;
  }
  int *d = nullptr;
  if (!nrpy_upload(&d, pre.data(), pre.size())) return NRPY_ERR_ALLOC;
  b->prefix[key] = d;
  b->ntiles[key] = pre[b->ninst];
  return NRPY_OK;
}

// Device prefix table and total tile count for a launch; NULL if not prepared.
static const int *nrpy_batch_prefix(const nrpy_batch *b, dim3 block, int range, unsigned *nblocks) {
  std::vector<int> key = {(int)block.x, (int)block.y, (int)block.z, range};
  auto it = b->prefix.find(key);
  if (it == b->prefix.end()) return nullptr;
  *nblocks = (unsigned)b->ntiles.at(key);
  return it->second;
}

extern "C" void nrpy_batch_free(nrpy_batch *b) {
  if (!b) return;
  (void)hipFree(b->d_inst);
  for (auto &kv : b->prefix) (void)hipFree(kv.second);
  (void)hipFree(b->d_row_inst);
  (void)hipFree(b->d_row_dest);
  (void)hipFree(b->d_rowptr);
  (void)hipFree(b->d_src);
  (void)hipFree(b->d_negmask);
  (void)hipFree(b->d_coeff);
  delete b;
}

// Builds prefix tables for the current block shape of every batched kernel. Call
// after changing shapes with nrpy_set_block_shape and before capturing a graph.
extern "C" int nrpy_batch_prepare(nrpy_batch *b) {
  if (!b) return NRPY_ERR_NULL;
  const int nk = (int)(sizeof(nrpy_batched_kernels) / sizeof(nrpy_batched_kernels[0]));
  for (int i = 0; i < nk; i++) {
    const dim3 block = nrpy_block(nrpy_batched_kernels[i][0], nrpy_batched_kernels[i][2], nrpy_batched_kernels[i][3],
                                  nrpy_batched_kernels[i][4]);
    const int rc = nrpy_batch_make_prefix(b, block, nrpy_batched_kernels[i][1]);
    if (rc) return rc;
  }
  return NRPY_OK;
}

// ninst instances: params[k] (host), rfms[k] and bcs[k] (as returned by rfm_struct_create /
// bcstruct_create for instance k's grid), arrays[6*k + role] device arrays of instance k in
// the roles above (auxevol/aux may be NULL if the system has none), dts[k] its time step.
extern "C" int nrpy_batch_create(const int ninst, const paramstruct *params, const rfm_struct *const *rfms,
                                 const bc_struct *const *bcs, double *const *arrays, const double *dts,
                                 nrpy_batch **out) {
  if (!params || !rfms || !bcs || !arrays || !dts || !out) return NRPY_ERR_NULL;
  *out = nullptr;
  if (ninst < 1) return NRPY_ERR_PARAM;
  nrpy_batch *b = new nrpy_batch();
  b->ninst = ninst;
  b->h_inst.resize(ninst);
  std::vector<int> row_inst, row_dest, rowptr(1, 0), src, negmask;
  std::vector<double> coeff;
  for (int k = 0; k < ninst; k++) {
    if (!rfms[k] || !bcs[k] || nrpy_check_params(&params[k])) {
      nrpy_batch_free(b);
      return rfms[k] && bcs[k] ? NRPY_ERR_SHAPE : NRPY_ERR_NULL;
    }
    const bc_struct *bc = bcs[k];
    if (bc->N[0] != params[k].Nxx_plus_2NGHOSTS0 || bc->N[1] != params[k].Nxx_plus_2NGHOSTS1 ||
        bc->N[2] != params[k].Nxx_plus_2NGHOSTS2) {
      nrpy_batch_free(b);
      return NRPY_ERR_SHAPE;
    }
    nrpy_batch_instance &I = b->h_inst[k];
    memset(&I, 0, sizeof(I));
    I.params = params[k];
    I.rfm = *rfms[k];
    for (int r = 0; r < NRPY_BATCH_ROLES; r++) I.arr[r] = arrays[NRPY_BATCH_ROLES * k + r];
    for (int r = 0; r <= NRPY_BATCH_ROLE_RT; r++)
      if (!I.arr[r]) {
        nrpy_batch_free(b);
        return NRPY_ERR_NULL;
      }
    I.dt = dts[k];
    I.Ntot = (long)bc->N[0] * bc->N[1] * bc->N[2];
    for (int r = 0; r < bc->nghost; r++) {
      row_inst.push_back(k);
      row_dest.push_back(bc->h_gdest[r]);
      for (int t = bc->h_rowptr[r]; t < bc->h_rowptr[r + 1]; t++) {
        src.push_back(bc->h_src[t]);
        coeff.push_back(bc->h_coeff[t]);
        negmask.push_back(bc->h_negmask[t]);
      }
      rowptr.push_back((int)src.size());
    }
  }
  b->nrows = (int)row_inst.size();
  bool ok = nrpy_upload(&b->d_inst, b->h_inst.data(), b->h_inst.size()) &&
            nrpy_upload(&b->d_row_inst, row_inst.data(), row_inst.size()) &&
            nrpy_upload(&b->d_row_dest, row_dest.data(), row_dest.size()) &&
            nrpy_upload(&b->d_rowptr, rowptr.data(), rowptr.size()) &&
            nrpy_upload(&b->d_src, src.data(), src.size()) &&
            nrpy_upload(&b->d_coeff, coeff.data(), coeff.size()) &&
            nrpy_upload(&b->d_negmask, negmask.data(), negmask.size());
  if (!ok) {
    nrpy_batch_free(b);
    return NRPY_ERR_ALLOC;
  }
  int rc = nrpy_batch_prepare(b);
  if (rc) {
    nrpy_batch_free(b);
    return rc;
  }
  *out = b;
  return NRPY_OK;
}

// Composed ghost fill of the array in role 'role' of every instance, one launch. One thread
// per ghost row fills NRPY_BATCH_BC_CHUNK gridfunctions per pass over the row's terms: the
// concatenated maps of a sweep (~100 MB for 256 Brill-Lindquist instances) are then read
// ceil(num_gfs / NRPY_BATCH_BC_CHUNK) times instead of once per gridfunction. Each
// gridfunction's sum runs over the same terms in the same order as one thread per
// (row, gridfunction) would, so the values are identical.
#define NRPY_BATCH_BC_CHUNK 8
__global__ void nrpy_batch_bc_kernel(int nrows, const nrpy_batch_instance *__restrict__ inst,
                                     const int *__restrict__ row_inst, const int *__restrict__ row_dest,
                                     const int *__restrict__ rowptr, const int *__restrict__ src,
                                     const double *__restrict__ coeff, const int *__restrict__ negmask, int role,
                                     int num_gfs, nrpy_gf_parity_args gfp) {
  const int r = (int)((size_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= nrows) return;
  const int k = row_inst[r];
  double *base = inst[k].arr[role];
  const size_t Ntot = inst[k].Ntot;
  const int q0 = rowptr[r], q1 = rowptr[r + 1], dest = row_dest[r];
  for (int gf0 = 0; gf0 < num_gfs; gf0 += NRPY_BATCH_BC_CHUNK) {
    double acc[NRPY_BATCH_BC_CHUNK];
#pragma unroll
    for (int j = 0; j < NRPY_BATCH_BC_CHUNK; j++) acc[j] = 0.0;
    for (int q = q0; q < q1; q++) {
      const int s = src[q];
      const double c = coeff[q];
      const int m = negmask[q];
#pragma unroll
      for (int j = 0; j < NRPY_BATCH_BC_CHUNK; j++) {
        const int gf = gf0 + j;
        if (gf < num_gfs) {
          const double cj = (m & (1 << gfp.p[gf])) ? -c : c;
          acc[j] += cj * base[(size_t)gf * Ntot + s];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NRPY_BATCH_BC_CHUNK; j++)
      if (gf0 + j < num_gfs) base[(size_t)(gf0 + j) * Ntot + dest] = acc[j];
  }
}

extern "C" int nrpy_batch_apply_bcs(const nrpy_batch *b, const int role, const int NUM_GFS, const int8_t *gfs_parity,
                                    hipStream_t stream) {
  if (!b || !gfs_parity) return NRPY_ERR_NULL;
  if (role < 0 || role >= NRPY_BATCH_ROLES || NUM_GFS < 1 || NUM_GFS > 64) return NRPY_ERR_PARAM;
  if (b->nrows == 0) return NRPY_OK;
  nrpy_gf_parity_args gfp;
  memset(&gfp, 0, sizeof(gfp));
  for (int g = 0; g < NUM_GFS; g++) {
    if (gfs_parity[g] < 0 || gfs_parity[g] > 9) return NRPY_ERR_PARAM;
    gfp.p[g] = gfs_parity[g];
  }
  const int B = 256;
  const size_t work = (size_t)b->nrows;
  hipLaunchKernelGGL(nrpy_batch_bc_kernel, dim3((unsigned)((work + B - 1) / B)), dim3(B), 0, stream, b->nrows,
                     b->d_inst, b->d_row_inst, b->d_row_dest, b->d_rowptr, b->d_src, b->d_coeff, b->d_negmask, role,
                     NUM_GFS, gfp);
  return hipGetLastError() == hipSuccess ? NRPY_OK : NRPY_ERR_LAUNCH;
}
