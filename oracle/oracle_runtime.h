/* Oracle (test infrastructure only): boundary conditions in the reference's
 * sequential form.
 *   oracle_bc_maps        restates set_up__bc_gz_map_and_parity_condns.h:15-113
 *   oracle_set_bcstruct   restates set_bcstruct.h:32-124
 *   oracle_apply_bcs      restates apply_bcs_curvilinear.h:7-64
 * (all under CurviBoundaryConditions/boundary_conditions/). */
typedef struct { short i0, i1, i2; } o_gz_map;
typedef struct { o_gz_map dest; int8_t face[3]; } o_outer;
typedef struct { o_gz_map dest, src; int8_t parity[10]; } o_inner;
typedef struct {
  o_outer *outer[NGHOSTS];
  o_inner *inner[NGHOSTS];
  int num_ob[NGHOSTS], num_ib[NGHOSTS];
} o_bcstruct;

/* Returns 0, or 1 where the reference prints an error and exits. */
static int oracle_bc_maps(const paramstruct *params, double *xx[3], o_gz_map *map, int8_t *par10) {
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;
  const double dxx0 = params->dxx0, dxx1 = params->dxx1, dxx2 = params->dxx2;
  const double xxmin[3] = { xx[0][0] - ((double)(0-NGHOSTS) + (1.0/2.0))*dxx0,
                            xx[1][0] - ((double)(0-NGHOSTS) + (1.0/2.0))*dxx1,
                            xx[2][0] - ((double)(0-NGHOSTS) + (1.0/2.0))*dxx2 };
  for (int i2 = 0; i2 < Nxx_plus_2NGHOSTS2; i2++)
    for (int i1 = 0; i1 < Nxx_plus_2NGHOSTS1; i1++)
      for (int i0 = 0; i0 < Nxx_plus_2NGHOSTS0; i0++) {
        double xCart[3], c[3], xCart2[3];
        EigenCoord_xx_to_Cart(params, xx, i0, i1, i2, xCart);
        EigenCoord_Cart_to_xx(params, xCart, c);
        int i0_inbounds = (int)( (c[0] - xxmin[0] - (1.0/2.0)*dxx0 + ((double)NGHOSTS)*dxx0)/dxx0 + 0.5 );
        int i1_inbounds = (int)( (c[1] - xxmin[1] - (1.0/2.0)*dxx1 + ((double)NGHOSTS)*dxx1)/dxx1 + 0.5 );
        int i2_inbounds = (int)( (c[2] - xxmin[2] - (1.0/2.0)*dxx2 + ((double)NGHOSTS)*dxx2)/dxx2 + 0.5 );
        if (i0_inbounds < 0 || i0_inbounds >= Nxx_plus_2NGHOSTS0 || i1_inbounds < 0 ||
            i1_inbounds >= Nxx_plus_2NGHOSTS1 || i2_inbounds < 0 || i2_inbounds >= Nxx_plus_2NGHOSTS2) return 1;
        EigenCoord_xx_to_Cart(params, xx, i0_inbounds, i1_inbounds, i2_inbounds, xCart2);
        for (int d = 0; d < 3; d++) if (fabs(xCart[d] - xCart2[d]) > 1e-8) return 1;
        const int idx = IDX3S(i0, i1, i2);
        if (i0_inbounds == i0 && i1_inbounds == i1 && i2_inbounds == i2) {
          map[idx].i0 = map[idx].i1 = map[idx].i2 = -1;
          for (int p = 0; p < 10; p++) par10[10*idx + p] = 1;
        } else {
          map[idx].i0 = i0_inbounds; map[idx].i1 = i1_inbounds; map[idx].i2 = i2_inbounds;
          double pv[10];
          parity_dot_products(params, pv, xx[0][i0], xx[1][i1], xx[2][i2],
                              xx[0][i0_inbounds], xx[1][i1_inbounds], xx[2][i2_inbounds]);
          for (int p = 0; p < 10; p++) {
            if ((pv[p] > 0 && fabs(pv[p] - 1) > 1e-8) || (pv[p] <= 0 && fabs(pv[p] + 1) > 1e-8)) return 1;
            par10[10*idx + p] = pv[p] < 0.0 ? -1 : 1;
          }
        }
      }
  return 0;
}

static void oracle_set_bcstruct(const paramstruct *params, const o_gz_map *map, const int8_t *par10, o_bcstruct *bc) {
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;
  const int maxpts = Nxx_plus_2NGHOSTS0*Nxx_plus_2NGHOSTS1*Nxx_plus_2NGHOSTS2;
  int imin[3] = {NGHOSTS, NGHOSTS, NGHOSTS};
  int imax[3] = {Nxx_plus_2NGHOSTS0-NGHOSTS, Nxx_plus_2NGHOSTS1-NGHOSTS, Nxx_plus_2NGHOSTS2-NGHOSTS};
#define O_LOOP(a0,b0,a1,b1,a2,b2) for (int i2=(a2); i2<(b2); i2++) for (int i1=(a1); i1<(b1); i1++) for (int i0=(a0); i0<(b0); i0++)
  for (int gz = 0; gz < NGHOSTS; gz++) {
    o_outer *ob = (o_outer *)malloc(sizeof(o_outer)*maxpts);
    o_inner *ib = (o_inner *)malloc(sizeof(o_inner)*maxpts);
    int no = 0, ni = 0;
#define O_VISIT(f0,f1,f2) { const int idx = IDX3S(i0,i1,i2); \
    if (map[idx].i0 == -1) { ob[no].dest.i0=i0; ob[no].dest.i1=i1; ob[no].dest.i2=i2; \
      ob[no].face[0]=f0; ob[no].face[1]=f1; ob[no].face[2]=f2; no++; } \
    else { ib[ni].dest.i0=i0; ib[ni].dest.i1=i1; ib[ni].dest.i2=i2; ib[ni].src = map[idx]; \
      for (int p = 0; p < 10; p++) ib[ni].parity[p] = par10[10*idx+p]; ni++; } }
    O_LOOP(imin[0]-1,imin[0], imin[1],imax[1], imin[2],imax[2]) O_VISIT(1,0,0) imin[0]--;
    O_LOOP(imax[0],imax[0]+1, imin[1],imax[1], imin[2],imax[2]) O_VISIT(-1,0,0) imax[0]++;
    O_LOOP(imin[0],imax[0], imin[1]-1,imin[1], imin[2],imax[2]) O_VISIT(0,1,0) imin[1]--;
    O_LOOP(imin[0],imax[0], imax[1],imax[1]+1, imin[2],imax[2]) O_VISIT(0,-1,0) imax[1]++;
    O_LOOP(imin[0],imax[0], imin[1],imax[1], imin[2]-1,imin[2]) O_VISIT(0,0,1) imin[2]--;
    O_LOOP(imin[0],imax[0], imin[1],imax[1], imax[2],imax[2]+1) O_VISIT(0,0,-1) imax[2]++;
    bc->outer[gz] = ob; bc->inner[gz] = ib; bc->num_ob[gz] = no; bc->num_ib[gz] = ni;
  }
}

void oracle_apply_bcs(const paramstruct *params, const o_bcstruct *bc, int NUM_GFS, const int8_t *gfs_parity,
                      double *gfs) {
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;
#pragma omp parallel for
  for (int gf = 0; gf < NUM_GFS; gf++)
    for (int gz = 0; gz < NGHOSTS; gz++) {
      for (int p = 0; p < bc->num_ob[gz]; p++) {
        const o_outer *o = &bc->outer[gz][p];
        const int i0 = o->dest.i0, i1 = o->dest.i1, i2 = o->dest.i2;
        const int f0 = o->face[0], f1 = o->face[1], f2 = o->face[2];
        gfs[IDX4S(gf,i0,i1,i2)] = +3.0*gfs[IDX4S(gf,i0+1*f0,i1+1*f1,i2+1*f2)]
                                  -3.0*gfs[IDX4S(gf,i0+2*f0,i1+2*f1,i2+2*f2)]
                                  +1.0*gfs[IDX4S(gf,i0+3*f0,i1+3*f1,i2+3*f2)];
      }
      for (int p = 0; p < bc->num_ib[gz]; p++) {
        const o_inner *n = &bc->inner[gz][p];
        gfs[IDX4S(gf,n->dest.i0,n->dest.i1,n->dest.i2)] =
            n->parity[gfs_parity[gf]] * gfs[IDX4S(gf,n->src.i0,n->src.i1,n->src.i2)];
      }
    }
}

/* Build maps + bcstruct; returns NULL on a map error. */
o_bcstruct *oracle_bcstruct_create(const paramstruct *params, double *xx[3]) {
  const size_t n = (size_t)params->Nxx_plus_2NGHOSTS0*params->Nxx_plus_2NGHOSTS1*params->Nxx_plus_2NGHOSTS2;
  o_gz_map *map = (o_gz_map *)malloc(sizeof(o_gz_map)*n);
  int8_t *par10 = (int8_t *)malloc(10*n);
  o_bcstruct *bc = NULL;
  if (oracle_bc_maps(params, xx, map, par10) == 0) {
    bc = (o_bcstruct *)calloc(1, sizeof(o_bcstruct));
    oracle_set_bcstruct(params, map, par10, bc);
  }
  free(map); free(par10);
  return bc;
}
void oracle_bcstruct_free(o_bcstruct *bc) {
  if (!bc) return;
  for (int g = 0; g < NGHOSTS; g++) { free(bc->outer[g]); free(bc->inner[g]); }
  free(bc);
}
int oracle_bcstruct_counts(const o_bcstruct *bc, int *num_ob, int *num_ib) {
  for (int g = 0; g < NGHOSTS; g++) { num_ob[g] = bc->num_ob[g]; num_ib[g] = bc->num_ib[g]; }
  return NGHOSTS;
}
/* Flatten layer gz: outer as (i0,i1,i2,f0,f1,f2), inner as (d0,d1,d2,s0,s1,s2,parity[10]). */
void oracle_bcstruct_layer(const o_bcstruct *bc, int gz, int *outer6, int *inner16) {
  for (int p = 0; p < bc->num_ob[gz]; p++) {
    const o_outer *o = &bc->outer[gz][p];
    int *r = outer6 + 6*p; r[0]=o->dest.i0; r[1]=o->dest.i1; r[2]=o->dest.i2; r[3]=o->face[0]; r[4]=o->face[1]; r[5]=o->face[2];
  }
  for (int p = 0; p < bc->num_ib[gz]; p++) {
    const o_inner *n = &bc->inner[gz][p];
    int *r = inner16 + 16*p; r[0]=n->dest.i0; r[1]=n->dest.i1; r[2]=n->dest.i2; r[3]=n->src.i0; r[4]=n->src.i1; r[5]=n->src.i2;
    for (int k = 0; k < 10; k++) r[6+k] = n->parity[k];
  }
}
int oracle_NGHOSTS(void) { return NGHOSTS; }
int oracle_sizeof_paramstruct(void) { return (int)sizeof(paramstruct); }

/* find_timestep (reference_metric.py:1412-1439): CFL_FACTOR * min over interior of scale factor * dxx,
 * divided by the wave speed. */
double oracle_find_timestep(const paramstruct *params, double *xx[3], double CFL_FACTOR) {
  double dsmin = 1e38;
  for (int i2 = NGHOSTS; i2 < params->Nxx_plus_2NGHOSTS2 - NGHOSTS; i2++)
    for (int i1 = NGHOSTS; i1 < params->Nxx_plus_2NGHOSTS1 - NGHOSTS; i1++)
      for (int i0 = NGHOSTS; i0 < params->Nxx_plus_2NGHOSTS0 - NGHOSTS; i0++) {
        double s[3];
        scale_factors(params, xx[0][i0], xx[1][i1], xx[2][i2], s);
        const double d0 = s[0]*params->dxx0, d1 = s[1]*params->dxx1, d2 = s[2]*params->dxx2;
        dsmin = fmin(dsmin, fmin(d0, fmin(d1, d2)));
      }
  return dsmin*CFL_FACTOR/ORACLE_WAVESPEED(params);
}

#ifdef ORACLE_HAS_SOMMERFELD
/* Sommerfeld radiation condition (CurviBoundaryConditions.py:259-658): d f/d r at a point
 * from one-sided stencils along directions where the point is on a face (FACE != 0) or
 * within NGHOSTS of the array edge, centred stencils otherwise, contracted with d x^i/d r
 * by oracle_r_and_dfdr (generated per coordinate system). Coefficients: one-sided
 * fd2 (-3/2, 2, -1/2), fd4 (-25/12, 4, -3, 4/3, -1/4), fd6 (-49/20, 6, -15/2, 20/3,
 * -15/4, 6/5, -1/6); centred fd2 (1/2), fd4 (2/3, 1/12), fd6 (3/4, 3/20, 1/60). */
static void o_sommerfeld_point(const paramstruct *params, double *xx[3], const double *f, const int8_t face[3],
                               int i0, int i1, int i2, int fd_order, double *r, double *dfdr) {
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int N[3] = {params->Nxx_plus_2NGHOSTS0, params->Nxx_plus_2NGHOSTS1, params->Nxx_plus_2NGHOSTS2};
  const double invdx[3] = {params->invdx0, params->invdx1, params->invdx2};
  static const double one2[3] = {-1.5, 2.0, -0.5};
  static const double one4[5] = {-25.0/12.0, 4.0, -3.0, 4.0/3.0, -1.0/4.0};
  static const double one6[7] = {-49.0/20.0, 6.0, -15.0/2.0, 20.0/3.0, -15.0/4.0, 6.0/5.0, -1.0/6.0};
  const double *one = fd_order == 2 ? one2 : (fd_order == 4 ? one4 : one6);
  const int ii[3] = {i0, i1, i2};
  double fdD[3];
  for (int d = 0; d < 3; d++) {
    int e[3] = {0, 0, 0};
    e[d] = 1;
#define O_F(k) f[IDX3S(i0 + (k)*e[0], i1 + (k)*e[1], i2 + (k)*e[2])]
    if (abs(face[d]) == 1 || ii[d] + NGHOSTS >= N[d] || ii[d] - NGHOSTS <= 0) {
      int S = face[d];
      if (ii[d] + NGHOSTS >= N[d]) S = -1;
      if (ii[d] - NGHOSTS <= 0) S = +1;
      double sum = 0.0;
      for (int k = 0; k <= fd_order; k++) sum = (k == 0) ? one[0]*O_F(0) : sum + one[k]*O_F(k*S);
      fdD[d] = S*sum*invdx[d];
    } else if (fd_order == 2) {
      fdD[d] = (O_F(1) - O_F(-1))*0.5*invdx[d];
    } else if (fd_order == 4) {
      const double c1 = 2.0/3.0, c2 = 1.0/12.0;
      fdD[d] = (-c2*O_F(2) + c1*O_F(1) - c1*O_F(-1) + c2*O_F(-2))*invdx[d];
    } else {
      const double c1 = 3.0/4.0, c2 = 3.0/20.0, c3 = 1.0/60.0;
      fdD[d] = (c3*O_F(3) - c2*O_F(2) + c1*O_F(1) - c1*O_F(-1) + c2*O_F(-2) - c3*O_F(-3))*invdx[d];
    }
#undef O_F
  }
  oracle_r_and_dfdr(params, xx[0][i0], xx[1][i1], xx[2][i2], fdD, r, dfdr);
}

/* apply_bcs_sommerfeld (CurviBoundaryConditions.py:538-620) in the reference's order:
 * per gridfunction, per layer from the innermost, outer points in list order, then the
 * parity copies of the RHS at inner points. Per-gridfunction parameters are arguments. */
void oracle_apply_bcs_sommerfeld(const paramstruct *params, double *xx[3], const o_bcstruct *bc, int NUM_GFS,
                                 const int8_t *gfs_parity, const double *gfs, double *rhs_gfs, int fd_order,
                                 const double *at_inf, const double *falloff_power, const double *speed) {
  const int Nxx_plus_2NGHOSTS0 = params->Nxx_plus_2NGHOSTS0;
  const int Nxx_plus_2NGHOSTS1 = params->Nxx_plus_2NGHOSTS1;
  const int Nxx_plus_2NGHOSTS2 = params->Nxx_plus_2NGHOSTS2;
  const size_t Ntot = (size_t)Nxx_plus_2NGHOSTS0*Nxx_plus_2NGHOSTS1*Nxx_plus_2NGHOSTS2;
#pragma omp parallel for
  for (int gf = 0; gf < NUM_GFS; gf++) {
    const double *f = gfs + gf*Ntot;
    double *rhs = rhs_gfs + gf*Ntot;
    for (int gz = 0; gz < NGHOSTS; gz++) {
      for (int p = 0; p < bc->num_ob[gz]; p++) {
        const o_outer *o = &bc->outer[gz][p];
        const int i0 = o->dest.i0, i1 = o->dest.i1, i2 = o->dest.i2;
        double r = 1e100, dfdr = 1e100;
        o_sommerfeld_point(params, xx, f, o->face, i0, i1, i2, fd_order, &r, &dfdr);
        const double invr = 1./r;
        const double source_rhs = -speed[gf]*(dfdr + invr*(f[IDX3S(i0,i1,i2)] - at_inf[gf]));
        rhs[IDX3S(i0,i1,i2)] = source_rhs;
        if (falloff_power[gf] > 0) {
          const int j0 = i0 + o->face[0], j1 = i1 + o->face[1], j2 = i2 + o->face[2];
          double r_off = 1e100, dfdr_off = 1e100;
          o_sommerfeld_point(params, xx, f, o->face, j0, j1, j2, fd_order, &r_off, &dfdr_off);
          const double invr_off = 1./r_off;
          const double extrap_rhs = speed[gf]*(dfdr_off + invr_off*(f[IDX3S(j0,j1,j2)] - at_inf[gf]));
          const double diff = rhs[IDX3S(j0,j1,j2)] + extrap_rhs;
          rhs[IDX3S(i0,i1,i2)] += diff*pow(r_off*invr, falloff_power[gf]);
        }
      }
      for (int p = 0; p < bc->num_ib[gz]; p++) {
        const o_inner *n = &bc->inner[gz][p];
        rhs[IDX3S(n->dest.i0,n->dest.i1,n->dest.i2)] =
            n->parity[gfs_parity[gf]] * rhs[IDX3S(n->src.i0,n->src.i1,n->src.i2)];
      }
    }
  }
}
#endif
