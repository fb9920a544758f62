/* Oracle (test infrastructure only): one RK4 step in the order of the
 * reference's generated RK_MoL.h (MoLtimestepping/C_Code_Generation.py:266-329,
 * RK4 table at RK_Butcher_Table_Dictionary.py:86-91) with the post-RHS calls of
 * Tutorial-Start_to_Finish-BSSNCurvilinear-Two_BHs_Collide.ipynb:213-219:
 * RHS into k, update loop over ALL points (ghost values are overwritten by the
 * boundary conditions that follow), apply_bcs_curvilinear, then (BSSN)
 * enforce_detgammabar_constraint. */
#ifndef ORACLE_HAS_RICCI
#define ORACLE_HAS_RICCI 0
#endif
#ifndef ORACLE_HAS_DETG
#define ORACLE_HAS_DETG 0
#endif
static const int8_t oracle_evol_parity_dummy = 0;
static void o_post(const paramstruct *params, double *xx[3], const o_bcstruct *bc, const int8_t *parity, double *gfs) {
  oracle_apply_bcs(params, bc, NUM_EVOL_GFS, parity, gfs);
  if (ORACLE_HAS_DETG) oracle_enforce_detgammabar_constraint(params, xx, gfs);
}
void oracle_MoL_step(const paramstruct *params, double *xx[3], const o_bcstruct *bc, const int8_t *evol_parity,
                     double *y_n_gfs, double *k_odd_gfs, double *k_even_gfs, double *y_nplus1_running_total_gfs,
                     double *auxevol_gfs, const double dt) {
  const size_t N = (size_t)params->Nxx_plus_2NGHOSTS0*params->Nxx_plus_2NGHOSTS1*params->Nxx_plus_2NGHOSTS2*NUM_EVOL_GFS;
  double *rt = y_nplus1_running_total_gfs;
  (void)oracle_evol_parity_dummy;
  /* k1 */
  if (ORACLE_HAS_RICCI) oracle_Ricci_eval(params, xx, y_n_gfs, auxevol_gfs);
  oracle_rhs_eval(params, xx, auxevol_gfs, y_n_gfs, k_odd_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    rt[i] = k_odd_gfs[i]*dt*(1.0/6.0);
    k_odd_gfs[i] = y_n_gfs[i] + k_odd_gfs[i]*dt*(1.0/2.0);
  }
  o_post(params, xx, bc, evol_parity, k_odd_gfs);
  /* k2 */
  if (ORACLE_HAS_RICCI) oracle_Ricci_eval(params, xx, k_odd_gfs, auxevol_gfs);
  oracle_rhs_eval(params, xx, auxevol_gfs, k_odd_gfs, k_even_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    rt[i] = rt[i] + k_even_gfs[i]*dt*(1.0/3.0);
    k_even_gfs[i] = y_n_gfs[i] + k_even_gfs[i]*dt*(1.0/2.0);
  }
  o_post(params, xx, bc, evol_parity, k_even_gfs);
  /* k3 */
  if (ORACLE_HAS_RICCI) oracle_Ricci_eval(params, xx, k_even_gfs, auxevol_gfs);
  oracle_rhs_eval(params, xx, auxevol_gfs, k_even_gfs, k_odd_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    rt[i] = rt[i] + k_odd_gfs[i]*dt*(1.0/3.0);
    k_odd_gfs[i] = y_n_gfs[i] + k_odd_gfs[i]*dt;
  }
  o_post(params, xx, bc, evol_parity, k_odd_gfs);
  /* k4 */
  if (ORACLE_HAS_RICCI) oracle_Ricci_eval(params, xx, k_odd_gfs, auxevol_gfs);
  oracle_rhs_eval(params, xx, auxevol_gfs, k_odd_gfs, k_even_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    y_n_gfs[i] = y_n_gfs[i] + rt[i] + k_even_gfs[i]*dt*(1.0/6.0);
  }
  o_post(params, xx, bc, evol_parity, y_n_gfs);
}

#ifdef ORACLE_HAS_SOMMERFELD
/* One RK4 step of the Start-to-Finish drivers' Sommerfeld branch
 * (Tutorial-Start_to_Finish-Solving_Maxwells_Equations_in_Vacuum-Cartesian.ipynb:240-246):
 * RHS_string = rhs_eval + apply_bcs_sommerfeld, post_RHS_string empty; the update loop
 * runs over all points, so ghost values evolve with the radiation-condition RHS. */
void oracle_MoL_step_sommerfeld(const paramstruct *params, double *xx[3], const o_bcstruct *bc,
                                const int8_t *evol_parity, double *y_n_gfs, double *k_odd_gfs, double *k_even_gfs,
                                double *y_nplus1_running_total_gfs, const double dt) {
  const size_t N = (size_t)params->Nxx_plus_2NGHOSTS0*params->Nxx_plus_2NGHOSTS1*params->Nxx_plus_2NGHOSTS2*NUM_EVOL_GFS;
  double *rt = y_nplus1_running_total_gfs;
#define O_SOM(IN, OUT) oracle_apply_bcs_sommerfeld(params, xx, bc, NUM_EVOL_GFS, evol_parity, IN, OUT, \
    oracle_sommerfeld_fd_order, oracle_sommerfeld_at_inf, oracle_sommerfeld_power, oracle_sommerfeld_speed)
  oracle_rhs_eval(params, xx, NULL, y_n_gfs, k_odd_gfs);
  O_SOM(y_n_gfs, k_odd_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    rt[i] = k_odd_gfs[i]*dt*(1.0/6.0);
    k_odd_gfs[i] = y_n_gfs[i] + k_odd_gfs[i]*dt*(1.0/2.0);
  }
  oracle_rhs_eval(params, xx, NULL, k_odd_gfs, k_even_gfs);
  O_SOM(k_odd_gfs, k_even_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    rt[i] = rt[i] + k_even_gfs[i]*dt*(1.0/3.0);
    k_even_gfs[i] = y_n_gfs[i] + k_even_gfs[i]*dt*(1.0/2.0);
  }
  oracle_rhs_eval(params, xx, NULL, k_even_gfs, k_odd_gfs);
  O_SOM(k_even_gfs, k_odd_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    rt[i] = rt[i] + k_odd_gfs[i]*dt*(1.0/3.0);
    k_odd_gfs[i] = y_n_gfs[i] + k_odd_gfs[i]*dt;
  }
  oracle_rhs_eval(params, xx, NULL, k_odd_gfs, k_even_gfs);
  O_SOM(k_odd_gfs, k_even_gfs);
#pragma omp parallel for
  for (size_t i = 0; i < N; i++) {
    y_n_gfs[i] = y_n_gfs[i] + rt[i] + k_even_gfs[i]*dt*(1.0/6.0);
  }
#undef O_SOM
}
#endif
