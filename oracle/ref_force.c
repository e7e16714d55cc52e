/* ref_force.c -- exposes MILC's HISQ force level 0 (outer products) to the
 * tests.  TEST INFRASTRUCTURE ONLY, built by Makefile.ref.
 *
 * fn_fermion_force_multi_hisq_smearing0 is file-static in
 * generic_ks/fermion_force_hisq_multi.c (lines 1478-1636), so this
 * translation unit includes that unmodified source file and adds a wrapper. */
#include "fermion_force_hisq_multi.c"

/* oprod1/oprod3: MILC link layout [i][dir][18]; coefficient 1 (the caller's
 * 2*eps*parity factor is applied by MILC after the smearing chain) */
void ref_force_oprod(double **multi_x, const double *res, int n, double *o1, double *o3) {
  info_t info = INFO_ZERO;
  Real r[64];
  su3_matrix *fa[4], *fn[4];
  size_t i;
  int dir, t;
  for (t = 0; t < n && t < 64; t++) r[t] = res[t];
  for (dir = 0; dir < 4; dir++) {
    fa[dir] = (su3_matrix *)malloc(sites_on_node * sizeof(su3_matrix));
    fn[dir] = (su3_matrix *)malloc(sites_on_node * sizeof(su3_matrix));
  }
  fn_fermion_force_multi_hisq_smearing0(&info, 1.0, r, (su3_vector **)multi_x, n, fa, fn);
  for (i = 0; i < sites_on_node; i++)
    for (dir = 0; dir < 4; dir++) {
      memcpy(o1 + (4 * i + dir) * 18, &fa[dir][i], 18 * sizeof(double));
      memcpy(o3 + (4 * i + dir) * 18, &fn[dir][i], 18 * sizeof(double));
    }
  for (dir = 0; dir < 4; dir++) {
    free(fa[dir]);
    free(fn[dir]);
  }
}

fermion_links_t *ref_get_fl(void);
int ref_hisq_force_naik(double **multi_x, const double *res, int n0, int n1, double eps, double *mom);

/* MILC's full HISQ thin-link force (eo_fermion_force_multi ->
 * fn_fermion_force_multi_hisq_wrapper_mx_cpu with KS_MULTIFF = FNMAT and
 * HISQ_FF_MULTI_WRAPPER) on the links of the last ref_hisq_links call.
 * multi_x[t] = X_t on all sites, one Naik set; mom: the anti_hermitmat
 * added to a zero momentum, [i][dir][10] doubles. */
int ref_hisq_force(double **multi_x, const double *res, int n, double eps, double *mom) {
  return ref_hisq_force_naik(multi_x, res, n, 0, eps, mom);
}

/* ... with the last n1 of the n0 + n1 terms in the second Naik set (links of
 * the last ref_hisq_links_naik call, or n1 = 0) */
int ref_hisq_force_naik(double **multi_x, const double *res, int n0, int n1, double eps, double *mom) {
  fermion_links_t *fl = ref_get_fl();
  Real r[64];
  size_t i;
  int dir, t, n = n0 + n1;
  if (!fl || n > 64 || n1 < 0 || (n1 > 0 && fermion_links_get_n_naiks(fl) != 2)) return -1;
  for (t = 0; t < n; t++) r[t] = res[t];
  n_order_naik_total = n;
  n_orders_naik[0] = n0;
  n_orders_naik[1] = n1;
  for (i = 0; i < sites_on_node; i++)
    for (dir = 0; dir < 4; dir++) memset(&lattice[i].mom[dir], 0, sizeof(anti_hermitmat));
  eo_fermion_force_multi(eps, r, (su3_vector **)multi_x, n, MILC_PRECISION, fl);
  for (i = 0; i < sites_on_node; i++)
    for (dir = 0; dir < 4; dir++) memcpy(mom + (4 * i + dir) * 10, &lattice[i].mom[dir], 10 * sizeof(double));
  return 0;
}
