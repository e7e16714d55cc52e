/* glue_driver.c -- drives MILC's own solver and force callers in a MILC build
 * that takes its GPU offload from the MI355X library.
 *
 * TEST INFRASTRUCTURE ONLY.  Makefile.glue compiles this file with unmodified
 * MILC 7.8.1 sources from the reference tree in four variants: MILC_PRECISION
 * 1 and 2, each once as MILC's CPU build and once with -DUSE_CG_GPU plus the
 * glue files of integration/ (which take the place of d_congrad5_fn_gpu.o,
 * ks_multicg_offset_gpu.o and fermion_force_hisq_multi.o) linked against
 * libmilc_hip.so.  The entry points below call MILC's ks_congrad_field
 * (d_congrad5_fn.c:16-36), ks_multicg_field (ks_multicg.c:65-102) and
 * eo_fermion_force_multi on the HISQ links MILC builds itself, so that a test
 * can run the same calls through both builds and compare.  Fields cross this
 * interface in MILC's layout and in the build's Real (float or double).
 */
#define CONTROL
#include "ks_imp_includes.h"

#include <string.h>

static fermion_links_t *g_fl = NULL;

static void third_nbr(int x, int y, int z, int t, int *dirpt, int FB, int *xp, int *yp, int *zp, int *tp) {
  int dir = (FB == FORWARDS) ? *dirpt : OPP_DIR(*dirpt);
  *xp = x; *yp = y; *zp = z; *tp = t;
  switch (dir) {
    case XUP: *xp = (x + 3) % nx; break;
    case XDOWN: *xp = (x + 4 * nx - 3) % nx; break;
    case YUP: *yp = (y + 3) % ny; break;
    case YDOWN: *yp = (y + 4 * ny - 3) % ny; break;
    case ZUP: *zp = (z + 3) % nz; break;
    case ZDOWN: *zp = (z + 4 * nz - 3) % nz; break;
    case TUP: *tp = (t + 3) % nt; break;
    case TDOWN: *tp = (t + 4 * nt - 3) % nt; break;
    default: printf("third_nbr: bad direction\n"); exit(1);
  }
}

/* lattice, gathers and KS phases as ks_imp_rhmc/setup.c:128-152 does */
int drv_init(int x, int y, int z, int t) {
  static int done = 0;
  int argc = 1;
  char *argv0 = "glue";
  char **argv = &argv0;
  int i;
  if (done) return -1;
  initialize_machine(&argc, &argv);
  nx = x; ny = y; nz = z; nt = t;
  number_of_nodes = numnodes();
  volume = (size_t)nx * ny * nz * nt;
  total_iters = 0;
  setup_layout();
  this_node = mynode();
  make_lattice();
  make_nn_gathers();
  for (i = XUP; i <= TUP; i++) make_gather(third_nbr, &i, WANT_INVERSE, ALLOW_EVEN_ODD, SWITCH_PARITY);
  sort_eight_gathers(X3UP);
  phaseset();
  phases_in = OFF;
  fermion_links_want_back(1);
  done = 1;
  return 0;
}

size_t drv_sites(void) { return sites_on_node; }
int drv_real_bytes(void) { return (int)sizeof(Real); }

/* thin links without phases, [i][dir][18] doubles -> site links, phases in */
void drv_set_thin_links(const double *U) {
  size_t i;
  int dir, k;
  if (phases_in == ON) rephase(OFF);
  for (i = 0; i < sites_on_node; i++)
    for (dir = XUP; dir <= TUP; dir++) {
      Real *m = (Real *)&lattice[i].link[dir];
      for (k = 0; k < 18; k++) m[k] = (Real)U[(4 * i + dir) * 18 + k];
    }
  rephase(ON);
}

/* MILC's HISQ links from the site links (one Naik set); fat/lng (may be
 * NULL) receive them as doubles */
int drv_hisq_links(double *fat, double *lng) {
  double eps_naik[1] = {0.0};
  imp_ferm_links_t *fn;
  size_t i;
  if (g_fl) destroy_fermion_links_hisq(g_fl);
  g_fl = create_fermion_links_from_site(MILC_PRECISION, 1, eps_naik);
  fn = get_fm_links(g_fl)[0];
  for (i = 0; i < 4 * sites_on_node * 18; i++) {
    if (fat) fat[i] = ((Real *)get_fatlinks(fn))[i];
    if (lng) lng[i] = ((Real *)get_lnglinks(fn))[i];
  }
  return 0;
}

/* ks_congrad_field on MILC's HISQ links; out = {final_rsq, final_relrsq,
 * converged, final_iters, final_restart} */
int drv_congrad_field(void *src, void *dest, int parity, double mass, int niter, int nrestart, double resid,
                      double relresid, int prec, double *out) {
  quark_invert_control qic;
  int it;
  if (!g_fl) return -1;
  memset(&qic, 0, sizeof qic);
  qic.prec = prec;
  qic.max = niter;
  qic.nrestart = nrestart;
  qic.parity = parity;
  qic.nsrc = 1;
  qic.resid = resid;
  qic.relresid = relresid;
  it = ks_congrad_field((su3_vector *)src, (su3_vector *)dest, &qic, (Real)mass, get_fm_links(g_fl)[0]);
  if (out) {
    out[0] = qic.final_rsq;
    out[1] = qic.final_relrsq;
    out[2] = qic.converged;
    out[3] = qic.final_iters;
    out[4] = qic.final_restart;
  }
  return it;
}

/* ks_multicg_field (KS_MULTICG = HYBRID) on MILC's HISQ links, every shift
 * on the same links; out[j] = qic[j].final_rsq */
int drv_multicg_field(void *src, void **psim, const double *offsets, int n, int parity, int niter, double resid,
                      int prec, double *out) {
  ks_param ksp[64];
  quark_invert_control qic[64];
  imp_ferm_links_t *fn[64];
  int j, it;
  if (!g_fl || n > 64) return -1;
  for (j = 0; j < n; j++) {
    memset(&ksp[j], 0, sizeof ksp[j]);
    memset(&qic[j], 0, sizeof qic[j]);
    ksp[j].offset = offsets[j];
    ksp[j].residue = 1.0;
    qic[j].prec = prec;
    qic[j].max = niter;
    qic[j].nrestart = 5;
    qic[j].parity = parity;
    qic[j].nsrc = 1;
    qic[j].resid = resid;
    fn[j] = get_fm_links(g_fl)[0];
  }
  it = ks_multicg_field((su3_vector *)src, (su3_vector **)psim, ksp, n, qic, fn);
  if (out)
    for (j = 0; j < n; j++) out[j] = qic[j].final_rsq;
  return it;
}

/* eo_fermion_force_multi on MILC's HISQ links (one Naik set): multi_x[t] on
 * all sites; mom gets the increment of a zero momentum, [i][dir][10] doubles */
int drv_force(void **multi_x, const double *res, int n, double eps, double *mom) {
  Real r[64];
  size_t i;
  int dir, t, k;
  if (!g_fl || n > 64) return -1;
  for (t = 0; t < n; t++) r[t] = (Real)res[t];
  n_order_naik_total = n;
  n_orders_naik[0] = n;
  for (i = 0; i < sites_on_node; i++)
    for (dir = 0; dir < 4; dir++) memset(&lattice[i].mom[dir], 0, sizeof(anti_hermitmat));
  eo_fermion_force_multi((Real)eps, r, (su3_vector **)multi_x, n, MILC_PRECISION, g_fl);
  for (i = 0; i < sites_on_node; i++)
    for (dir = 0; dir < 4; dir++)
      for (k = 0; k < 10; k++) mom[(4 * i + dir) * 10 + k] = ((Real *)&lattice[i].mom[dir])[k];
  return 0;
}
