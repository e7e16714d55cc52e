/* ref_driver.c -- thin driver around MILC 7.8.1's own CPU routines.
 *
 * TEST INFRASTRUCTURE ONLY.  Makefile.ref compiles this file together with
 * unmodified MILC sources taken from the reference tree into
 * oracle/_ref/libmilc_ref.so.  It does what the ks_imp_rhmc application's
 * setup code does (ks_imp_rhmc/setup.c:128-152: layout, lattice, nearest and
 * third-nearest gathers, KS phases) and exposes MILC's Dslash, CG,
 * multi-shift CG, HISQ link construction and force level 0 to ctypes, with
 * fields in MILC's own layout.  It is used as a checker and as the CPU
 * baseline, never by the product.
 */
#define CONTROL
#include "ks_imp_includes.h"

#include <string.h>

static fermion_links_t *g_fl = NULL; /* links built by MILC from the site links */
static fn_links_t *g_ext = NULL;     /* links supplied by the caller */
static fn_links_t *g_cur = NULL;     /* links used by the solver entry points */

/* third-nearest-neighbour gathers, the same pattern ks_imp_rhmc/setup.c
 * builds for the Naik term */
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

int ref_init(int x, int y, int z, int t) {
  static int done = 0;
  int argc = 1;
  char *argv0 = "ref";
  char **argv = &argv0;
  int i;
  if (done) return -1; /* MILC's lattice globals are set up once per process */
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
  fermion_links_want_back(1); /* DBLSTORE_FN Dslash needs the back links */
  done = 1;
  return 0;
}

size_t ref_sites(void) { return sites_on_node; }

/* thin links without phases, MILC order [i][dir][18] -> site links, phases in */
void ref_set_thin_links(const double *U) {
  size_t i;
  int dir;
  if (phases_in == ON) rephase(OFF);
  for (i = 0; i < sites_on_node; i++)
    for (dir = XUP; dir <= TUP; dir++) memcpy(&lattice[i].link[dir], U + (4 * i + dir) * 18, 18 * sizeof(double));
  rephase(ON);
}

/* MILC's HISQ link construction from the site links (create_fermion_links_from_site) */
int ref_hisq_links(double *fat, double *lng) {
  double eps_naik[1] = {0.0};
  imp_ferm_links_t *fn;
  if (g_fl) destroy_fermion_links_hisq(g_fl);
  g_fl = create_fermion_links_from_site(MILC_PRECISION, 1, eps_naik);
  fn = get_fm_links(g_fl)[0];
  memcpy(fat, get_fatlinks(fn), sizeof(su3_matrix) * 4 * sites_on_node);
  memcpy(lng, get_lnglinks(fn), sizeof(su3_matrix) * 4 * sites_on_node);
  g_cur = fn;
  return 0;
}

/* the same with a second Naik set (n_naiks = 2, eps_naik = {0, eps}, as a
 * 2+1+1 run with a charm quark): set 0 into fat/lng, set 1 into fat1/lng1
 * (load_hisq_fn_links, fermion_links_hisq_load_milc.c:574-640) */
int ref_hisq_links_naik(double eps, double *fat, double *lng, double *fat1, double *lng1) {
  double eps_naik[2] = {0.0, eps};
  imp_ferm_links_t **fn;
  if (g_fl) destroy_fermion_links_hisq(g_fl);
  g_fl = create_fermion_links_from_site(MILC_PRECISION, 2, eps_naik);
  fn = get_fm_links(g_fl);
  memcpy(fat, get_fatlinks(fn[0]), sizeof(su3_matrix) * 4 * sites_on_node);
  memcpy(lng, get_lnglinks(fn[0]), sizeof(su3_matrix) * 4 * sites_on_node);
  memcpy(fat1, get_fatlinks(fn[1]), sizeof(su3_matrix) * 4 * sites_on_node);
  memcpy(lng1, get_lnglinks(fn[1]), sizeof(su3_matrix) * 4 * sites_on_node);
  g_cur = fn[0];
  return 0;
}

/* use caller-supplied fat/long links (MILC layout) for the solver entry points */
int ref_set_fn_links(const double *fat, const double *lng) {
  if (!g_ext) g_ext = create_fn_links();
  memcpy(g_ext->fat, fat, sizeof(su3_matrix) * 4 * sites_on_node);
  memcpy(g_ext->lng, lng, sizeof(su3_matrix) * 4 * sites_on_node);
  load_fn_backlinks(g_ext);
  g_cur = g_ext;
  return 0;
}

int ref_dslash(const double *src, double *dst, int parity) {
  if (!g_cur) return -1;
  dslash_fn_field((su3_vector *)src, (su3_vector *)dst, parity, g_cur);
  return 0;
}

/* ks_congrad_parity_cpu (d_congrad5_fn_milc.c:60): dest is the initial guess */
int ref_congrad(const double *src, double *dest, int parity, double mass, int niter, int nrestart, double resid,
                double *final_rsq, int *restarts, int *converged) {
  quark_invert_control qic;
  int it;
  if (!g_cur) return -1;
  memset(&qic, 0, sizeof qic);
  qic.prec = MILC_PRECISION;
  qic.max = niter;
  qic.nrestart = nrestart;
  qic.parity = parity;
  qic.resid = resid;
  qic.relresid = 0;
  it = ks_congrad_parity_cpu((su3_vector *)src, (su3_vector *)dest, &qic, mass, g_cur);
  if (final_rsq) *final_rsq = qic.final_rsq;
  if (restarts) *restarts = qic.final_restart;
  if (converged) *converged = qic.converged;
  return it;
}

/* the same with MILC's Fermilab relative residual test (qic.relresid,
 * d_congrad5_fn_milc.c:37-56,92); *final_relrsq = qic.final_relrsq */
int ref_congrad_rel(const double *src, double *dest, int parity, double mass, int niter, int nrestart,
                    double resid, double relresid, double *final_rsq, double *final_relrsq, int *converged) {
  quark_invert_control qic;
  int it;
  if (!g_cur) return -1;
  memset(&qic, 0, sizeof qic);
  qic.prec = MILC_PRECISION;
  qic.max = niter;
  qic.nrestart = nrestart;
  qic.parity = parity;
  qic.resid = resid;
  qic.relresid = relresid;
  it = ks_congrad_parity_cpu((su3_vector *)src, (su3_vector *)dest, &qic, mass, g_cur);
  if (final_rsq) *final_rsq = qic.final_rsq;
  if (final_relrsq) *final_relrsq = qic.final_relrsq;
  if (converged) *converged = qic.converged;
  return it;
}

/* ks_multicg_offset_field_cpu (ks_multicg_offset.c:63) */
int ref_multicg(const double *src, double **psim, const double *offsets, int n, int parity, int niter, double resid,
                double *final_rsq) {
  ks_param ksp[64];
  quark_invert_control qic[64];
  int j, it;
  if (!g_cur || n > 64) return -1;
  for (j = 0; j < n; j++) {
    memset(&ksp[j], 0, sizeof ksp[j]);
    memset(&qic[j], 0, sizeof qic[j]);
    ksp[j].offset = offsets[j];
    qic[j].prec = MILC_PRECISION;
    qic[j].max = niter;
    qic[j].nrestart = 1;
    qic[j].parity = parity;
    qic[j].resid = resid;
  }
  it = ks_multicg_offset_field_cpu((su3_vector *)src, (su3_vector **)psim, ksp, n, qic, g_cur);
  if (final_rsq) *final_rsq = qic[0].final_rsq;
  return it;
}

/* the HISQ fermion links built by ref_hisq_links (for the force entry point) */
fermion_links_t *ref_get_fl(void) { return g_fl; }
