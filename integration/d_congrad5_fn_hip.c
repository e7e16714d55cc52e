/* d_congrad5_fn_hip.c -- MILC's single-mass staggered CG on the MI355X
 * library (libmilc_hip.so): replaces generic_ks/d_congrad5_fn_gpu.c.
 *
 * Built inside a MILC 7.8.1 application with -DUSE_CG_GPU, so that
 * ks_congrad_parity / ks_congrad_block_parity name the _gpu functions below
 * (include/imp_ferm_links.h:73-77) and ks_congrad_field, ks_congrad_site,
 * mat_invert_* and ks_multicg_hybrid_field reach them unchanged.
 *
 * Mixed precision follows MILC's make options as the QUDA glue does
 * (d_congrad5_fn_gpu.c:104-110): MAX_MIXED = 16-bit sloppy iterations,
 * HALF_MIXED = FP32 sloppy iterations, both with FP64 reliable updates.
 */
#include "generic_ks_includes.h"
#include <math.h>
#include <string.h>

#include "milc_hip.h"
#include "milc_hip_glue.h"

#ifdef SINGLE_FOR_DOUBLE /* backward compatibility, d_congrad5_fn_gpu.c:24-27 */
#define HALF_MIXED
#endif

#ifdef CGTIME
static const char *prec_label[2] = {"F", "D"};
#endif

int milc_hip_glue_init(void) {
  static int done = 0;
  if (done) return 0;
  if (numnodes() != 1) {
    node0_printf("milc_hip_glue_init: the MI355X library runs one lattice on one GPU (%d ranks)\n", numnodes());
    terminate(1);
  }
  int dims[4] = {nx, ny, nz, nt};
  int rc = milc_hip_init(dims, 0);
  if (rc < 0) milc_hip_glue_fail("milc_hip_init", rc);
  done = 1;
  return 0;
}

void milc_hip_glue_fail(const char *what, int rc) {
  printf("%s: MI355X library returned %d\n", what, rc);
  fflush(stdout);
  terminate(1);
}

static int glue_mixed(void) {
#if defined(MAX_MIXED)
  return 2;
#elif defined(HALF_MIXED)
  return 1;
#else
  return 0;
#endif
}

static double glue_tadpole(void) {
#if (FERM_ACTION == HISQ)
  return 1.0;
#else
  return u0;
#endif
}

/* sites of one parity: even sites first (FORSOMEFIELDPARITY, include/macros.h) */
void milc_hip_glue_parity_range(int parity, int *lo, int *hi) {
  *lo = parity == ODD ? (int)even_sites_on_node : 0;
  *hi = parity == EVEN ? (int)even_sites_on_node : (int)sites_on_node;
}

/* source norm on one parity; the trivial solution for a zero source
 * (d_congrad5_fn_gpu.c:60-90) */
static double parity_norm(su3_vector *v, int parity) {
  int i, lo, hi;
  double s = 0.0;
  milc_hip_glue_parity_range(parity, &lo, &hi);
  for (i = lo; i < hi; i++) s += (double)magsq_su3vec(&v[i]);
  g_doublesum(&s);
  return s;
}

static void zero_parity(su3_vector *v, int parity) {
  int lo, hi;
  milc_hip_glue_parity_range(parity, &lo, &hi);
  if (hi > lo) memset(v + lo, 0, (size_t)(hi - lo) * sizeof(su3_vector));
}

static int glue_args(const quark_invert_control *qic, imp_ferm_links_t *fn, const char *myname,
                     milc_hip_inv_args *args) {
  if (qic->parity == EVEN) args->evenodd = MILC_HIP_EVEN_PARITY;
  else if (qic->parity == ODD) args->evenodd = MILC_HIP_ODD_PARITY;
  else {
    printf("%s: Unrecognised parity\n", myname);
    terminate(2);
  }
  args->max_iter = qic->max * qic->nrestart;
  args->mixed_precision = glue_mixed();
  args->naik_epsilon = fn->eps_naik;
  args->tadpole = glue_tadpole();
  /* MILC's restart interval: the FP64 solve recomputes the true residual
   * every qic->max iterations, as ks_congrad_parity_cpu does */
  args->restart_iters = qic->max;
  /* the links changed or another fn: the library re-uploads them
   * (d_congrad5_fn_gpu.c:120-126) */
  if (fn != get_fn_last() || fresh_fn_links(fn)) {
    cancel_quda_notification(fn);
    set_fn_last(fn);
    return -1;
  }
  return 0;
}

int ks_congrad_parity_gpu(su3_vector *t_src, su3_vector *t_dest, quark_invert_control *qic, Real mass,
                          imp_ferm_links_t *fn) {
  char myname[] = "ks_congrad_parity_gpu";
  double dtimec = -dclock();
#ifdef CGTIME
  double nflop = 1187;
#endif
  qic->size_r = 0;
  qic->size_relr = 0;
  qic->final_iters = 0;
  qic->final_restart = 0;
  qic->converged = 1;
  qic->final_rsq = 0.;
  qic->final_relrsq = 0.;

  if (parity_norm(t_src, qic->parity) == 0.0) {
    zero_parity(t_dest, qic->parity);
    return 0;
  }
  milc_hip_glue_init();
  milc_hip_inv_args args;
  int num_iters = glue_args(qic, fn, myname, &args);
  double residual = 0, relative_residual = 0;
  /* qic->relresid: MILC's Fermilab relative residual (d_congrad5_fn_milc.c:37-56,92);
   * the library stops on it as ks_congrad_parity_cpu does */
  int rc = milc_hip_invert(MILC_PRECISION, qic->prec, mass, args, qic->resid, qic->relresid, get_fatlinks(fn),
                           get_lnglinks(fn), t_src, t_dest, &residual, &relative_residual, &num_iters);
  if (rc < 0) milc_hip_glue_fail(myname, rc);

  qic->final_rsq = residual * residual;
  /* = my_relative_residue of the solution, as on the CPU path */
  qic->final_relrsq = relative_residual * relative_residual;
  qic->final_iters = num_iters;
  /* converged: every target in use was met (the CPU path's meaning; with
   * relresid = 0 this is the QUDA glue's residual < resid test) */
  qic->converged = (qic->resid <= 0 || residual < qic->resid) &&
                   (qic->relresid <= 0 || qic->final_relrsq < qic->relresid * qic->relresid);
  qic->size_r = 0.0;
  qic->size_relr = 0.0;

  dtimec += dclock();
#ifdef CGTIME
  if (this_node == 0) {
    printf("CONGRAD5: time = %e (fn_HIP %s) masses = 1 iters = %d mflops = %e\n", dtimec,
           prec_label[qic->prec - 1], qic->final_iters,
           (double)(nflop * volume * qic->final_iters / (1.0e6 * dtimec * numnodes())));
    fflush(stdout);
  }
#else
  (void)dtimec;
#endif
  return num_iters;
}

int ks_congrad_block_parity_gpu(int nsrc, su3_vector **t_src, su3_vector **t_dest, quark_invert_control *qic,
                                Real mass, imp_ferm_links_t *fn) {
  char myname[] = "ks_congrad_block_parity_gpu";
  int s;
  qic->size_r = 0;
  qic->size_relr = 0;
  qic->final_iters = 0;
  qic->final_restart = 0;
  qic->converged = 1;
  qic->final_rsq = 0.;
  qic->final_relrsq = 0.;
  if (nsrc <= 0) return 0;
  /* the batched solver has no relative-residual test: the sources go one
   * by one through the single-source solver then */
  if (qic->relresid != 0.) {
    int iters = 0, conv = 1;
    double rsq = 0, relrsq = 0;
    for (s = 0; s < nsrc; s++) {
      iters += ks_congrad_parity_gpu(t_src[s], t_dest[s], qic, mass, fn);
      conv = conv && qic->converged;
      if (qic->final_rsq > rsq) rsq = qic->final_rsq;
      if (qic->final_relrsq > relrsq) relrsq = qic->final_relrsq;
    }
    qic->final_iters = iters;
    qic->final_rsq = rsq;
    qic->final_relrsq = relrsq;
    qic->converged = conv;
    return iters;
  }
  if (parity_norm(t_src[0], qic->parity) == 0.0) {
    /* d_congrad5_fn_gpu.c:192-205 tests the first source only */
    for (s = 0; s < nsrc; s++) zero_parity(t_dest[s], qic->parity);
    return 0;
  }
  milc_hip_glue_init();
  milc_hip_inv_args args;
  int num_iters = glue_args(qic, fn, myname, &args);
  double *res = (double *)malloc(nsrc * sizeof(double));
  if (!res) {
    printf("%s: no room\n", myname);
    terminate(1);
  }
  int rc = milc_hip_invert_block(MILC_PRECISION, nsrc, mass, args, qic->resid, get_fatlinks(fn), get_lnglinks(fn),
                                 (void *const *)t_src, (void *const *)t_dest, res, &num_iters);
  if (rc < 0) milc_hip_glue_fail(myname, rc);
  double worst = 0;
  for (s = 0; s < nsrc; s++)
    if (res[s] > worst) worst = res[s];
  free(res);
  qic->final_rsq = worst * worst;
  qic->final_iters = num_iters;
  qic->converged = worst < qic->resid ? 1 : 0;
  return num_iters;
}
