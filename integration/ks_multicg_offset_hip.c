/* ks_multicg_offset_hip.c -- MILC's multi-shift staggered CG on the MI355X
 * library (libmilc_hip.so): replaces generic_ks/ks_multicg_offset_gpu.c.
 *
 * Built with -DUSE_CG_GPU, so that ks_multicg_offset_field names
 * ks_multicg_offset_field_gpu (include/imp_ferm_links.h:238-239) and
 * ks_multicg_field / ks_multicg_hybrid_field (generic_ks/ks_multicg.c:65-102,
 * 170-236) reach it unchanged.  Also defines get_fn_last / set_fn_last, the
 * record of the fn links the library holds (ks_multicg_offset_gpu.c:25-36).
 */
#include "generic_ks_includes.h"
#include <math.h>
#include <string.h>

#include "milc_hip.h"
#include "milc_hip_glue.h"

#ifdef SINGLE_FOR_DOUBLE
#define HALF_MIXED
#endif

static imp_ferm_links_t *fn_last = NULL;

imp_ferm_links_t *get_fn_last(void) { return fn_last; }

void set_fn_last(imp_ferm_links_t *fn_last_new) { fn_last = fn_last_new; }

int ks_multicg_offset_field_gpu(su3_vector *src, su3_vector **psim, ks_param *ksp, int num_offsets,
                                quark_invert_control *qic, imp_ferm_links_t *fn) {
  char myname[] = "ks_multicg_offset_field_gpu";
  int i, j;
#ifdef CGTIME
  double dtimec = -dclock();
  double nflop = 1205 + 15 * num_offsets;
#endif
  /* no Fermilab relative residual in the multi-shift solve (:55-58) */
  if (qic[0].relresid != 0.) {
    printf("%s: GPU code does not yet support a Fermilab-type relative residual\n", myname);
    terminate(1);
  }
  for (j = 0; j < num_offsets; j++) {
    qic[j].final_rsq = 0.;
    qic[j].final_relrsq = 0.;
    qic[j].size_r = 0.;
    qic[j].size_relr = 0.;
    qic[j].final_iters = 0;
    qic[j].final_restart = 0;
    qic[j].converged = 1;
  }
  if (num_offsets == 0) return 0;

  int lo, hi;
  milc_hip_glue_parity_range(qic[0].parity, &lo, &hi);
  double source_norm = 0.0;
  for (i = lo; i < hi; i++) source_norm += (double)magsq_su3vec(&src[i]);
  g_doublesum(&source_norm);
  if (source_norm == 0.0) {
    for (j = 0; j < num_offsets; j++)
      if (hi > lo) memset(psim[j] + lo, 0, (size_t)(hi - lo) * sizeof(su3_vector));
    return 0;
  }

  milc_hip_inv_args args;
  memset(&args, 0, sizeof args);
  if (qic[0].parity == EVEN) args.evenodd = MILC_HIP_EVEN_PARITY;
  else if (qic[0].parity == ODD) args.evenodd = MILC_HIP_ODD_PARITY;
  else {
    node0_printf("%s: EVENANDODD not supported\n", myname);
    terminate(1);
  }

  double *offset = (double *)malloc(num_offsets * sizeof(double));
  double *residual = (double *)malloc(num_offsets * sizeof(double));
  double *relative_residual = (double *)malloc(num_offsets * sizeof(double));
  double *final_residual = (double *)malloc(num_offsets * sizeof(double));
  double *final_relative_residual = (double *)malloc(num_offsets * sizeof(double));
  if (!offset || !residual || !relative_residual || !final_residual || !final_relative_residual) {
    printf("%s: no room\n", myname);
    terminate(1);
  }
  /* per-shift targets as the QUDA glue sets them (:133-158): the mixed
   * solver scales the higher shifts' targets by the residues; the FP64
   * solver runs every shift with the base system */
  for (i = 0; i < num_offsets; i++) {
    offset[i] = ksp[i].offset;
    residual[i] = qic[i].resid;
    if (i > 0) {
#if defined(MAX_MIXED) || defined(HALF_MIXED)
      if (ksp[i].residue != 0) {
        residual[i] = fabs(ksp[0].residue / ksp[i].residue) * residual[0];
        if (residual[i] < 1e-14) residual[i] = 1e-14;
      }
#else
      residual[i] = 0;
#endif
    }
    relative_residual[i] = qic[i].relresid;
  }
  args.max_iter = qic[0].max * qic[0].nrestart;
#if defined(MAX_MIXED) || defined(HALF_MIXED)
  args.mixed_precision = 1; /* never 16-bit in the multi-shift stage (:160-164) */
#else
  args.mixed_precision = 0;
#endif
  milc_hip_glue_init();

  int num_iters = 0;
  if (fn != get_fn_last() || fresh_fn_links(fn)) {
    cancel_quda_notification(fn);
    set_fn_last(fn);
    num_iters = -1;
  }
  /* a change of Naik term also refreshes the links (:179-184) */
  static int naik_term_epsilon_index = -1;
  if (naik_term_epsilon_index != ksp[0].naik_term_epsilon_index) {
    num_iters = -1;
    naik_term_epsilon_index = ksp[0].naik_term_epsilon_index;
  }
  args.naik_epsilon = ksp[0].naik_term_epsilon;
  if (args.naik_epsilon != fn->eps_naik) {
    node0_printf("%s: naik_epsilon in action (%e) does not match value in link (%e)\n", myname, args.naik_epsilon,
                 fn->eps_naik);
    terminate(1);
  }
#if (FERM_ACTION == HISQ)
  args.tadpole = 1.0;
#else
  args.tadpole = u0;
#endif

  int rc = milc_hip_multishift_invert(MILC_PRECISION, qic[0].prec, num_offsets, offset, args, residual,
                                      relative_residual, get_fatlinks(fn), get_lnglinks(fn), (void *)src,
                                      (void **)psim, final_residual, final_relative_residual, &num_iters);
  if (rc < 0) milc_hip_glue_fail(myname, rc);

  for (i = 0; i < num_offsets; i++) {
    qic[i].final_rsq = final_residual[i] * final_residual[i];
    qic[i].final_relrsq = final_relative_residual[i] * final_relative_residual[i];
    qic[i].final_iters = num_iters;
    qic[i].size_r = 0.0;
    qic[i].size_relr = 0.0;
  }
  free(offset);
  free(residual);
  free(relative_residual);
  free(final_residual);
  free(final_relative_residual);
#ifdef CGTIME
  dtimec += dclock();
  if (this_node == 0) {
    printf("CONGRAD5: time = %e (multicg_offset_HIP %s) masses = %d iters = %d mflops = %e\n", dtimec,
           qic[0].prec == 1 ? "F" : "D", num_offsets, num_iters,
           (double)(nflop)*volume * num_iters / (1.0e6 * dtimec * numnodes()));
    fflush(stdout);
  }
#endif
  return num_iters;
}
