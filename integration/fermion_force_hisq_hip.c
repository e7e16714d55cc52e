/* fermion_force_hisq_hip.c -- MILC's HISQ fermion force on the MI355X
 * library (libmilc_hip.so).
 *
 * Replaces the external entry points of generic_ks/fermion_force_hisq_multi.o
 * in a HISQ build (FERM_ACTION == HISQ, KS_MULTIFF = FNMAT): the GPU wrapper
 * of that file (fn_fermion_force_multi_hisq_wrapper_mx_gpu, :2168-2290) is
 * file-static and needs the QUDA headers, so the file is swapped as a whole.
 * eo_fermion_force_multi hands the same arguments the wrapper hands to
 * qudaHisqForce to milc_hip_hisq_force: per-term one-hop and three-hop
 * weights (2 r_t and naik * 2 r_t), the rows of the further Naik sets
 * (eps_naik (1/8, -1/24) 2 r_t), the two smearing levels' path coefficients,
 * and MILC's own W (U(3)-projected), V (fat7) and thin links of the HISQ
 * auxiliary links.  The momentum increments are added to the site momenta.
 */
#include "generic_ks_includes.h"
#include <string.h>

#include "milc_hip.h"
#include "milc_hip_glue.h"

#ifndef KS_MULTIFF
#define KS_MULTIFF FNMAT
#endif

/* {one_link, naik, three_staple, five_staple, seven_staple, lepage}, the
 * order qudaHisqForce takes (fermion_force_hisq_multi.c:2156-2166) */
static void path_coeffs(ks_component_paths paths, double a[6]) {
  a[0] = paths.act_path_coeff.one_link;
  a[1] = paths.act_path_coeff.naik;
  a[2] = paths.act_path_coeff.three_staple;
  a[3] = paths.act_path_coeff.five_staple;
  a[4] = paths.act_path_coeff.seven_staple;
  a[5] = paths.act_path_coeff.lepage;
}

void eo_fermion_force_multi(Real eps, Real *residues, su3_vector **xxx, int nterms, int prec, fermion_links_t *fl) {
  char myname[] = "eo_fermion_force_multi";
  double dtime = -dclock();
  int i, j, dir;
  site *s;
  if (prec != MILC_PRECISION)
    node0_printf("%s: WARNING, precision requests not supported. Using %d.\n", myname, MILC_PRECISION);
  if (nterms != n_order_naik_total) {
    node0_printf("%s: %d terms, but the Naik sets hold %d\n", myname, nterms, n_order_naik_total);
    terminate(1);
  }
  ks_action_paths_hisq *ap = get_action_paths_hisq(fl);
  hisq_auxiliary_t *aux = get_hisq_auxiliary(fl);
  int n_naiks = fermion_links_get_n_naiks(fl);
  double *eps_naik = fermion_links_get_eps_naik(fl);

  /* terms of the further Naik sets: their fields are the last ones */
  int n_naik_terms = 0;
  for (i = 1; i < n_naiks; i++) n_naik_terms += n_orders_naik[i];
  int nrows = n_order_naik_total + n_naik_terms;
  double *rows = (double *)malloc(2 * nrows * sizeof(double));
  double **coeff = (double **)malloc(nrows * sizeof(double *));
  Real *momentum = (Real *)malloc(sites_on_node * 4 * sizeof(anti_hermitmat));
  if (!rows || !coeff || !momentum) {
    printf("%s: no room\n", myname);
    terminate(1);
  }
  for (i = 0; i < nrows; i++) coeff[i] = rows + 2 * i;
  for (i = 0; i < n_order_naik_total; i++) {
    coeff[i][0] = 2.0 * residues[i];
    coeff[i][1] = ap->p2.act_path_coeff.naik * 2.0 * residues[i];
  }
  int row = n_order_naik_total, shift = n_orders_naik[0];
  for (i = 1; i < n_naiks; i++) {
    for (j = 0; j < n_orders_naik[i]; j++) {
      coeff[row][0] = ap->p3.act_path_coeff.one_link * eps_naik[i] * 2.0 * residues[shift + j];
      coeff[row][1] = ap->p3.act_path_coeff.naik * eps_naik[i] * 2.0 * residues[shift + j];
      row++;
    }
    shift += n_orders_naik[i];
  }
  double level2_coeff[6], fat7_coeff[6];
  path_coeffs(ap->p2, level2_coeff);
  path_coeffs(ap->p1, fat7_coeff);

  milc_hip_glue_init();
  memset(momentum, 0, sites_on_node * 4 * sizeof(anti_hermitmat));
  int rc = milc_hip_hisq_force(MILC_PRECISION, n_order_naik_total, n_naik_terms, eps, coeff, (void **)xxx,
                               level2_coeff, fat7_coeff, aux->W_unitlink, aux->V_link, aux->U_link, momentum);
  if (rc < 0) milc_hip_glue_fail(myname, rc);

  /* append the result (:2270-2277) */
  FORALLSITES(i, s) {
    for (dir = XUP; dir <= TUP; dir++) {
      Real *m = (Real *)&(s->mom[dir]);
      const Real *inc = momentum + (4 * i + dir) * 10;
      for (j = 0; j < 10; j++) m[j] += inc[j];
    }
  }
  free(momentum);
  free(coeff);
  free(rows);
  dtime += dclock();
#ifdef FFTIME
  node0_printf("FFTIME:  time = %e (HISQ HIP %s) terms = %d\n", dtime, MILC_PRECISION == 1 ? "F" : "D", nterms);
#else
  (void)dtime;
#endif
}

/* the one- and two-term forms of ks_imp_dyn (:223-275) */
void eo_fermion_force_oneterm(Real eps, Real weight, su3_vector *temp_x, int prec, fermion_links_t *fl) {
  su3_vector *xxx[1] = {temp_x};
  Real residues[1] = {weight};
  eo_fermion_force_multi(eps, residues, xxx, 1, prec, fl);
}

void eo_fermion_force_oneterm_site(Real eps, Real weight, field_offset x_off, int prec, fermion_links_t *fl) {
  su3_vector *temp_x = create_v_field_from_site_member(x_off);
  eo_fermion_force_oneterm(eps, weight, temp_x, prec, fl);
  destroy_v_field(temp_x);
}

void eo_fermion_force_twoterms(Real eps, Real weight1, Real weight2, su3_vector *x1_off, su3_vector *x2_off,
                               int prec, fermion_links_t *fl) {
  su3_vector *xxx[2] = {x1_off, x2_off};
  Real residues[2] = {weight1, weight2};
  eo_fermion_force_multi(eps, residues, xxx, 2, prec, fl);
}

void eo_fermion_force_twoterms_site(Real eps, Real weight1, Real weight2, field_offset x1_off,
                                    field_offset x2_off, int prec, fermion_links_t *fl) {
  su3_vector *x1 = create_v_field_from_site_member(x1_off);
  su3_vector *x2 = create_v_field_from_site_member(x2_off);
  eo_fermion_force_twoterms(eps, weight1, weight2, x1, x2, prec, fl);
  destroy_v_field(x1);
  destroy_v_field(x2);
}

const char *ks_multiff_opt_chr(void) { return "FNMAT"; }
