/* milc_hip.h -- C-ABI of the MI355X-native HISQ/asqtad staggered library.
 *
 * The entry points mirror the GPU-offload calls that MILC 7.8.1's generic_ks
 * "_gpu" glue makes (the QUDA interface used at the call sites cited below),
 * with plain pointers and sizes only.  MILC host fields are passed exactly as
 * MILC holds them:
 *   - colour vectors: su3_vector[sites_on_node], even sites first
 *     (node_index order, generic/layout_hyper_prime.c:509-520), complex
 *     re/im interleaved, 6 Reals per site;
 *   - links: su3_matrix[4*sites_on_node], element (4*i + dir), 18 Reals, KS
 *     phases and the antiperiodic time boundary already multiplied in
 *     (generic_ks/rephase.c:122-143);
 *   - "Real" is float when external_precision == 1, double when 2
 *     (MILC_PRECISION).
 * Operator conventions are MILC's: D has no factor 1/2; M = 2m + D;
 * the single-parity solver inverts M^dag M = 4m^2 - D_eo D_oe; multi-shift
 * offsets already include 4m^2 (ks_multicg_offset.c:184-194).
 *
 * Return value: >= 0 on success (iteration count where the mirrored call
 * returns one), negative on error.  The library keeps one device context per
 * process (the reference's GPU path is single-threaded per rank and not
 * reentrant, d_congrad5_fn_gpu.c:120-126); calls must not overlap.
 */
#ifndef MILC_HIP_H
#define MILC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* parity codes of inv_args.evenodd */
#define MILC_HIP_EVEN_PARITY 0
#define MILC_HIP_ODD_PARITY 1

/* Mirrors the QudaInvertArgs_t members MILC sets before qudaInvert /
 * qudaMultishiftInvert / qudaDslash (d_congrad5_fn_gpu.c:95-131,
 * ks_multicg_offset_gpu.c:161-199, dslash_fn.c:313-325).
 * mixed_precision: 0 = solve entirely in internal precision; 1 = FP32 sloppy
 * iterations with FP64 reliable updates; 2 = 16-bit (per-site normed) sloppy
 * iterations with FP64 reliable updates (MILC's HALF_MIXED / MAX_MIXED,
 * Makefile:463-466). */
typedef struct {
  int max_iter;
  int evenodd;
  int mixed_precision;
  double naik_epsilon;
  double tadpole;
  /* Addition without a QUDA counterpart: MILC's per-restart iteration count
   * (qic->max, d_congrad5_fn_milc.c:87); the FP64 solver recomputes the true
   * residual every restart_iters iterations exactly as the CPU path does.
   * 0 = max_iter. */
  int restart_iters;
} milc_hip_inv_args;

/* ---- lifecycle ----------------------------------------------------------- */
/* Replaces initialize_quda()/qudaInit + layout set-up
 * (generic/milc_to_quda_utilities.c:11-43): one lattice nx*ny*nz*nt on one
 * device.  nx and the volume must be even. */
int milc_hip_init(const int dims[4], int device);
/* Replaces qudaFinalize (ks_imp_rhmc/control.c:189). */
int milc_hip_finalize(void);
/* Library version string. */
const char* milc_hip_version(void);

/* ---- solvers --------------------------------------------------------------- */
/* Replaces qudaInvert (generic_ks/d_congrad5_fn_gpu.c:136-148): CG on
 * M^dag M for parity args.evenodd.  `solution` is the initial guess and is
 * overwritten on that parity only.  On entry *num_iters == -1 signals that the
 * links changed (d_congrad5_fn_gpu.c:120-126); otherwise the device copy of
 * the links from the previous call is reused when the pointers match.
 * internal_precision 2 with mixed_precision 0 reproduces MILC's FP64
 * restarted CG (d_congrad5_fn_milc.c:60-407).  final_resid = |r|/|b| of the
 * FP64 true residual.  target_relresid > 0 adds MILC's Fermilab relative
 * residual test (qic->relresid, d_congrad5_fn_milc.c:37-56,92): the solve also
 * runs until my_relative_residue(r, x) = sqrt(sum_i |r_i|^2/|x_i|^2 / (V/2)) <
 * target_relresid^2, tested where the CPU path tests it (after every update and
 * at every restart); 0 = no such test.  final_rel_resid =
 * sqrt(my_relative_residue) of the solution, so that the glue's square
 * (d_congrad5_fn_gpu.c:151) gives the CPU path's final_relrsq. */
int milc_hip_invert(int external_precision, int internal_precision, double mass, milc_hip_inv_args args,
                    double target_resid, double target_relresid, const void* fatlink, const void* longlink,
                    const void* source, void* solution, double* final_resid, double* final_rel_resid,
                    int* num_iters);

/* Replaces qudaMultishiftInvert (generic_ks/ks_multicg_offset_gpu.c:203-217):
 * Jegerlehner multi-mass CG for (M^dag M + offset_j - 4m^2) x_j = b, i.e.
 * (-D_eo D_oe + offset_j) x_j = b, all offsets > 0.  Solutions are zeroed
 * first.  The base system (smallest offset) stops at target_resid[0];
 * internal_precision 1 runs the vector updates in FP32 with FP64
 * recurrences, 2 runs everything in FP64.  With args.mixed_precision set,
 * the FP32 bulk stops at max(target_resid[0], 1e-8) and every shift is then
 * refined to its own FP64 true residual target_resid[j] (the per-shift CG of
 * ks_multicg_hybrid_field), up to four consecutive shifts in one batched mixed
 * CG with an FP32 (mixed_precision 1) or 16-bit (mixed_precision 2) sloppy
 * operator; final_resid[j] is then that true |r|/|b|.
 * A nonzero target_relresid[j] returns -4: the multi-shift solve has no
 * relative-residual test (the reference glue stops on it,
 * ks_multicg_offset_gpu.c:55-58). */
int milc_hip_multishift_invert(int external_precision, int internal_precision, int num_offsets,
                               const double* offsets, milc_hip_inv_args args, const double* target_resid,
                               const double* target_relresid, const void* fatlink, const void* longlink,
                               const void* source, void** solutions, double* final_resid,
                               double* final_rel_resid, int* num_iters);

/* Replaces qudaDslash (generic_ks/dslash_fn.c:336-341): dst = D src on
 * parity args.evenodd (dslash_fn_dblstore.c:311-563 semantics). */
int milc_hip_dslash(int external_precision, int internal_precision, milc_hip_inv_args args,
                    const void* fatlink, const void* longlink, const void* src, void* dst, int* num_iters);

/* ---- fermion force --------------------------------------------------------- */
/* Replaces qudaComputeOprod (generic_ks/fermion_force_asqtad_gpu.c:60) and
 * restates the level-0 outer products of the HISQ force
 * (fermion_force_hisq_multi.c:1478-1636):
 *   oprod[0][dir][i] = sum_t coeff[t][0] * Z_t(x_i + dir)  Z_t(x_i)^dag
 *   oprod[1][dir][i] = sum_t coeff[t][1] * Z_t(x_i + 3dir) Z_t(x_i)^dag
 * quark_field[t] = Z_t on all sites (X_t on even, D_oe X_t on odd,
 * update_h_rhmc.c:82-84); oprod[k] points to an array of 4 su3_matrix[V]
 * fields, overwritten. */
int milc_hip_compute_oprod(int precision, int num_terms, double** coeff, void** quark_field, void* oprod[2]);

/* ---- FN (fat + Naik) fermion force on the thin links ------------------------ */
/* Replaces qudaAsqtadForce (generic_ks/fermion_force_asqtad_gpu.c:133-136):
 * the asqtad path chain from the outer products to the momentum.
 * act_path_coeff = {one_link, naik, three_staple, five_staple, seven_staple,
 * lepage} (ks_action_paths.act_path_coeff); one_link_oprod[dir] /
 * naik_oprod[dir] = su3_matrix[V] outer products that already carry 2 eps
 * and the term weights (fermion_force_asqtad_gpu.c:53-60); gauge = thin links
 * [i][dir][18] with the KS phases in.  momentum = anti_hermitmat [i][dir]
 * packed as 10 Reals (m01, m02, m12 complex; m00im, m11im, m22im; space);
 * the force is ADDED to it.  precision 1 = float, 2 = double for every array. */
int milc_hip_asqtad_force(int precision, const double act_path_coeff[6], const void* const one_link_oprod[4],
                          const void* const naik_oprod[4], const void* gauge, void* momentum);
/* Replaces fermion_force_fn_multi (generic_ks/fermion_force_fn_multi.c:81-313,
 * dispatcher fermion_force_multi.c:29-45): multi_x[t] = X_t on all sites
 * (Real su3_vector[V]), residues[t] (double), eps; outer products, path chain
 * and mom += TA(2 eps U G) in one call, arrays as in milc_hip_asqtad_force. */
int milc_hip_fermion_force_fn_multi(int precision, double eps, const double* residues, void** multi_x, int nterms,
                                    const double act_path_coeff[6], const void* gauge, void* momentum);

/* HISQ thin-link force.  Replaces qudaHisqForce (generic_ks/fermion_force_hisq_multi.c:2266-2268,
 * fn_fermion_force_multi_hisq_wrapper_mx_gpu): coeff[t] = {one-hop, three-hop}
 * weights of term t (MILC passes 2 r_t and naik * 2 r_t), quark_field[t] = X_t on
 * all sites; level2_coeff / fat7_coeff = the six path coefficients of the two
 * smearing levels (write_path_coeffs_to_array order); w_link, v_link, u_link =
 * HISQ W (U(3)-projected), V (fat7) and thin links, [i][dir][18], phases in.
 * Level 0 outer products -> level-2 adjoint on W (the three-hop product enters
 * the Naik path with unit weight) -> derivative of W = V (V^dag V)^-1/2 ->
 * fat7 adjoint on U; TA(eps U F) is ADDED to momentum ([i][dir][10]).
 * num_naik_terms (0..num_terms): the last num_naik_terms fields belong to a
 * second Naik set (a charm quark, links X_2 + eps_naik X_3); coeff then has
 * num_terms + num_naik_terms rows, row num_terms + i = eps_naik (1/8, -1/24)
 * 2 r of field num_terms - num_naik_terms + i (fermion_force_hisq_multi.c:
 * 2195-2222), and those products take the "1-link + Naik" paths on W. */
int milc_hip_hisq_force(int precision, int num_terms, int num_naik_terms, double eps, double** coeff,
                        void** quark_field, const double level2_coeff[6], const double fat7_coeff[6],
                        const void* w_link, const void* v_link, const void* u_link, void* momentum);
/* The same force in the form of eo_fermion_force_multi for the HISQ action
 * (fermion_force_hisq_multi.c:164-209 -> ..._wrapper_mx_cpu :1182-1475):
 * residues[t], multi_x[t] on all sites, thin links with phases in; V and W are
 * rebuilt on the device; TA(2 eps U F) is ADDED to momentum. */
int milc_hip_eo_fermion_force_multi(int precision, double eps, const double* residues, void** multi_x, int nterms,
                                    const void* thin_links, void* momentum);

/* dst slot (FP64) = D src slot on output parity parity_out, the stencil run in
 * precision prec (0 FP64, 1 FP32, 2 16-bit: the source converted to and the
 * result converted back from that precision) on the current device links. */
int milc_hip_dev_dslash(int prec, int slot_src, int parity_out, int slot_dst);

/* ---- gauge action and molecular-dynamics link update ---------------------
 * Links: su3_matrix[4 V] in MILC layout and precision (1 float, 2 double),
 * phases_in = the KS phases are in (they are removed for the gauge loops).
 * Momenta: anti_hermitmat[4 V] (m01, m02, m12, m00im, m11im, m22im, space).
 * loop_coeff: the three loop coefficients of the Symanzik 1x1 + 1x2 + 1x1x1
 * action (num_loop_types must be 3; generic/imp_actions/symanzik_1loop_*). */

/* momentum += -eb3 TA(U G), G = sum over the action's loops through each link
 * of the loop coefficient times the rest of the loop.  Replaces
 * qudaGaugeForcePhased (generic/gauge_force_imp_gpu.c:31-34, eb3 = eps beta/3)
 * and imp_gauge_force_cpu (generic/gauge_force_imp.c:37-219). */
int milc_hip_gauge_force(int precision, int num_loop_types, const double* loop_coeff, double eb3, const void* links,
                         int phases_in, void* momentum);
/* *action = sum over loops and sites of coeff (3 - Re Tr P): imp_gauge_action
 * (generic/gauge_stuff.c:248-322); MILC's gauge action is beta/3 times it
 * (ks_imp_rhmc/d_action_rhmc.c:27). */
int milc_hip_gauge_action(int precision, int num_loop_types, const double* loop_coeff, const void* links,
                          int phases_in, double* action);
/* links <- exp(eps P) links to eighth order (Horner), in place: replaces
 * qudaUpdateUPhased / update_u (ks_imp_rhmc/update_u.c:14-34, :66-111). */
int milc_hip_update_u(int precision, double eps, void* links, const void* momentum, int phases_in);
/* reunitarize_ks (generic_ks/reunitarize_ks.c:9-29): every link made SU(3) by
 * reunit_su3 (generic/reunitarize2.c:133-205; rows 0 and 1 by Gram-Schmidt,
 * row 2 = conj(row0 x row1)), KS phases taken out and put back when
 * phases_in; replaces qudaUnitarizeSU3Phased (reunitarize2.c:210-227) and
 * reunitarize_cpu (:231).  *errors = the number of deviations at or above
 * tolerance (MILC's TOLERANCE is 1e-4; it stops above 100), *max_deviation the
 * largest. */
int milc_hip_reunitarize(int precision, double tolerance, void* links, int phases_in, double* max_deviation,
                         int* errors);
/* *action = sum over links of |P|^2/2 - 4: hmom_action (ks_imp_rhmc/d_action_rhmc.c:86-101). */
int milc_hip_mom_action(int precision, const void* momentum, double* action);

/* ---- one RHMC trajectory on the device -------------------------------------
 * Rational functions in MILC's layout (params_ratfunc, ks_imp_rhmc/params_rhmc.h):
 * R(x) = res[0] + sum_{k=1..order} res[k] / (x + pole[k]), x = M^dag M. */
typedef struct {
  int order;
  double res[33];
  double pole[33];
} milc_hip_rational;

#define MILC_HIP_MAX_PSEUDO 8
/* one pseudofermion (rparam[i] and the cgresid_md_fa_gr line of MILC's input) */
typedef struct {
  double mass;                          /* 4 mass^2 is added to every pole; 0 when the
                                           poles already hold it, as in MILC's files */
  milc_hip_rational md, gr, fa;         /* force, heatbath, action */
  double resid_md, resid_gr, resid_fa;  /* solver targets |r|/|b| (MILC: sqrt(rsqmin)) */
  double naik_epsilon;                  /* naik_term_epsilon of the rational file (HISQ only):
                                           != 0 puts the pseudofermion on the links
                                           X_2 + eps X_3 of its Naik set */
} milc_hip_pseudofermion;

typedef struct {
  int steps;             /* steps_per_trajectory (fermion-force steps); even except for
                            the leapfrog */
  int integrator;        /* MILC's int_alg_t (ks_imp_rhmc/ks_imp_includes.h:58-60):
                            0 INT_LEAPFROG, 1 INT_OMELYAN (lambda = 0.8), 4 INT_2G1F,
                            5 INT_3G1F (the HISQ build's), 6 INT_5G1F, 7 INT_6G1F
                            (alpha = beta = 0.1) */
  double epsilon;        /* microcanonical_time_step */
  double beta;           /* gauge coupling */
  double loop_coeff[3];  /* Symanzik loop coefficients (1x1, 1x2, 1x1x1) */
  int action;            /* quark action: 0 HISQ (ks_imp_rhmc's su3_rhmc_hisq builds),
                            1 asqtad with tadpole factor u0 (su3_rhmc / su3_omelyan_rhmc) */
  double u0;             /* asqtad tadpole factor (unused for HISQ) */
  int n_pseudo;          /* 1..MILC_HIP_MAX_PSEUDO */
  milc_hip_pseudofermion pf[MILC_HIP_MAX_PSEUDO];
  int mixed_md;          /* MD solves: FP32 multi-shift + 16-bit mixed refinement */
  int mixed_action;      /* heatbath and action solves the same way (each shift still
                            refined to its FP64 true residual) */
  int action_from_md;    /* 1: fermion action phi^dag R_MD phi (the MD force's own
                            action, for integrator tests); 0: |R_FA phi|^2 */
  int accept;            /* 0 keep the new links, 1 Metropolis, 2 restore the old */
  int site_rng;          /* 1: MILC's per-site generators (milc_hip_dev_seed_rng), which
                            continue from call to call, and MILC's node generator for
                            the Metropolis draw; 0: this library's streams from seed */
  unsigned long long seed;  /* site_rng = 0: momenta, pseudofermion noise, accept draw */
} milc_hip_rhmc_params;

/* One trajectory of update() (ks_imp_rhmc/update_rhmc.c:157-745) on the
 * current device links (milc_hip_dev_gauge_hisq or milc_hip_dev_load_thin_links):
 * Gaussian momenta (ranmom), phi_i = R_GR,i chi_i for each pseudofermion
 * (grsource_imp_rhmc), the actions (d_action_rhmc: beta/3 gauge loops, momenta,
 * sum over pseudofermions of the fermion action), the integrator's link
 * updates (update_u) and kicks (update_h_gauge: gauge force; update_h_fermion:
 * every pseudofermion's MD solve, then one HISQ force from all terms), with the
 * fat and Naik links rebuilt when a fermion kick or action needs them and the
 * links reunitarised (reunitarize_ks, tolerance 1e-4) after each leapfrog step
 * or step pair, the actions again, accept/reject.
 * stats: 32 doubles = g0, h0, f0, g1, h1, f1, dH, accepted, seconds, MD solver
 * iterations, fermion kicks, seconds in fermion kicks, in gauge kicks, in link
 * updates/rebuilds/reunitarisation, in actions, heatbath solver iterations,
 * gauge kicks, largest reunitarisation deviation, reunitarisation errors, link
 * rebuilds, [20..27] |chi_i|^2 of the heatbath noise (MILC's "GRSOURCE: sum";
 * with site_rng and single-precision generators summed as MILC's
 * single-precision build does), [28] action solver iterations, seconds in
 * [29] set-up, [30] heatbath, [31] MD solves (part of the fermion kicks).
 * Returns -6 when a reunitarisation finds more than 100 deviations (MILC
 * terminates there); after any error the device links are the ones the call
 * started with. */
int milc_hip_dev_rhmc_trajectory(const milc_hip_rhmc_params* p, double* stats);
/* Test hook for the error contract above, no counterpart in the reference:
 * the next trajectory fails with -90 at `stage` (1 after the heatbath, 2 after
 * the starting actions, 3 after the molecular dynamics; 0 = none), once. */
int milc_hip_debug_fail_trajectory(int stage);

/* MILC's per-site random number generators on the device: seeded as
 * initialize_site_prn_from_seed(iseed) (generic/ranstuff.c:117-131, index =
 * lexicographic site index) and the node generator as initialize_prn(&node_prn,
 * iseed, volume) (ks_imp_rhmc/setup.c:140).  precision 1 rounds the Gaussian
 * draws as a MILC_PRECISION=1 build, 2 as a double-precision build. */
int milc_hip_dev_seed_rng(int iseed, int precision);
/* ranmom() (generic/ranmom.c:10-41) from the site generators: momentum =
 * anti_hermitmat[4 V] as 10 doubles each, node_index order. */
int milc_hip_dev_ranmom(double* momentum);
/* three complex Gaussians per site of `parity` from the site generators
 * (grsource_imp_rhmc, ks_imp_rhmc/grsource_rhmc.c:30-37) into device slot
 * `slot`; *sum = their |.|^2 summed over the sites ("GRSOURCE: sum"). */
int milc_hip_dev_gaussian_source(int parity, int slot, double* sum);

/* 16-bit sloppy long links in the two-row form (on = 1, the default): a long
 * link that is a real multiple of a U(3) matrix, L = c U (HISQ and asqtad
 * Naik links), is stored as rows 0 and 1 plus the phase of det(U) in 16
 * bits, and row 2 = z conj(row0 x row1), z = exp(i phase) / c with c from the
 * row norms, is rebuilt in the stencil: 26 instead of 36 bytes per link, 528
 * instead of 608 bytes per output site of the 16-bit Dslash (the 16-bit
 * batched stencil of the batched CG uses the same links).  The form is used
 * only when every long link passes |L^dag L - c^2 1| <= 1e-5 c^2;
 * other links stay in the nine-element form.  No counterpart in the
 * reference (QUDA's reconstruct-13 plays this role under MILC's QUDA glue). */
int milc_hip_set_link_compression(int on);
/* Long-link format of the precision-prec device links (building them if
 * needed): 1 two-row form, 0 nine elements, negative on error. */
int milc_hip_link_compression(int prec);

/* Evaluation form of the thin-link force adjoints (FN and HISQ): 0 = gather
 * over the path table for every smearing; 1 (default) = reverse mode through
 * the generalised staples of the fattening wherever that is exact for any
 * links (no Lepage term, or the U(3)-projected HISQ links), path table
 * otherwise; 2 = staple form for every smearing (exact for unitary links: the
 * optimised Lepage term retraces U U^+).  Returns -2 for another value.  No
 * counterpart in the reference (MILC's CPU code walks sorted paths). */
int milc_hip_set_force_form(int form);

/* ---- many right-hand sides ------------------------------------------------- */
/* Block CG: replaces ks_congrad_block_parity_gpu (generic_ks/d_congrad5_fn_gpu.c:174-176):
 * nsrc sources on parity args.evenodd, solution[s] holds the initial guess on
 * entry (d_congrad5_fn_gpu.c:104-110) and the solution on exit (only that
 * parity's sites are written).  Batched FP32 iterations with FP64 reliable
 * updates to the FP64 true residual |b - M^dag M x| / |b| < target_residual;
 * final_residual[s] (may be NULL) gets each source's true residual.  Returns
 * the total iteration count. */
int milc_hip_invert_block(int external_precision, int nsrc, double mass, milc_hip_inv_args args,
                          double target_residual, const void* fatlink, const void* longlink, void* const* source,
                          void* const* solution, double* final_residual, int* num_iters);
/* Multi-shift over many sources: the batched form of ks_multicg_offset_field_gpu
 * (generic_ks/ks_multicg_offset_gpu.c:38-45), one call for the sources that
 * MILC solves one after another (the pseudofermion loop of
 * ks_imp_rhmc/update_h_rhmc.c:64-84, source loops of ks_spectrum).
 * solution[s * num_offsets + j] = (M^dag M + offsets[j] - 4m^2)^-1 source[s]
 * (offsets include 4m^2, ks_multicg_offset.c:184-194).  The multi-shift stage
 * runs to target_residual[0] in FP32 with FP64 recurrences; refine_residual > 0
 * then refines every shift to that FP64 true residual (mixed-precision CG,
 * the per-shift refinement of ks_multicg_hybrid_field, ks_multicg.c:170-236).
 * final_residual[s * num_offsets + j] gets the true residuals (after refinement). */
int milc_hip_multishift_invert_block(int external_precision, int nsrc, int num_offsets, const double* offsets,
                                     milc_hip_inv_args args, const double* target_residual, double refine_residual,
                                     const void* fatlink, const void* longlink, void* const* source,
                                     void* const* solution, double* final_residual, int* num_iters);

/* ---- link construction ----------------------------------------------------- */
/* Replaces qudaLoadUnitarizedLink + qudaLoadKSLink for the HISQ U(3) action
 * (generic_ks/fermion_links_hisq_load_milc.c:546-702,
 * fermion_links_fn_load_gpu.c:18-123): thin links U (phases in) ->
 * fat7 V -> analytic U(3) projection W -> fat and Naik links.  n_svd, if not
 * NULL, receives the number of links at which MILC would have used its SVD
 * fallback (the analytic result is kept for them). */
int milc_hip_load_hisq_links(int precision, const void* thin, void* fatlink, void* longlink, void* W_link,
                             int* n_svd);
/* The links of a Naik set with correction eps_naik (a charm quark of a 2+1+1
 * run): X_2 + eps_naik X_3, X_3 = the "1-link + Naik" table on W (1/8, -1/24),
 * as load_hisq_fn_links builds set inaik (fermion_links_hisq_load_milc.c:
 * 574-640, the qudaLoadKSLink call of that table on the GPU path);
 * eps_naik = 0 gives milc_hip_load_hisq_links. */
int milc_hip_load_hisq_links_naik(int precision, const void* thin, double eps_naik, void* fatlink, void* longlink,
                                  void* W_link, int* n_svd);

/* ---- device-resident workloads (inputs generated in HBM) -------------------- */
/* Thin links U = exp(i*scale*H) from `seed` (H traceless Hermitian, Gaussian
 * entries; counter-based generator shared with the oracle), KS phases in,
 * then the HISQ fat/Naik links, kept on the device as the current links. */
int milc_hip_dev_gauge_hisq(uint64_t seed, double scale, int* n_svd);
/* Thin links from the host (MILC layout, precision 1 float / 2 double), e.g.
 * as read by milc_io_read_gauge: the KS phases are applied on the device
 * unless phases_in (rephase.c:64-143), the HISQ links are built on the device
 * (load_hisq_aux_links_cpu + load_hisq_fn_links,
 * fermion_links_hisq_load_milc.c:546-702), and both become the current device
 * links -- the device side of reload_lattice followed by load_ferm_links. */
int milc_hip_dev_load_thin_links(int precision, const void* thin, int phases_in, int* n_svd);
/* The current device thin links, MILC layout, FP64, with the KS phases
 * (with_phases = 1) or without (0, as save_lattice writes them). */
int milc_hip_dev_read_thin_links(double* thin, int with_phases);
/* Gaussian even/odd source (<re^2> = 1/2, grsource_imp.c:169-195) into device
 * slot `slot` (0..63). */
int milc_hip_dev_source(uint64_t seed, int parity, int slot);
/* Solve on device slots: x[slot_x] = (M^dag M)^-1 b[slot_b]; returns
 * iterations; stats[0..5] = final |r|/|b|, seconds, restarts/reliable
 * updates, converged, iterations, final_rsq. */
int milc_hip_dev_invert(int slot_b, int slot_x, double mass, int parity, double resid, int mixed_precision,
                        int max_iter, double* stats);
/* Allocation-free form of milc_hip_dev_invert; no counterpart in the
 * reference.  For callers that cannot allocate device memory (tight solve
 * loops, harnesses that forbid hipMalloc): the work vectors come from one
 * static device array of 256 MiB per device (held until the process ends),
 * and the call makes no hipMalloc, hipFree or other device allocation.
 * Preconditions (else -3): slot_x already exists (written by
 * milc_hip_dev_write, milc_hip_dev_source or an earlier solve) and the links of
 * the sloppy precision are resident (one milc_hip_dev_invert with the same
 * mixed_precision after the links were loaded builds them).  Work vectors
 * needing more than 256 MiB (FP64: 4 x 48 B, mixed: 200 B per site of one
 * parity, i.e. up to about 2.6M lattice sites mixed, 2.8M FP64) return -6
 * with a message.  parity must be EVEN or ODD (else -2).  Same results and
 * stats as milc_hip_dev_invert.  Calls share the array, so calls that overlap
 * on one device race (the library's calls must not overlap anyway). */
int milc_hip_dev_invert_noalloc(int slot_b, int slot_x, double mass, int parity, double resid, int mixed_precision,
                                int max_iter, double* stats);
/* RHMC fermion action (ks_ratinv + ks_rateval + FP64 sum, d_action_rhmc.c
 * fermion_action): action = sum_even |residues[0] phi + sum_k residues[k] x_k|^2,
 * x_k = (M^dag M + roots[k] - 4m^2)^-1 phi (roots already include 4m^2, k = 1..order).
 * mixed_precision 0 = FP64 multi-shift, 1 = FP32 multi-shift + mixed per-shift
 * refinement.  Returns the total iterations. */
int milc_hip_dev_rhmc_action(int slot_phi, int order, const double* roots, const double* residues, double resid,
                             int mixed_precision, double* action);
/* Level-0 fermion force of S = sum_t residues[t] phi^dag (M^dag M + offsets[t] - 4m^2)^-1 phi
 * on the current device links (update_h_rhmc.c:77-107 up to the force call):
 * X_t by multi-shift, Z_t = (X_t even, D_oe X_t odd), then oprod1/oprod3 as in
 * milc_hip_compute_oprod with coeff[t] = (residues[t], residues[t]).  *S gets
 * the action value; oprod1/oprod3 (MILC layout [i][dir][18], may be NULL) the
 * outer products.  dS/dF_mu(x) = 2 eps(x) O1_mu(x)^dag (eps = +1 even, -1 odd). */
int milc_hip_dev_force_level0(int slot_phi, int n, const double* offsets, const double* residues, double resid,
                              int mixed_precision, double* S, double* oprod1, double* oprod3);
/* Make MILC-layout host fat/long links (double) the current device links. */
int milc_hip_dev_load_links(const double* fat, const double* lng);
/* Copy a device slot to / from a MILC-layout host field (double). */
int milc_hip_dev_read(int slot, int parity, double* host);
int milc_hip_dev_write(int slot, int parity, const double* host);
/* Current device links to host MILC arrays (double). */
int milc_hip_dev_read_links(double* fat, double* lng);
/* Time `reps` Dslash applications on device slots (HIP events on the
 * library stream); returns the average in *usec. */
int milc_hip_dev_dslash_bench(int prec, int parity, int reps, double* usec);
/* The kernels of the 16-bit mixed CG iteration timed inside their own
 * sequence: from the residual of source slot slot_src (x = 0, on `parity`)
 * `reps` iterations run with HIP events between the kernels, then `reps`
 * without; us[0..4] = stencil D p, stencil + M^dag M p and its dots, CG
 * scalars, vector update, one whole iteration (the last without events). */
int milc_hip_dev_cg_iteration_bench(int slot_src, double mass, int parity, int reps, double* us);
/* STREAM-style device copy of `bytes` bytes, `reps` times (HIP events):
 * the measured HBM reference next to the vendor peak. */
int milc_hip_dev_copy_bench(size_t bytes, int reps, double* usec);
/* Batched stencil autotuner: times every variant (VALU threads-per-site
 * shapes and the MFMA 4x4x1 form) for prec (1 float, 2 double, 3 the 16-bit
 * records and links of the batched CG's sloppy operator) and nrhs
 * (4, 8, 16) right-hand sides and keeps the fastest; returns the number of
 * variants, the chosen index, its time and a description. */
int milc_hip_dev_block_tune(int prec, int nrhs, int* variant, double* usec, char* desc, int ndesc);
/* Description of stencil variant `variant` (threads per site, rhs per thread or
 * MFMA, block size, site loop, site-order block shape); returns the count. */
int milc_hip_dev_block_describe(int prec, int nrhs, int variant, char* desc, int ndesc);
/* Time `reps` batched stencil applications of one variant (HIP events). */
int milc_hip_dev_block_dslash_bench(int prec, int nrhs, int variant, int reps, double* usec);
/* dst_slots[s] = D src_slots[s] on out_parity for s < nsrc <= nrhs, through
 * the batched stencil (variant -1: the autotuned one). */
int milc_hip_dev_block_dslash(int prec, int nrhs, int variant, int nsrc, const int* src_slots, int out_parity,
                              const int* dst_slots);
/* Batched multi-shift on device slots (see milc_hip_multishift_invert_block);
 * solutions of the first out_nsrc sources to slots out_slot0 + s * num_offsets + j
 * when out_slot0 >= 0.
 * stats[0..5] = multi-shift iterations, multi-shift seconds, refinement
 * seconds, converged, stencil variant, its time (us); stats[6 + j] = max true
 * residual of shift j, stats[6 + num_offsets + j] = its refinement iterations.
 * Returns the multi-shift iteration count. */
int milc_hip_dev_multishift_block(int nsrc, const int* src_slots, int num_offsets, const double* offsets,
                                  double resid, double refine_resid, int parity, int max_iters, int out_slot0,
                                  int out_nsrc, double* stats);
/* One source on device slots, as milc_hip_multishift_invert does for host
 * arrays: FP64 multi-shift (mixed_precision 0) or FP32 multi-shift + mixed
 * per-shift refinement (1) to the true residual resid; solutions to slots
 * out_slot0 + j.  stats[0..2] = iterations, seconds, converged. */
int milc_hip_dev_multishift(int slot_src, int parity, int num_offsets, const double* offsets, double resid,
                            int mixed_precision, int out_slot0, double* stats);
/* One RHMC force evaluation on the current device links and the FP64 action
 * (the fermion part of update_h_rhmc.c:77-107 and d_action_rhmc.c):
 * X_t = (M^dag M + offsets[t] - 4m^2)^-1 phi (mixed_precision 1: FP32
 * multi-shift + mixed refinement to resid; 0: FP64), outer products O1/O3 with
 * coefficients residues[t]; thin_force != 0 (links built by
 * milc_hip_dev_gauge_hisq) carries them through the HISQ chain to the
 * thin-link momentum TA(2 U F_U); then S_act = sum_even |phi + sum_t
 * residues[t] X_t|^2 summed in FP64, its solve in the precision of the force
 * solve to the FP64 true residual min(resid, 1e-10).  compare_fp64 != 0
 * repeats the force in FP64 (target min(resid, 1e-12)) and the action with an
 * FP64 solve, and reports relative differences.  stats must hold 13 values:
 * stats[0..12] = force seconds, force iterations, sum_t r_t Re<phi, X_t>,
 * action, action seconds, action iterations, rel. error O1, rel. error O3,
 * FP64 force seconds, thin-link force seconds, rel. error of the thin-link
 * momentum, FP64-solve action, rel. error of the action (-1 when not
 * computed). */
int milc_hip_dev_hmc_step(int slot_phi, int n, const double* offsets, const double* residues, double resid,
                          int mixed_precision, int compare_fp64, int thin_force, double* stats);

/* HISQ thin-link force from device-resident solutions: the part of
 * update_h_rhmc.c:77-107 after the solve, i.e. eo_fermion_force_multi
 * (generic_ks/fermion_force_hisq_multi.c:1182-1475 through
 * fermion_force_fn_multi_hisq) with multi_x on the device.  Slots slot0 ..
 * slot0+n-1 hold X_t on even sites (FP64); the odd halves D_oe X_t are
 * computed here, the outer products take coefficients residues[t], and
 * TA(2 eps U F_U) goes through the HISQ chain on the links built by
 * milc_hip_dev_gauge_hisq.  momentum (nullable: FP64 [V][4][10] packed
 * anti_hermitmat, MILC site order) is ADDED to.  stats[0..3] = seconds for
 * D_oe + outer products, seconds for the thin-link chain, |momentum
 * increment|^2, seconds to rebuild the HISQ intermediate links V, W. */
int milc_hip_dev_hisq_force(int slot0, int n, const double* residues, double eps, void* momentum, double* stats);
/* |a - b| / |b| over the parity half held in two device slots. */
int milc_hip_dev_diff(int slot_a, int slot_b, double* rel);

#ifdef __cplusplus
}
#endif
#endif
