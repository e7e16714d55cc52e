/* ks_oracle.h -- FP64 CPU restatement of MILC's staggered (HISQ/asqtad) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in milc_qcd_amd/ links or calls it.
 *
 * Field layout follows MILC exactly (so buffers can be compared byte for byte
 * with a MILC build):
 *   - site order = node_index(): lexicographic i = x + nx(y + ny(z + nz t)),
 *     even sites (x+y+z+t even) first at i/2, odd sites at (i+V)/2
 *     (generic/layout_hyper_prime.c:509-520);
 *   - su3_vector  = 3 complex = 6 doubles, re/im interleaved
 *     (include/milc_datatypes.h:56);
 *   - su3_matrix  = e[3][3] complex = 18 doubles, row major
 *     (include/milc_datatypes.h:55);
 *   - link fields = su3_matrix[4*i + dir], dir = XUP,YUP,ZUP,TUP = 0..3
 *     (generic_ks/dslash_fn_dblstore.c:505, fat4 = &t_fatlink[4*i]).
 * Parities use MILC's codes: EVEN = 2, ODD = 1, EVENANDODD = 3
 * (include/macros.h:17-19).
 *
 * Pinning status of each function is recorded in the header comment of the
 * function in ks_oracle.c ("pinned by ..." / "parity unpinned").
 */
#ifndef KS_ORACLE_H
#define KS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KSO_EVEN 2
#define KSO_ODD 1
#define KSO_EVENANDODD 3

/* Lattice geometry with precomputed neighbour tables.
 * nbr[k][i], k = 0..7  : one-hop neighbours in MILC direction order
 *                        XUP,YUP,ZUP,TUP,TDOWN,ZDOWN,YDOWN,XDOWN (include/dirs.h:8-20)
 * nbr[k][i], k = 8..15 : three-hop neighbours X3UP..T3UP, T3DOWN..X3DOWN
 *                        (include/dirs.h:22-34). */
typedef struct {
  int dims[4];
  size_t V, Vh;
  int *coords;      /* [V][4] x,y,z,t of MILC site index i */
  size_t *nbr;      /* [16][V] */
} kso_geom;

int  kso_geom_init(kso_geom *g, const int dims[4]);
void kso_geom_free(kso_geom *g);
size_t kso_node_index(const kso_geom *g, int x, int y, int z, int t);

/* --- gauge field generation and phases ---------------------------------- */
void kso_gauge_random(const kso_geom *g, uint64_t seed, double scale, double *links);
void kso_gaussian_vector(const kso_geom *g, uint64_t seed, int parity, double *v);
void kso_rephase(const kso_geom *g, double *links);
double kso_plaquette(const kso_geom *g, const double *links);

/* --- link smearing ------------------------------------------------------- */
/* coeff[6] = {one_link, naik, three_staple, five_staple, seven_staple, lepage}
 * (asqtad-style act_path_coeff, generic_ks/ks_action_paths_hisq.c:130-175) */
void kso_fat_links(const kso_geom *g, const double coeff[6], const double *U, double *fat);
void kso_long_links(const kso_geom *g, double naik, const double *U, double *lng);
int  kso_u3_unitarize(const double *V, double *W);
/* Full HISQ U(3) chain: U -> V (fat7) -> Y=W (U(3), analytic) -> (fat, lng).
 * Returns the number of links whose analytic unitarisation would have been
 * routed to MILC's SVD fallback. */
int  kso_hisq_links(const kso_geom *g, const double *U, double *fat, double *lng,
                    double *W_out);
/* Asqtad links with tadpole factor u0, as check_ks_invert builds them. */
void kso_asqtad_links(const kso_geom *g, double u0, const double *U, double *fat, double *lng);

/* --- Dslash and solvers --------------------------------------------------- */
void kso_dslash(const kso_geom *g, const double *fat, const double *lng,
                const double *src, double *dst, int parity);

typedef struct {
  int max, nrestart, parity;
  double resid, relresid;
  /* outputs */
  double final_rsq, final_relrsq, size_r, size_relr;
  int final_iters, final_restart, converged;
} kso_qic;

int kso_congrad(const kso_geom *g, const double *fat, const double *lng,
                const double *src, double *dest, kso_qic *qic, double mass);
int kso_multicg_offset(const kso_geom *g, const double *fat, const double *lng,
                       const double *src, double **psim, const double *offsets,
                       int num_offsets, kso_qic *qic);

/* --- fermion force level 0 (outer products) ------------------------------- */
void kso_force_oprod(const kso_geom *g, const double *const *multi_x, const double *residues,
                     int nterms, double *oprod1, double *oprod3);

/* FN (fat + Naik) thin-link force, fermion_force_fn_multi: mom = TA(2 eps U G),
 * stored as full anti-Hermitian 3x3 matrices [V][4][18]. */
void kso_fn_force(const kso_geom *g, const double coeff[6], const double *U, const double *const *multi_x,
                  const double *residues, int nterms, double eps, double *mom);
/* HISQ thin-link force (level 0 -> level-2 adjoint on W -> U(3) projection
 * adjoint -> fat7 adjoint on U): mom = TA(2 eps U F_U), full matrices. */
/* Symanzik-improved gauge action, its momentum force, momentum action, link update */
void kso_gauge_coeffs(double u0, int nf, int hisq, double coeff[3]);
int kso_gauge_loops(int *dirs, int *len, int *shape);
double kso_gauge_action(const kso_geom *g, const double coeff[3], const double *U);
void kso_gauge_force(const kso_geom *g, const double coeff[3], double eb3, const double *U, double *mom);
double kso_mom_action(const kso_geom *g, const double *mom);
void kso_update_u(const kso_geom *g, double eps, double *U, const double *mom);
void kso_hisq_force(const kso_geom *g, const double *U, const double *const *multi_x,
                    const double *residues, int nterms, double eps, double *mom);
/* the same with the last n_naik of the nterms terms in a second Naik set
 * (links X_2 + eps_naik X_3) */
void kso_hisq_force_naik(const kso_geom *g, const double *U, const double *const *multi_x,
                         const double *residues, int nterms, int n_naik, double eps_naik, double eps,
                         double *mom);
/* links of a second Naik set from the level-2 links and W:
 * fat += eps_naik/8 W, lng += eps_naik lng (the Naik paths of both tables
 * have coefficient -1/24) */
void kso_hisq_links_naik(const kso_geom *g, double eps_naik, const double *W, double *fat, double *lng);
/* 0: force adjoints through the path tables (default), 1: staple form */
void kso_set_force_form(int form);
/* force through W = V (V^dag V)^-1/2 at one link: F_V from F_W */
void kso_u3_adjoint(const double *V, const double *FW, double *FV);
/* the force's path table for direction mu (dirs[7*p+i], -1 padded) */
int kso_fn_paths(int mu, const double coeff[6], int *dirs, double *c, int *len, int *naik);

/* MILC's per-site generator (9 words per site, node_index order), its
 * momentum and Gaussian-vector draws (single = 1: MILC_PRECISION=1 rounding),
 * and reunit_su3 */
void kso_prn_seed(const kso_geom *g, int iseed, uint64_t *st);
void kso_prn_seed_one(int seed, int index, uint64_t *st);
uint32_t kso_prn_next(uint64_t *st);
void kso_ranmom_prn(const kso_geom *g, uint64_t *st, int single, double *mom);
double kso_gaussian_prn(const kso_geom *g, uint64_t *st, int single, int parity, double *v);
int kso_reunit_su3(double *M, double tol, double *maxdev);
int kso_reunitarize(const kso_geom *g, double tol, double *U, double *maxdev);

/* --- reductions and helpers used by tests ---------------------------------- */
double kso_norm2(const kso_geom *g, const double *v, int parity);
void kso_mdagm(const kso_geom *g, const double *fat, const double *lng, double offset,
               const double *src, double *dst, double *tmp, int parity);

#ifdef __cplusplus
}
#endif
#endif
