/* ks_oracle.c -- FP64 CPU restatement of MILC's staggered hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see ks_oracle.h.
 *
 * Every function cites the MILC 7.8.1 source it restates.  Pinning:
 *   kso_dslash        pinned by: MILC golden (M^dag M)^-1 src on 8^4 asqtad
 *                     (binary_samples/ksvector.ans.MdaggerMinv.sample.l8),
 *                     free-field closed form, anti-Hermiticity, gauge
 *                     covariance, compiled MILC dslash_fn_field (oracle/_ref).
 *   kso_congrad       pinned by: the same golden, dense solve on 4^4,
 *                     iteration-for-iteration against compiled MILC.
 *   kso_multicg_offset pinned by: per-shift true residuals, agreement with
 *                     independent single-shift CG, compiled MILC.
 *   kso_fat_links / kso_long_links / kso_asqtad_links
 *                     pinned by: MILC golden asqtad fat links
 *                     (binary_samples/fat.l8 from lat.sample.l8, u0=0.502),
 *                     free-field sums of coefficients, gauge covariance.
 *   kso_u3_unitarize  pinned by: unitarity, polar-decomposition identity
 *                     W = V (V^dag V)^-1/2 checked against numpy SVD.
 *   kso_hisq_links    pinned by: compiled MILC load_hisq_aux_links/load_X
 *                     (oracle/_ref) and the free-field closed form.
 *   kso_force_oprod   pinned by: central finite differences of the action,
 *                     compiled MILC smearing0 outer products.
 *   kso_gauge_random  pinned by: unitarity, det = 1, numpy expm of i*0.4*H.
 */
#include "ks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* complex 3x3 helpers on double[18] (row-major e[i][j], re/im)               */
/* ------------------------------------------------------------------------ */
#define RE(m, i, j) (m)[2 * (3 * (i) + (j))]
#define IM(m, i, j) (m)[2 * (3 * (i) + (j)) + 1]

static inline void m_nn(const double *a, const double *b, double *c) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, i, k) * RE(b, k, j) - IM(a, i, k) * IM(b, k, j);
        im += RE(a, i, k) * IM(b, k, j) + IM(a, i, k) * RE(b, k, j);
      }
      RE(t, i, j) = re; IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}
/* c = a * b^dag  (libraries/m_mat_na.c) */
static inline void m_na(const double *a, const double *b, double *c) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, i, k) * RE(b, j, k) + IM(a, i, k) * IM(b, j, k);
        im += IM(a, i, k) * RE(b, j, k) - RE(a, i, k) * IM(b, j, k);
      }
      RE(t, i, j) = re; IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}
/* c = a^dag * b  (libraries/m_mat_an.c) */
static inline void m_an(const double *a, const double *b, double *c) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += RE(a, k, i) * RE(b, k, j) + IM(a, k, i) * IM(b, k, j);
        im += RE(a, k, i) * IM(b, k, j) - IM(a, k, i) * RE(b, k, j);
      }
      RE(t, i, j) = re; IM(t, i, j) = im;
    }
  memcpy(c, t, sizeof t);
}
static inline void m_adj(const double *a, double *c) {
  double t[18];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) { RE(t, i, j) = RE(a, j, i); IM(t, i, j) = -IM(a, j, i); }
  memcpy(c, t, sizeof t);
}
static inline void m_axpy(double *y, double s, const double *x) {
  for (int k = 0; k < 18; k++) y[k] += s * x[k];
}
/* out (+)= a * v */
static inline void mv(const double *a, const double *v, double *out, int accumulate) {
  double t[6];
  for (int i = 0; i < 3; i++) {
    double re = 0, im = 0;
    for (int k = 0; k < 3; k++) {
      re += RE(a, i, k) * v[2 * k] - IM(a, i, k) * v[2 * k + 1];
      im += RE(a, i, k) * v[2 * k + 1] + IM(a, i, k) * v[2 * k];
    }
    t[2 * i] = re; t[2 * i + 1] = im;
  }
  if (accumulate) for (int k = 0; k < 6; k++) out[k] += t[k];
  else memcpy(out, t, sizeof t);
}
/* out (+)= a^dag * v */
static inline void amv(const double *a, const double *v, double *out, int accumulate) {
  double t[6];
  for (int i = 0; i < 3; i++) {
    double re = 0, im = 0;
    for (int k = 0; k < 3; k++) {
      re += RE(a, k, i) * v[2 * k] + IM(a, k, i) * v[2 * k + 1];
      im += RE(a, k, i) * v[2 * k + 1] - IM(a, k, i) * v[2 * k];
    }
    t[2 * i] = re; t[2 * i + 1] = im;
  }
  if (accumulate) for (int k = 0; k < 6; k++) out[k] += t[k];
  else memcpy(out, t, sizeof t);
}

/* ------------------------------------------------------------------------ */
/* geometry (generic/layout_hyper_prime.c:509-520, include/dirs.h)            */
/* ------------------------------------------------------------------------ */
size_t kso_node_index(const kso_geom *g, int x, int y, int z, int t) {
  const int *d = g->dims;
  x = ((x % d[0]) + d[0]) % d[0];
  y = ((y % d[1]) + d[1]) % d[1];
  z = ((z % d[2]) + d[2]) % d[2];
  t = ((t % d[3]) + d[3]) % d[3];
  size_t i = (size_t)x + (size_t)d[0] * ((size_t)y + (size_t)d[1] * ((size_t)z + (size_t)d[2] * (size_t)t));
  if ((x + y + z + t) % 2 == 0) return i / 2;
  return (i + g->V) / 2;
}

int kso_geom_init(kso_geom *g, const int dims[4]) {
  memcpy(g->dims, dims, sizeof g->dims);
  g->V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  g->Vh = g->V / 2;
  if (dims[0] % 2 || g->V % 2) return -1;
  g->coords = (int *)malloc(sizeof(int) * 4 * g->V);
  g->nbr = (size_t *)malloc(sizeof(size_t) * 16 * g->V);
  if (!g->coords || !g->nbr) return -2;
  for (int t = 0; t < dims[3]; t++)
    for (int z = 0; z < dims[2]; z++)
      for (int y = 0; y < dims[1]; y++)
        for (int x = 0; x < dims[0]; x++) {
          size_t i = kso_node_index(g, x, y, z, t);
          int *c = g->coords + 4 * i;
          c[0] = x; c[1] = y; c[2] = z; c[3] = t;
        }
#pragma omp parallel for
  for (size_t i = 0; i < g->V; i++) {
    const int *c = g->coords + 4 * i;
    for (int dir = 0; dir < 4; dir++) {
      for (int hop = 1; hop <= 3; hop += 2) {
        int cp[4], cm[4];
        memcpy(cp, c, sizeof cp); memcpy(cm, c, sizeof cm);
        cp[dir] += hop; cm[dir] -= hop;
        size_t ip = kso_node_index(g, cp[0], cp[1], cp[2], cp[3]);
        size_t im = kso_node_index(g, cm[0], cm[1], cm[2], cm[3]);
        int base = hop == 1 ? 0 : 8;
        g->nbr[(size_t)(base + dir) * g->V + i] = ip;         /* XUP..TUP / X3UP..T3UP */
        g->nbr[(size_t)(base + 7 - dir) * g->V + i] = im;     /* OPP_DIR = 7-dir */
      }
    }
  }
  return 0;
}

void kso_geom_free(kso_geom *g) {
  free(g->coords); free(g->nbr);
  g->coords = NULL; g->nbr = NULL;
}

#define NB(g, k, i) ((g)->nbr[(size_t)(k) * (g)->V + (i)])

static void parity_range(const kso_geom *g, int parity, size_t *lo, size_t *hi) {
  switch (parity) {
    case KSO_EVEN: *lo = 0; *hi = g->Vh; break;
    case KSO_ODD: *lo = g->Vh; *hi = g->V; break;
    default: *lo = 0; *hi = g->V; break;
  }
}

/* ------------------------------------------------------------------------ */
/* counter-based RNG shared (by specification) with the HIP generator       */
/*   u64 draw(seed, ctr) = mix(mix(seed ^ K0) + ctr), mix = splitmix64 final */
/*   uniform = ((draw >> 11) + 0.5) * 2^-53                                   */
/*   gaussian pair p: Box-Muller on uniforms 2p, 2p+1                         */
/* ------------------------------------------------------------------------ */
static inline uint64_t sm_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline double uni(uint64_t key, uint64_t ctr) {
  uint64_t r = sm_mix(key + ctr);
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static inline void gauss_pair(uint64_t key, uint64_t p, double *g0, double *g1) {
  double u1 = uni(key, 2 * p), u2 = uni(key, 2 * p + 1);
  double r = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586476925286766559 * u2;
  *g0 = r * cos(th);
  *g1 = r * sin(th);
}

/* exp(X) for anti-Hermitian 3x3 X: 4 halvings, degree-14 Taylor (Horner),
 * 4 squarings.  Only + and * after the Gaussians, so CPU and GPU agree to
 * rounding. */
static void expm3(const double *X, double *E) {
  double Y[18], T[18];
  for (int k = 0; k < 18; k++) Y[k] = X[k] * 0.0625;
  /* Horner: E = I + Y(I + Y/2(I + Y/3(...))) */
  memset(E, 0, sizeof(double) * 18);
  RE(E, 0, 0) = RE(E, 1, 1) = RE(E, 2, 2) = 1.0;
  for (int n = 14; n >= 1; n--) {
    m_nn(Y, E, T);
    for (int k = 0; k < 18; k++) E[k] = T[k] / (double)n;
    RE(E, 0, 0) += 1.0; RE(E, 1, 1) += 1.0; RE(E, 2, 2) += 1.0;
  }
  for (int s = 0; s < 4; s++) { m_nn(E, E, T); memcpy(E, T, sizeof T); }
}

/* U = exp(i*scale*H), H = (A+A^dag)/2 - Tr/3, A_jk ~ N(0,1)+iN(0,1); per link
 * (lexicographic site s, dir mu) the 9 Gaussian pairs use p = (4s+mu)*9+n.
 * Parity unpinned against MILC (MILC has no such generator): pinned by
 * unitarity, det U = 1 and numpy's expm of the same H in tests. */
void kso_gauge_random(const kso_geom *g, uint64_t seed, double scale, double *links) {
  uint64_t key = sm_mix(seed ^ 0x5851F42D4C957F2DULL);
#pragma omp parallel for
  for (size_t i = 0; i < g->V; i++) {
    const int *c = g->coords + 4 * i;
    uint64_t s = (uint64_t)c[0] + (uint64_t)g->dims[0] * ((uint64_t)c[1] + (uint64_t)g->dims[1] *
                 ((uint64_t)c[2] + (uint64_t)g->dims[2] * (uint64_t)c[3]));
    for (int mu = 0; mu < 4; mu++) {
      double A[18], H[18], X[18];
      for (int n = 0; n < 9; n++) gauss_pair(key, (4 * s + mu) * 9 + n, &A[2 * n], &A[2 * n + 1]);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          RE(H, a, b) = 0.5 * (RE(A, a, b) + RE(A, b, a));
          IM(H, a, b) = 0.5 * (IM(A, a, b) - IM(A, b, a));
        }
      double tr = (RE(H, 0, 0) + RE(H, 1, 1) + RE(H, 2, 2)) / 3.0;
      RE(H, 0, 0) -= tr; RE(H, 1, 1) -= tr; RE(H, 2, 2) -= tr;
      /* X = i*scale*H */
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          RE(X, a, b) = -scale * IM(H, a, b);
          IM(X, a, b) = scale * RE(H, a, b);
        }
      expm3(X, links + (4 * i + mu) * 18);
    }
  }
}

/* Gaussian complex vector on one parity, <re^2> = <im^2> = 1/2 as in
 * grsource_plain_field (generic_ks/grsource_imp.c:169-195); site s (lex) uses
 * pairs p = 3s + c.  Other-parity sites are zeroed. */
void kso_gaussian_vector(const kso_geom *g, uint64_t seed, int parity, double *v) {
  uint64_t key = sm_mix(seed ^ 0x2545F4914F6CDD1DULL);
  size_t lo, hi;
  parity_range(g, parity, &lo, &hi);
  const double s2 = 0.70710678118654752440;
#pragma omp parallel for
  for (size_t i = 0; i < g->V; i++) {
    double *vi = v + 6 * i;
    if (i < lo || i >= hi) { memset(vi, 0, 6 * sizeof(double)); continue; }
    const int *c = g->coords + 4 * i;
    uint64_t s = (uint64_t)c[0] + (uint64_t)g->dims[0] * ((uint64_t)c[1] + (uint64_t)g->dims[1] *
                 ((uint64_t)c[2] + (uint64_t)g->dims[2] * (uint64_t)c[3]));
    for (int k = 0; k < 3; k++) {
      double a, b;
      gauss_pair(key, 3 * s + k, &a, &b);
      vi[2 * k] = s2 * a; vi[2 * k + 1] = s2 * b;
    }
  }
}

/* KS phases and antiperiodic time boundary (generic_ks/rephase.c:64-118);
 * multiplying twice restores the input, as rephase(OFF) does. */
void kso_rephase(const kso_geom *g, double *links) {
#pragma omp parallel for
  for (size_t i = 0; i < g->V; i++) {
    const int *c = g->coords + 4 * i;
    double ph[4];
    ph[3] = 1.0;
    ph[0] = (c[3] % 2 == 1) ? -1.0 : 1.0;
    ph[1] = (c[0] % 2 == 1) ? -ph[0] : ph[0];
    ph[2] = (c[1] % 2 == 1) ? -ph[1] : ph[1];
    if (c[3] == g->dims[3] - 1) ph[3] = -ph[3];
    for (int mu = 0; mu < 4; mu++) {
      double *m = links + (4 * i + mu) * 18;
      for (int k = 0; k < 18; k++) m[k] *= ph[mu];
    }
  }
}

/* Average plaquette Re Tr/3 over all 6 planes (for test diagnostics). */
double kso_plaquette(const kso_geom *g, const double *L) {
  double sum = 0;
#pragma omp parallel for reduction(+ : sum)
  for (size_t i = 0; i < g->V; i++) {
    for (int mu = 0; mu < 4; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        double a[18], b[18];
        m_nn(L + (4 * i + mu) * 18, L + (4 * NB(g, mu, i) + nu) * 18, a);
        m_nn(L + (4 * i + nu) * 18, L + (4 * NB(g, nu, i) + mu) * 18, b);
        m_na(a, b, a);
        sum += (RE(a, 0, 0) + RE(a, 1, 1) + RE(a, 2, 2)) / 3.0;
      }
  }
  return sum / (6.0 * g->V);
}

/* ------------------------------------------------------------------------ */
/* link smearing                                                             */
/* ------------------------------------------------------------------------ */

/* Generalised staple (generic/general_staple.c:42-124): for the "mu link"
 * field L (stride floats of 18 per site) and thin links U,
 *   upper(x) = U_nu(x) [L(x+nu) U_nu(x+mu)^dag]
 *   lower(x) = [U_nu(x-nu)^dag L(x-nu)] U_nu(x-nu+mu)
 * If staple != NULL: staple = upper + lower and fat_mu += coef*staple;
 * else fat_mu += coef*upper then fat_mu += coef*lower. */
static void gen_staple(const kso_geom *g, double *staple, int mu, int nu, const double *L,
                       int stride, double *fat, double coef, const double *U, double *tmp) {
  const size_t V = g->V;
#pragma omp parallel for
  for (size_t i = 0; i < V; i++) {
    double t1[18], t2[18];
    size_t xn = NB(g, nu, i), xm = NB(g, mu, i);
    m_na(L + (size_t)stride * 18 * xn, U + (4 * xm + nu) * 18, t1);
    if (staple) m_nn(U + (4 * i + nu) * 18, t1, staple + 18 * i);
    else {
      m_nn(U + (4 * i + nu) * 18, t1, t2);
      m_axpy(fat + (4 * i + mu) * 18, coef, t2);
    }
  }
#pragma omp parallel for
  for (size_t i = 0; i < V; i++) {
    double t1[18];
    m_an(U + (4 * i + nu) * 18, L + (size_t)stride * 18 * i, t1);
    m_nn(t1, U + (4 * NB(g, mu, i) + nu) * 18, tmp + 18 * i);
  }
#pragma omp parallel for
  for (size_t i = 0; i < V; i++) {
    const double *low = tmp + 18 * NB(g, 7 - nu, i);
    double *f = fat + (4 * i + mu) * 18;
    if (staple) {
      for (int k = 0; k < 18; k++) staple[18 * i + k] += low[k];
      m_axpy(f, coef, staple + 18 * i);
    } else {
      m_axpy(f, coef, low);
    }
  }
}

/* Optimised asqtad fattening (generic_ks/fermion_links_fn_load_milc.c:222-252,
 * ASQ_OPTIMIZED_FATTENING): one-link term corrected by -6*lepage because the
 * Lepage staple-of-staple contains six retracing paths equal to U_mu. */
void kso_fat_links(const kso_geom *g, const double coeff[6], const double *U, double *fat) {
  const size_t V = g->V;
  double one_link = coeff[0] - 6.0 * coeff[5];
  double *staple = (double *)malloc(sizeof(double) * 18 * V);
  double *tempmat1 = (double *)malloc(sizeof(double) * 18 * V);
  double *tmp = (double *)malloc(sizeof(double) * 18 * V);
  for (int dir = 0; dir < 4; dir++) {
#pragma omp parallel for
    for (size_t i = 0; i < V; i++) {
      const double *u = U + (4 * i + dir) * 18;
      double *f = fat + (4 * i + dir) * 18;
      for (int k = 0; k < 18; k++) f[k] = one_link * u[k];
    }
    if (coeff[2] == 0.0 && coeff[5] == 0.0 && coeff[3] == 0.0) continue;
    for (int nu = 0; nu < 4; nu++) {
      if (nu == dir) continue;
      gen_staple(g, staple, dir, nu, U + 18 * dir, 4, fat, coeff[2], U, tmp);
      gen_staple(g, NULL, dir, nu, staple, 1, fat, coeff[5], U, tmp);
      for (int rho = 0; rho < 4; rho++) {
        if (rho == dir || rho == nu) continue;
        gen_staple(g, tempmat1, dir, rho, staple, 1, fat, coeff[3], U, tmp);
        for (int sig = 0; sig < 4; sig++) {
          if (sig == dir || sig == nu || sig == rho) continue;
          gen_staple(g, NULL, dir, sig, tempmat1, 1, fat, coeff[4], U, tmp);
        }
      }
    }
  }
  free(staple); free(tempmat1); free(tmp);
}

/* Naik links (generic_ks/fermion_links_fn_load_milc.c:51-131): the only path
 * with net displacement 3*mu in the HISQ/asqtad tables is the straight one,
 * lng_mu(x) = naik * U_mu(x) U_mu(x+mu) U_mu(x+2mu). */
void kso_long_links(const kso_geom *g, double naik, const double *U, double *lng) {
#pragma omp parallel for
  for (size_t i = 0; i < g->V; i++) {
    for (int mu = 0; mu < 4; mu++) {
      size_t x1 = NB(g, mu, i), x2 = NB(g, mu, x1);
      double t[18];
      m_nn(U + (4 * x1 + mu) * 18, U + (4 * x2 + mu) * 18, t);
      m_nn(U + (4 * i + mu) * 18, t, t);
      double *l = lng + (4 * i + mu) * 18;
      for (int k = 0; k < 18; k++) l[k] = naik * t[k];
    }
  }
}

/* Analytic U(3) projection W = V (V^dag V)^-1/2
 * (generic_ks/su3_mat_op.c:828-1211, Hasenfratz-Hoffmann-Schaefer):
 * eigenvalues g_k of Q = V^dag V from the trigonometric cubic solution, then
 * Q^-1/2 = f0 + f1 Q + f2 Q^2.  Returns 1 when MILC's SVD fallback condition
 * (relative det mismatch > 1e-8 or |det|^2 < 1e-8) holds; the analytic result
 * is still returned then (callers count and report these links). */
int kso_u3_unitarize(const double *V, double *W) {
  double Q[18], Q2[18];
  m_an(V, V, Q);
  m_nn(Q, Q, Q2);
  double c0 = RE(Q, 0, 0) + RE(Q, 1, 1) + RE(Q, 2, 2);
  double c1 = (RE(Q2, 0, 0) + RE(Q2, 1, 1) + RE(Q2, 2, 2)) / 2;
  double q3 = 0;
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) q3 += RE(Q2, i, k) * RE(Q, k, i) - IM(Q2, i, k) * IM(Q, k, i);
  double c2 = q3 / 3;
  double g0, g1, g2;
  double S = c1 / 3 - c0 * (c0 / 18);
  if (fabs(S) < 1.0e-14) {
    g0 = g1 = g2 = c0 / 3;
  } else {
    double R = c2 / 2 - c0 * (c1 / 3) + c0 * c0 * (c0 / 27);
    S = sqrt(S);
    double RoS = R / (S * S * S);
    double theta;
    if (!(fabs(RoS) < 1.0)) theta = R > 0 ? 0.0 : 3.14159265358979323846264338328;
    else theta = acos(RoS);
    double th3 = theta / 3, pi23 = 6.28318530717958647692528676656 / 3;
    g0 = c0 / 3 + 2 * S * cos(th3);
    g1 = c0 / 3 + 2 * S * cos(th3 + pi23);
    g2 = c0 / 3 + 2 * S * cos(th3 + 2 * pi23);
  }
  /* determinant check used by MILC to decide on the SVD fallback */
  double dre = 0, dim = 0;
  {
    double a1re = RE(V,1,1)*RE(V,2,2) - IM(V,1,1)*IM(V,2,2) - RE(V,1,2)*RE(V,2,1) + IM(V,1,2)*IM(V,2,1);
    double a1im = RE(V,1,1)*IM(V,2,2) + IM(V,1,1)*RE(V,2,2) - RE(V,1,2)*IM(V,2,1) - IM(V,1,2)*RE(V,2,1);
    double a2re = RE(V,1,0)*RE(V,2,2) - IM(V,1,0)*IM(V,2,2) - RE(V,1,2)*RE(V,2,0) + IM(V,1,2)*IM(V,2,0);
    double a2im = RE(V,1,0)*IM(V,2,2) + IM(V,1,0)*RE(V,2,2) - RE(V,1,2)*IM(V,2,0) - IM(V,1,2)*RE(V,2,0);
    double a3re = RE(V,1,0)*RE(V,2,1) - IM(V,1,0)*IM(V,2,1) - RE(V,1,1)*RE(V,2,0) + IM(V,1,1)*IM(V,2,0);
    double a3im = RE(V,1,0)*IM(V,2,1) + IM(V,1,0)*RE(V,2,1) - RE(V,1,1)*IM(V,2,0) - IM(V,1,1)*RE(V,2,0);
    dre = RE(V,0,0)*a1re - IM(V,0,0)*a1im - RE(V,0,1)*a2re + IM(V,0,1)*a2im + RE(V,0,2)*a3re - IM(V,0,2)*a3im;
    dim = IM(V,0,0)*a1re + RE(V,0,0)*a1im - IM(V,0,1)*a2re - RE(V,0,1)*a2im + IM(V,0,2)*a3re + RE(V,0,2)*a3im;
  }
  double det_check = dre * dre + dim * dim;
  int svd = 0;
  if (det_check != 0 && fabs(det_check - g0 * g1 * g2) / fabs(det_check) > 1e-8) svd = 1;
  if (det_check < 1e-8) svd = 1;

  double g0s = sqrt(g0), g1s = sqrt(g1), g2s = sqrt(g2);
  double us = g1s + g2s, ws = g1s * g2s;
  double vs = g0s * us + ws;
  us += g0s;
  ws *= g0s;
  double denom = ws * (us * vs - ws);
  double f0 = (us * vs * vs - ws * (us * us + vs)) / denom;
  double f1 = (2 * us * vs - ws - us * us * us) / denom;
  double f2 = us / denom;
  double S2[18];
  for (int k = 0; k < 18; k++) S2[k] = f1 * Q[k] + f2 * Q2[k];
  RE(S2, 0, 0) += f0; RE(S2, 1, 1) += f0; RE(S2, 2, 2) += f0;
  m_nn(V, S2, W);
  return svd;
}

/* HISQ U(3) link chain (generic_ks/fermion_links_hisq_load_milc.c:546-702)
 * with the path coefficients of generic_ks/imp_actions/hisq/hisq_u3_action.h:
 *   level 1 (fat7):  one-link 1/8, 3-staple -1/16, 5-staple 1/64, 7-staple -1/384
 *   level 2:         one-link 1, Naik -1/24, 3-staple -1/16, 5-staple 1/64,
 *                    7-staple -1/384, Lepage -1/8;
 * unitarisation with phases out (rephase_field_offset OFF/ON around
 * u3_unitarize_analytic, :251-270), U(3) group so W = Y (WeqY).
 * eps_naik = 0, a single Naik set. */
/* load_hisq_fn_links with n_naiks = 2 (generic_ks/fermion_links_hisq_load_milc.c:
 * 574-640): X_3 = the "1-link + Naik" table on W, set 1 = X_2 + eps_naik X_3;
 * fat/lng hold X_2 on entry and set 1 on return */
void kso_hisq_links_naik(const kso_geom *g, double eps_naik, const double *W, double *fat, double *lng) {
  const size_t n = 4 * g->V * 18;
  const double c3[6] = {1.0 / 8.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double *f3 = (double *)malloc(sizeof(double) * n);
  double *l3 = (double *)malloc(sizeof(double) * n);
  kso_fat_links(g, c3, W, f3);
  kso_long_links(g, -1.0 / 24.0, W, l3);
  for (size_t k = 0; k < n; k++) {
    fat[k] += eps_naik * f3[k];
    lng[k] += eps_naik * l3[k];
  }
  free(f3);
  free(l3);
}

int kso_hisq_links(const kso_geom *g, const double *U, double *fat, double *lng, double *W_out) {
  const size_t n = 4 * g->V * 18;
  double *Vl = (double *)malloc(sizeof(double) * n);
  double *W = W_out ? W_out : (double *)malloc(sizeof(double) * n);
  const double c1[6] = {1.0 / 8.0, 0.0, -1.0 / 8.0 * 0.5, 1.0 / 8.0 * 0.25 * 0.5,
                        -1.0 / 8.0 * 0.125 * (1.0 / 6.0), 0.0};
  const double c2[6] = {(1.0 / 8.0) + (2.0 * 6.0 / 16.0) + (1.0 / 8.0), -1.0 / 24.0,
                        -1.0 / 8.0 * 0.5, 1.0 / 8.0 * 0.25 * 0.5,
                        -1.0 / 8.0 * 0.125 * (1.0 / 6.0), -2.0 / 16.0};
  kso_fat_links(g, c1, U, Vl);
  kso_rephase(g, Vl);
  int nsvd = 0;
#pragma omp parallel for reduction(+ : nsvd)
  for (size_t l = 0; l < 4 * g->V; l++) nsvd += kso_u3_unitarize(Vl + 18 * l, W + 18 * l);
  kso_rephase(g, W);
  kso_fat_links(g, c2, W, fat);
  kso_long_links(g, c2[1], W, lng);
  free(Vl);
  if (!W_out) free(W);
  return nsvd;
}

/* Asqtad links as ks_imp_utilities/check_ks_invert builds them
 * (generic_ks/imp_actions/asqtad_action.h path coefficients, tadpole
 * factors u0^-(length-1), ks_action_paths.c build_act_path_coeff). */
void kso_asqtad_links(const kso_geom *g, double u0, const double *U, double *fat, double *lng) {
  double u2 = u0 * u0, u4 = u2 * u2, u6 = u4 * u2;
  const double c[6] = {5.0 / 8.0, -1.0 / 24.0 / u2, -1.0 / 16.0 / u2, 1.0 / 64.0 / u4,
                       -1.0 / 384.0 / u6, -1.0 / 16.0 / u4};
  kso_fat_links(g, c, U, fat);
  kso_long_links(g, c[1], U, lng);
}

/* ------------------------------------------------------------------------ */
/* Dslash (generic_ks/dslash_fn_dblstore.c:311-563)                           */
/*   dst(x) = sum_mu F_mu(x) s(x+mu) + L_mu(x) s(x+3mu)                       */
/*          - F_mu(x-mu)^dag s(x-mu) - L_mu(x-3mu)^dag s(x-3mu)               */
/* on sites of `parity`; the back links are the adjoints of the neighbours'    */
/* forward links (fn_links_milc.c:115-195).  src and dst may alias (the two   */
/* parities are disjoint halves).                                             */
/* ------------------------------------------------------------------------ */
void kso_dslash(const kso_geom *g, const double *fat, const double *lng, const double *src,
                double *dst, int parity) {
  size_t lo, hi;
  parity_range(g, parity, &lo, &hi);
#pragma omp parallel for
  for (size_t i = lo; i < hi; i++) {
    double fwd[6], t[6], b[6], bl[6];
    /* forward fat 4dir sum, then forward long 4dir sum (tvec), then add */
    mv(fat + (4 * i + 0) * 18, src + 6 * NB(g, 0, i), fwd, 0);
    for (int mu = 1; mu < 4; mu++) mv(fat + (4 * i + mu) * 18, src + 6 * NB(g, mu, i), fwd, 1);
    mv(lng + (4 * i + 0) * 18, src + 6 * NB(g, 8, i), t, 0);
    for (int mu = 1; mu < 4; mu++) mv(lng + (4 * i + mu) * 18, src + 6 * NB(g, 8 + mu, i), t, 1);
    for (int k = 0; k < 6; k++) fwd[k] += t[k];
    /* backward: fatback(x,mu) = F_mu(x-mu)^dag, lngback(x,mu) = L_mu(x-3mu)^dag */
    for (int mu = 0; mu < 4; mu++) {
      size_t xm = NB(g, 7 - mu, i), x3m = NB(g, 15 - mu, i);
      amv(fat + (4 * xm + mu) * 18, src + 6 * xm, b, mu > 0);
      amv(lng + (4 * x3m + mu) * 18, src + 6 * x3m, bl, mu > 0);
    }
    for (int k = 0; k < 6; k++) { fwd[k] -= b[k]; fwd[k] -= bl[k]; }
    memcpy(dst + 6 * i, fwd, sizeof fwd);
  }
}

double kso_norm2(const kso_geom *g, const double *v, int parity) {
  size_t lo, hi;
  parity_range(g, parity, &lo, &hi);
  double s = 0;
#pragma omp parallel for reduction(+ : s)
  for (size_t i = lo; i < hi; i++)
    for (int k = 0; k < 6; k++) s += v[6 * i + k] * v[6 * i + k];
  return s;
}

/* dst = (offset - D_{p,p'} D_{p',p}) src on `parity`, i.e. (M^dag M) for
 * offset = 4m^2 (d_congrad5_fn_milc.c:196-211 computes the negative). */
void kso_mdagm(const kso_geom *g, const double *fat, const double *lng, double offset,
               const double *src, double *dst, double *tmp, int parity) {
  int other = parity == KSO_EVEN ? KSO_ODD : KSO_EVEN;
  kso_dslash(g, fat, lng, src, tmp, other);
  kso_dslash(g, fat, lng, tmp, tmp, parity);
  size_t lo, hi;
  parity_range(g, parity, &lo, &hi);
#pragma omp parallel for
  for (size_t i = lo; i < hi; i++)
    for (int k = 0; k < 6; k++) dst[6 * i + k] = offset * src[6 * i + k] - tmp[6 * i + k];
}

/* ------------------------------------------------------------------------ */
/* CG (generic_ks/d_congrad5_fn_milc.c:60-407, FEWSUMS variant)               */
/* ------------------------------------------------------------------------ */
int kso_congrad(const kso_geom *g, const double *fat, const double *lng, const double *src,
                double *dest, kso_qic *qic, double mass) {
  const size_t V = g->V;
  int parity = qic->parity;
  int otherparity = parity == KSO_EVEN ? KSO_ODD : KSO_EVEN;
  int niter = qic->max, max_restarts = qic->nrestart;
  double rsqmin = qic->resid * qic->resid;
  double relrsqmin = qic->relresid * qic->relresid;
  int max_cg = max_restarts * niter;
  double msq_x4 = 4.0 * mass * mass;
  size_t lo, hi;
  parity_range(g, parity, &lo, &hi);

  double source_norm = kso_norm2(g, src, parity);
  int nrestart = 0, iteration = 0;
  qic->size_r = 0; qic->size_relr = 1.0;
  qic->final_iters = 0; qic->final_restart = 0; qic->converged = 1;
  qic->final_rsq = 0; qic->final_relrsq = 0;
  if (source_norm == 0.0) {
    for (size_t i = lo; i < hi; i++) memset(dest + 6 * i, 0, 6 * sizeof(double));
    return 0;
  }
  double *ttt = (double *)calloc(6 * V, sizeof(double));
  double *cg_p = (double *)calloc(6 * V, sizeof(double));
  double *resid = (double *)calloc(6 * V, sizeof(double));
  double rsq = 0, actual_rsq = 999., oldrsq, pkp, c_tr, c_tt, relrsq = 1.0;
  double a, b;
  for (;;) {
    if ((iteration % niter == 0) ||
        ((rsqmin <= 0 || rsqmin > qic->size_r) && (relrsqmin <= 0 || relrsqmin > qic->size_relr))) {
      rsq = 0.0;
      kso_dslash(g, fat, lng, dest, ttt, otherparity);
      kso_dslash(g, fat, lng, ttt, ttt, parity);
#pragma omp parallel for reduction(+ : rsq)
      for (size_t i = lo; i < hi; i++)
        for (int k = 0; k < 6; k++) {
          size_t j = 6 * i + k;
          ttt[j] = ttt[j] + (-msq_x4) * dest[j];
          resid[j] = src[j] + ttt[j];
          cg_p[j] = resid[j];
          rsq += resid[j] * resid[j];
        }
      actual_rsq = rsq;
      qic->final_rsq = rsq / source_norm;
      qic->final_relrsq = relrsq;
      iteration++;
      if (iteration >= max_cg || nrestart >= max_restarts ||
          ((rsqmin <= 0 || rsqmin > qic->final_rsq) && (relrsqmin <= 0 || relrsqmin > qic->final_relrsq)))
        break;
      nrestart++;
    }
    oldrsq = actual_rsq;
    kso_dslash(g, fat, lng, cg_p, ttt, otherparity);
    kso_dslash(g, fat, lng, ttt, ttt, parity);
    pkp = 0; c_tr = 0; c_tt = 0;
#pragma omp parallel for reduction(+ : pkp, c_tr, c_tt)
    for (size_t i = lo; i < hi; i++)
      for (int k = 0; k < 6; k++) {
        size_t j = 6 * i + k;
        ttt[j] = ttt[j] + (-msq_x4) * cg_p[j];
        pkp += cg_p[j] * ttt[j];
        c_tr += ttt[j] * resid[j];
        c_tt += ttt[j] * ttt[j];
      }
    iteration++;
    a = -rsq / pkp;
    actual_rsq = 0;
#pragma omp parallel for reduction(+ : actual_rsq)
    for (size_t i = lo; i < hi; i++)
      for (int k = 0; k < 6; k++) {
        size_t j = 6 * i + k;
        dest[j] = dest[j] + a * cg_p[j];
        resid[j] = resid[j] + a * ttt[j];
        actual_rsq += resid[j] * resid[j];
      }
    rsq = oldrsq + 2.0 * a * c_tr + a * a * c_tt;
    qic->size_r = rsq / source_norm;
    qic->size_relr = relrsq;
    qic->final_iters = iteration;
    qic->final_restart = nrestart;
    qic->converged = 1;
    b = rsq / oldrsq;
#pragma omp parallel for
    for (size_t i = lo; i < hi; i++)
      for (int k = 0; k < 6; k++) {
        size_t j = 6 * i + k;
        cg_p[j] = resid[j] + b * cg_p[j];
      }
  }
  qic->final_iters = iteration;
  qic->final_restart = nrestart;
  qic->converged = 1;
  if (nrestart == max_restarts || iteration == max_cg) qic->converged = 0;
  free(ttt); free(cg_p); free(resid);
  return iteration;
}

/* ------------------------------------------------------------------------ */
/* multi-shift CG (generic_ks/ks_multicg_offset.c:63-505), one parity          */
/* ------------------------------------------------------------------------ */
int kso_multicg_offset(const kso_geom *g, const double *fat, const double *lng, const double *src,
                       double **psim, const double *offsets, int num_offsets, kso_qic *qic) {
  const size_t V = g->V;
  int niter = qic[0].max * qic[0].nrestart;
  double rsqmin = qic[0].resid * qic[0].resid;
  int parity = qic[0].parity;
  int l_parity = parity == KSO_ODD ? KSO_ODD : KSO_EVEN;
  int l_other = l_parity == KSO_EVEN ? KSO_ODD : KSO_EVEN;
  for (int j = 0; j < num_offsets; j++) {
    qic[j].final_rsq = 0; qic[j].final_relrsq = 0; qic[j].size_r = 0; qic[j].size_relr = 0;
    qic[j].final_iters = 0; qic[j].final_restart = 0; qic[j].converged = 1;
  }
  if (num_offsets == 0) return 0;
  for (int j = 0; j < num_offsets; j++)
    if (offsets[j] <= 0) return -1;
  size_t lo, hi;
  parity_range(g, l_parity, &lo, &hi);

  double *shifts = (double *)malloc(sizeof(double) * num_offsets);
  double *zeta_i = (double *)malloc(sizeof(double) * num_offsets);
  double *zeta_im1 = (double *)malloc(sizeof(double) * num_offsets);
  double *zeta_ip1 = (double *)malloc(sizeof(double) * num_offsets);
  double *beta_i = (double *)malloc(sizeof(double) * num_offsets);
  double *beta_im1 = (double *)malloc(sizeof(double) * num_offsets);
  double *alpha = (double *)malloc(sizeof(double) * num_offsets);
  int *finished = (int *)calloc(num_offsets, sizeof(int));
  double **pm = (double **)malloc(sizeof(double *) * num_offsets);
  double offset_low = 1.0e20;
  int j_low = -1;
  for (int j = 0; j < num_offsets; j++) {
    shifts[j] = offsets[j];
    if (offsets[j] < offset_low) { offset_low = offsets[j]; j_low = j; }
  }
  for (int j = 0; j < num_offsets; j++) {
    pm[j] = (double *)calloc(6 * V, sizeof(double));
    if (j != j_low) shifts[j] -= shifts[j_low];
  }
  double shift0 = -shifts[j_low];
  double *ttt = (double *)calloc(6 * V, sizeof(double));
  double *cg_p = (double *)calloc(6 * V, sizeof(double));
  double *resid = (double *)calloc(6 * V, sizeof(double));

  int iteration = 0;
  int num_offsets_now = num_offsets;
  double source_norm = 0;
#pragma omp parallel for reduction(+ : source_norm)
  for (size_t i = lo; i < hi; i++)
    for (int k = 0; k < 6; k++) {
      size_t q = 6 * i + k;
      source_norm += src[q] * src[q];
      resid[q] = src[q];
      cg_p[q] = src[q];
      for (int j = 0; j < num_offsets; j++) { psim[j][q] = 0.0; pm[j][q] = src[q]; }
    }
  double rsq = source_norm, oldrsq, pkp, c1, c2;
  int ret = 0;
  if (source_norm == 0.0) goto done;
  iteration++;
  double rsqstop = rsqmin * source_norm;
  for (int j = 0; j < num_offsets; j++) {
    zeta_im1[j] = zeta_i[j] = 1.0;
    beta_im1[j] = -1.0;
    alpha[j] = 0.0;
  }
  do {
    oldrsq = rsq;
    kso_dslash(g, fat, lng, cg_p, ttt, l_other);
    kso_dslash(g, fat, lng, ttt, ttt, l_parity);
    pkp = 0.0;
#pragma omp parallel for reduction(+ : pkp)
    for (size_t i = lo; i < hi; i++)
      for (int k = 0; k < 6; k++) {
        size_t q = 6 * i + k;
        ttt[q] = ttt[q] + shift0 * cg_p[q];
        pkp += cg_p[q] * ttt[q];
      }
    iteration++;
    beta_i[j_low] = -rsq / pkp;
    zeta_ip1[j_low] = 1.0;
    for (int j = 0; j < num_offsets_now; j++) {
      if (j == j_low) continue;
      zeta_ip1[j] = zeta_i[j] * zeta_im1[j] * beta_im1[j_low];
      c1 = beta_i[j_low] * alpha[j_low] * (zeta_im1[j] - zeta_i[j]);
      c2 = zeta_im1[j] * beta_im1[j_low] * (1.0 + shifts[j] * beta_i[j_low]);
      if (c1 + c2 != 0.0) zeta_ip1[j] /= c1 + c2;
      else { zeta_ip1[j] = 0.0; finished[j] = 1; }
      if (zeta_i[j] != 0.0) beta_i[j] = beta_i[j_low] * zeta_ip1[j] / zeta_i[j];
      else {
        zeta_ip1[j] = 0.0; beta_i[j] = 0.0; finished[j] = 1;
        if (j == num_offsets_now - 1 && j > j_low) num_offsets_now--;
      }
    }
    rsq = 0.0;
#pragma omp parallel for reduction(+ : rsq)
    for (size_t i = lo; i < hi; i++)
      for (int k = 0; k < 6; k++) {
        size_t q = 6 * i + k;
        for (int j = 0; j < num_offsets_now; j++) psim[j][q] = psim[j][q] + beta_i[j] * pm[j][q];
        resid[q] = resid[q] + beta_i[j_low] * ttt[q];
        rsq += resid[q] * resid[q];
      }
    if (rsqstop > 0 && rsq <= rsqstop) {
      for (int j = 0; j < num_offsets; j++) {
        qic[j].final_rsq = rsq / source_norm;
        qic[j].size_r = qic[j].final_rsq;
        qic[j].size_relr = qic[j].final_relrsq;
        qic[j].final_iters = iteration;
      }
      goto done;
    }
    alpha[j_low] = rsq / oldrsq;
    for (int j = 0; j < num_offsets_now; j++) {
      if (j == j_low) continue;
      if (zeta_i[j] * beta_i[j_low] != 0.0)
        alpha[j] = alpha[j_low] * zeta_ip1[j] * beta_i[j] / (zeta_i[j] * beta_i[j_low]);
      else { alpha[j] = 0.0; finished[j] = 1; }
    }
#pragma omp parallel for
    for (size_t i = lo; i < hi; i++)
      for (int k = 0; k < 6; k++) {
        size_t q = 6 * i + k;
        for (int j = 0; j < num_offsets_now; j++) pm[j][q] = zeta_ip1[j] * resid[q] + alpha[j] * pm[j][q];
        cg_p[q] = pm[j_low][q];
      }
    for (int j = 0; j < num_offsets_now; j++) {
      beta_im1[j] = beta_i[j];
      zeta_im1[j] = zeta_i[j];
      zeta_i[j] = zeta_ip1[j];
    }
  } while (iteration < niter);
  for (int j = 0; j < num_offsets; j++) {
    qic[j].final_rsq = rsq / source_norm;
    qic[j].size_r = qic[j].final_rsq;
    qic[j].final_iters = iteration;
    qic[j].converged = 0;
  }
done:
  ret = iteration;
  for (int j = 0; j < num_offsets; j++) free(pm[j]);
  free(pm); free(shifts); free(zeta_i); free(zeta_im1); free(zeta_ip1);
  free(beta_i); free(beta_im1); free(alpha); free(finished);
  free(ttt); free(cg_p); free(resid);
  return ret;
}

/* ------------------------------------------------------------------------ */
/* Force level 0 (generic_ks/fermion_force_hisq_multi.c:1478-1636):          */
/*   oprod1[4x+mu] = sum_t r_t Z_t(x+mu)  Z_t(x)^dag                          */
/*   oprod3[4x+mu] = sum_t r_t Z_t(x+3mu) Z_t(x)^dag                          */
/* over all sites (EVENANDODD), Z_t = multi_x[t] on both parities.            */
/* ------------------------------------------------------------------------ */
void kso_force_oprod(const kso_geom *g, const double *const *multi_x, const double *residues,
                     int nterms, double *oprod1, double *oprod3) {
#pragma omp parallel for
  for (size_t i = 0; i < g->V; i++) {
    for (int mu = 0; mu < 4; mu++) {
      for (int hop = 0; hop < 2; hop++) {
        size_t y = NB(g, hop ? 8 + mu : mu, i);
        double acc[18];
        memset(acc, 0, sizeof acc);
        for (int t = 0; t < nterms; t++) {
          const double *a = multi_x[t] + 6 * y, *b = multi_x[t] + 6 * i;
          double r = residues[t];
          for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) {
              /* a_p * conj(b_q) (libraries/su3_proj.c) */
              double re = a[2 * p] * b[2 * q] + a[2 * p + 1] * b[2 * q + 1];
              double im = a[2 * p + 1] * b[2 * q] - a[2 * p] * b[2 * q + 1];
              acc[2 * (3 * p + q)] += r * re;
              acc[2 * (3 * p + q) + 1] += r * im;
            }
        }
        memcpy((hop ? oprod3 : oprod1) + (4 * i + mu) * 18, acc, sizeof acc);
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FN (fat + Naik) fermion force on the thin links                           */
/* (generic_ks/fermion_force_fn_multi.c:81-313, eo_fermion_force_multi)       */
/*                                                                           */
/* The fat link is F_mu(x) = sum_p c_p P_p(x), a sum over the action's paths  */
/* from x to x+mu (rotations and reflections of the basic paths of            */
/* generic_ks/imp_actions/asqtad_action.h), and the Naik link is              */
/* L_mu(x) = c_N U_mu(x) U_mu(x+mu) U_mu(x+2mu).  With                        */
/* dS = sum 2 Re Tr[dF_mu(x) eps(x) O1_mu(x) + dL_mu(x) eps(x) O3_mu(x)]        */
/* (O1/O3 = kso_force_oprod) every link of every path gets                   */
/* M_j = (L_j+1 ... L_n-1) eps(x) O (L_0 ... L_j-1): a forward link U         */
/* collects c_p M_j, a backward link U^dag collects c_p M_j^dag, into G.     */
/* The momentum is mom_mu(y) = TA(2 eps_md U_mu(y) G_mu(y)), TA = traceless   */
/* anti-Hermitian part (libraries/make_ahmat.c).                              */
/* Pinned by MILC's golden binary_samples/colormatrix.mom.sample.l8.          */
/* ------------------------------------------------------------------------ */
typedef struct { int n; int d[7]; double c; int naik; } kso_path;

static int add_path(kso_path *t, int k, int n, const int *d, double c, int naik) {
  t[k].n = n; t[k].c = c; t[k].naik = naik;
  for (int i = 0; i < n; i++) t[k].d[i] = d[i];
  return k + 1;
}
/* signed direction: s*nu with s=+1 -> nu, s=-1 -> OPP_DIR(nu) = 7-nu */
static int sdir(int nu, int s) { return s > 0 ? nu : 7 - nu; }

/* all paths of direction mu for coefficients {one_link, naik, 3-staple,
 * 5-staple, 7-staple, Lepage}; returns the count */
static int fn_paths(int mu, const double coeff[6], kso_path *t) {
  int k = 0, d[7];
  d[0] = mu; k = add_path(t, k, 1, d, coeff[0], 0);
  d[0] = d[1] = d[2] = mu; k = add_path(t, k, 3, d, coeff[1], 1);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    for (int s = -1; s <= 1; s += 2) {
      d[0] = sdir(nu, s); d[1] = mu; d[2] = sdir(nu, -s); k = add_path(t, k, 3, d, coeff[2], 0);
      d[0] = d[1] = sdir(nu, s); d[2] = mu; d[3] = d[4] = sdir(nu, -s); k = add_path(t, k, 5, d, coeff[5], 0);
      for (int rho = 0; rho < 4; rho++) {
        if (rho == mu || rho == nu) continue;
        for (int s2 = -1; s2 <= 1; s2 += 2) {
          d[0] = sdir(nu, s); d[1] = sdir(rho, s2); d[2] = mu; d[3] = sdir(rho, -s2); d[4] = sdir(nu, -s);
          k = add_path(t, k, 5, d, coeff[3], 0);
          for (int sig = 0; sig < 4; sig++) {
            if (sig == mu || sig == nu || sig == rho) continue;
            for (int s3 = -1; s3 <= 1; s3 += 2) {
              d[0] = sdir(nu, s); d[1] = sdir(rho, s2); d[2] = sdir(sig, s3); d[3] = mu;
              d[4] = sdir(sig, -s3); d[5] = sdir(rho, -s2); d[6] = sdir(nu, -s);
              k = add_path(t, k, 7, d, coeff[4], 0);
            }
          }
        }
      }
    }
  }
  return k;
}

int kso_fn_paths(int mu, const double coeff[6], int *dirs, double *c, int *len, int *naik) {
  kso_path t[128];
  int n = fn_paths(mu, coeff, t);
  for (int p = 0; p < n; p++) {
    len[p] = t[p].n; c[p] = t[p].c; naik[p] = t[p].naik;
    for (int i = 0; i < 7; i++) dirs[7 * p + i] = i < t[p].n ? t[p].d[i] : -1;
  }
  return n;
}

/* Adjoint of a path-table smearing X_mu(x) = sum_p c_p P_p(x): given the
 * force on X (non-Naik paths read F1, Naik paths F3; with_eps multiplies the
 * force at the path start x by eps(x)), add to G the force on the links L of
 * the paths, all in the convention dS = sum 2 Re Tr[dA_mu(x) F_mu(x)]. */
static void path_adjoint(const kso_geom *g, const double coeff[6], const double *L, const double *F1,
                         const double *F3, int with_eps, double *G) {
  const size_t V = g->V;
  for (int mu = 0; mu < 4; mu++) {
    kso_path t[128];
    int np = fn_paths(mu, coeff, t);
    for (int p = 0; p < np; p++) {
      const kso_path *P = &t[p];
      if (P->c == 0.0) continue;
      /* serial over sites: contributions scatter to neighbours */
      for (size_t x = 0; x < V; x++) {
        size_t y[8];
        double Lk[7][18], pre[8][18], suf[8][18], Ob[18];
        y[0] = x;
        for (int j = 0; j < P->n; j++) {
          int d = P->d[j];
          if (d < 4) { memcpy(Lk[j], L + (4 * y[j] + d) * 18, sizeof Lk[j]); y[j + 1] = NB(g, d, y[j]); }
          else { int dp = 7 - d; y[j + 1] = NB(g, d, y[j]); m_adj(L + (4 * y[j + 1] + dp) * 18, Lk[j]); }
        }
        /* prefix products pre[j] = L_0..L_j-1, suffix suf[j] = L_j+1..L_n-1 */
        memset(pre[0], 0, sizeof pre[0]);
        RE(pre[0], 0, 0) = RE(pre[0], 1, 1) = RE(pre[0], 2, 2) = 1.0;
        for (int j = 0; j < P->n; j++) m_nn(pre[j], Lk[j], pre[j + 1]);
        memset(suf[P->n - 1], 0, sizeof suf[0]);
        RE(suf[P->n - 1], 0, 0) = RE(suf[P->n - 1], 1, 1) = RE(suf[P->n - 1], 2, 2) = 1.0;
        for (int j = P->n - 2; j >= 0; j--) m_nn(Lk[j + 1], suf[j + 1], suf[j]);
        double ex = (with_eps && x >= g->Vh) ? -1.0 : 1.0;
        const double *O = (P->naik ? F3 : F1) + (4 * x + mu) * 18;
        for (int k = 0; k < 18; k++) Ob[k] = ex * O[k];
        for (int j = 0; j < P->n; j++) {
          double A[18], M[18];
          m_nn(suf[j], Ob, A);
          m_nn(A, pre[j], M);
          int d = P->d[j];
          if (d < 4) m_axpy(G + (4 * y[j] + d) * 18, P->c, M);
          else {
            double Md[18];
            m_adj(M, Md);
            m_axpy(G + (4 * y[j + 1] + (7 - d)) * 18, P->c, Md);
          }
        }
      }
    }
  }
}

/* The same adjoint in staple form: reverse mode through the optimised
 * fattening of kso_fat_links (X = (c1 - 6 c_lep) U + sum_nu [c3 S1 + c_lep S_nu(S1)
 * + sum_rho (c5 S2 + sum_sig c7 S_sig(S2))], S1 = S_nu(U_mu), S2 = S_rho(S1),
 * S_nu(A)(x) = U_nu(x) A(x+nu) U_nu(x+mu)^dag + U_nu(x-nu)^dag A(x-nu) U_nu(x-nu+mu))
 * plus the Naik product.  For unitary links it differs from path_adjoint only
 * by components that the traceless anti-Hermitian projection (and the U(3)
 * projection derivative) remove: the retracing Lepage terms are constant on
 * the group. */
static void staple_fwd(const kso_geom *g, int mu, int nu, const double *A, int strideA, const double *U,
                       double *out) {
#pragma omp parallel for
  for (size_t x = 0; x < g->V; x++) {
    double t1[18], up[18], lo[18];
    size_t xn = NB(g, nu, x), xm = NB(g, mu, x), xd = NB(g, 7 - nu, x), xdm = NB(g, mu, xd);
    m_na(A + (size_t)strideA * 18 * xn, U + (4 * xm + nu) * 18, t1);
    m_nn(U + (4 * x + nu) * 18, t1, up);
    m_an(U + (4 * xd + nu) * 18, A + (size_t)strideA * 18 * xd, t1);
    m_nn(t1, U + (4 * xdm + nu) * 18, lo);
    for (int k = 0; k < 18; k++) out[18 * x + k] = up[k] + lo[k];
  }
}

/* reverse of O = S_nu(A) (mu-staple) for the force Gs on O, scaled by cg:
 * FA += force on A (stride strideFA, may be NULL), FU[nu] += force on U_nu */
static void staple_adj(const kso_geom *g, int mu, int nu, const double *A, int strideA, const double *Gs,
                       double cg, const double *U, double *FA, int strideFA, double *FU) {
#pragma omp parallel for
  for (size_t y = 0; y < g->V; y++) {
    double t1[18], t2[18], acc[18];
    size_t ypn = NB(g, nu, y), ymn = NB(g, 7 - nu, y), ypm = NB(g, mu, y), ymm = NB(g, 7 - mu, y);
    size_t ymnpm = NB(g, mu, ymn), ymmpn = NB(g, nu, ymm);
    const double *Unu = U + nu * 18;
#define UN(s) (Unu + 4 * 18 * (s))
#define AA(s) (A + (size_t)strideA * 18 * (s))
#define GG(s) (Gs + 18 * (s))
    if (FA) {
      /* FA(y) += U_nu(y-nu+mu)^dag G(y-nu) U_nu(y-nu) + U_nu(y+mu) G(y+nu) U_nu(y)^dag */
      m_an(UN(ymnpm), GG(ymn), t1);
      m_nn(t1, UN(ymn), acc);
      m_nn(UN(ypm), GG(ypn), t1);
      m_na(t1, UN(y), t2);
      for (int k = 0; k < 18; k++) acc[k] += t2[k];
      m_axpy(FA + (size_t)strideFA * 18 * y, cg, acc);
    }
    /* FU_nu(y) += A(y+nu) U_nu(y+mu)^dag G(y) + [G(y-mu) U_nu(y-mu) A(y-mu+nu)]^dag
     *           + [A(y) U_nu(y+mu) G(y+nu)]^dag + G(y-mu+nu) U_nu(y-mu)^dag A(y-mu) */
    m_na(AA(ypn), UN(ypm), t1);
    m_nn(t1, GG(y), acc);
    m_nn(GG(ymm), UN(ymm), t1);
    m_nn(t1, AA(ymmpn), t2);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) { RE(acc, i, j) += RE(t2, j, i); IM(acc, i, j) -= IM(t2, j, i); }
    m_nn(AA(y), UN(ypm), t1);
    m_nn(t1, GG(ypn), t2);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) { RE(acc, i, j) += RE(t2, j, i); IM(acc, i, j) -= IM(t2, j, i); }
    m_na(GG(ymmpn), UN(ymm), t1);
    m_nn(t1, AA(ymm), t2);
    for (int k = 0; k < 18; k++) acc[k] += t2[k];
    m_axpy(FU + (4 * y + nu) * 18, cg, acc);
#undef UN
#undef AA
#undef GG
  }
}

static void smear_adjoint(const kso_geom *g, const double coeff[6], const double *L, const double *F1,
                          const double *F3, int with_eps, double *G) {
  const size_t V = g->V;
  double *F = (double *)malloc(sizeof(double) * 18 * V);
  double *S1 = (double *)malloc(sizeof(double) * 18 * V);
  double *S2 = (double *)malloc(sizeof(double) * 18 * V);
  double *G1 = (double *)malloc(sizeof(double) * 18 * V);
  double *G2 = (double *)malloc(sizeof(double) * 18 * V);
  const double c1 = coeff[0] - 6.0 * coeff[5], cN = coeff[1], c3 = coeff[2], c5 = coeff[3], c7 = coeff[4],
               cl = coeff[5];
  for (int mu = 0; mu < 4; mu++) {
    for (size_t x = 0; x < V; x++) {
      double e = (with_eps && x >= g->Vh) ? -1.0 : 1.0;
      for (int k = 0; k < 18; k++) F[18 * x + k] = e * F1[(4 * x + mu) * 18 + k];
      m_axpy(G + (4 * x + mu) * 18, c1, F + 18 * x);
    }
    if (c3 != 0.0 || cl != 0.0 || c5 != 0.0 || c7 != 0.0) {
      for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        staple_fwd(g, mu, nu, L + 18 * mu, 4, L, S1);
        for (size_t k = 0; k < 18 * V; k++) G1[k] = c3 * F[k];
        if (cl != 0.0) staple_adj(g, mu, nu, S1, 1, F, cl, L, G1, 1, G);
        if (c5 != 0.0 || c7 != 0.0) {
          for (int rho = 0; rho < 4; rho++) {
            if (rho == mu || rho == nu) continue;
            staple_fwd(g, mu, rho, S1, 1, L, S2);
            for (size_t k = 0; k < 18 * V; k++) G2[k] = c5 * F[k];
            if (c7 != 0.0)
              for (int sig = 0; sig < 4; sig++) {
                if (sig == mu || sig == nu || sig == rho) continue;
                staple_adj(g, mu, sig, S2, 1, F, c7, L, G2, 1, G);
              }
            staple_adj(g, mu, rho, S1, 1, G2, 1.0, L, G1, 1, G);
          }
        }
        staple_adj(g, mu, nu, L + 18 * mu, 4, G1, 1.0, L, G + 18 * mu, 4, G);
      }
    }
    if (cN != 0.0) {
      /* L_mu(x) = U(x) U(x+mu) U(x+2mu): gather the three positions */
      for (size_t x = 0; x < V; x++) {
        double e = (with_eps && x >= g->Vh) ? -1.0 : 1.0;
        for (int k = 0; k < 18; k++) F[18 * x + k] = e * F3[(4 * x + mu) * 18 + k];
      }
#pragma omp parallel for
      for (size_t y = 0; y < V; y++) {
        double t1[18], t2[18], acc[18];
        size_t p1 = NB(g, mu, y), p2 = NB(g, mu, p1), m1 = NB(g, 7 - mu, y), m2 = NB(g, 7 - mu, m1);
        const double *Um = L + mu * 18;
        m_nn(Um + 72 * p1, Um + 72 * p2, t1);
        m_nn(t1, F + 18 * y, acc);
        m_nn(Um + 72 * p1, F + 18 * m1, t1);
        m_nn(t1, Um + 72 * m1, t2);
        for (int k = 0; k < 18; k++) acc[k] += t2[k];
        m_nn(F + 18 * m2, Um + 72 * m2, t1);
        m_nn(t1, Um + 72 * m1, t2);
        for (int k = 0; k < 18; k++) acc[k] += t2[k];
        m_axpy(G + (4 * y + mu) * 18, cN, acc);
      }
    }
  }
  free(F); free(S1); free(S2); free(G1); free(G2);
}

static int g_force_form = 0; /* 0: path table (gather), 1: staple form */
void kso_set_force_form(int form) { g_force_form = form; }

static void force_adjoint(const kso_geom *g, const double coeff[6], const double *L, const double *F1,
                          const double *F3, int with_eps, double *G) {
  if (g_force_form) smear_adjoint(g, coeff, L, F1, F3, with_eps, G);
  else path_adjoint(g, coeff, L, F1, F3, with_eps, G);
}

/* mom = TA(scale U G) per link, full anti-Hermitian matrices (libraries/make_ahmat.c) */
static void ta_force(const kso_geom *g, const double *U, const double *G, double scale, double *mom) {
#pragma omp parallel for
  for (size_t i = 0; i < 4 * g->V; i++) {
    double UG[18], m[18];
    m_nn(U + 18 * i, G + 18 * i, UG);
    for (int k = 0; k < 18; k++) UG[k] *= scale;
    double tr = (IM(UG, 0, 0) + IM(UG, 1, 1) + IM(UG, 2, 2)) / 3.0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        RE(m, a, b) = 0.5 * (RE(UG, a, b) - RE(UG, b, a));
        IM(m, a, b) = 0.5 * (IM(UG, a, b) + IM(UG, b, a));
      }
    for (int a = 0; a < 3; a++) IM(m, a, a) -= tr;
    memcpy(mom + 18 * i, m, sizeof m);
  }
}

void kso_fn_force(const kso_geom *g, const double coeff[6], const double *U, const double *const *multi_x,
                  const double *residues, int nterms, double eps, double *mom) {
  const size_t V = g->V;
  double *o1 = (double *)malloc(sizeof(double) * 18 * 4 * V);
  double *o3 = (double *)malloc(sizeof(double) * 18 * 4 * V);
  double *G = (double *)calloc(18 * 4 * V, sizeof(double));
  kso_force_oprod(g, multi_x, residues, nterms, o1, o3);
  force_adjoint(g, coeff, U, o1, o3, 1, G);
  ta_force(g, U, G, 2.0 * eps, mom);
  free(o1); free(o3); free(G);
}

/* ------------------------------------------------------------------------ */
/* Derivative of the U(3) projection W = V (V^dag V)^-1/2 (the role of        */
/* u3_unit_der_analytic, generic_ks/su3_mat_op.c:1379, and                    */
/* fn_fermion_force_multi_hisq_reunit, fermion_force_hisq_multi.c:1877-1998). */
/* With M = V^dag V = E diag(l) E^dag, Q = M^-1/2 and                          */
/* dQ = E [ f(l_i,l_j) (E^dag dM E)_ij ] E^dag,                               */
/* f(a,b) = (a^-1/2 - b^-1/2)/(a - b) = -1/(sqrt a sqrt b (sqrt a + sqrt b)),  */
/* a force F_W (dS = 2 Re Tr[dW F_W]) becomes                                  */
/*   F_V = Q F_W + L(K + K^dag) V^dag,  K = F_W V,  L(H) = E [f_ij H'_ij] E^dag */
/* MILC's force filter: when min l < 5e-5 all l are raised by 5e-5            */
/* (HISQ_FORCE_FILTER of ks_imp_rhmc/Make_template).  E from cyclic complex  */
/* Jacobi rotations.                                                          */
/* ------------------------------------------------------------------------ */
#include <complex.h>

static void jacobi_herm3(double complex a[3][3], double complex e[3][3], double l[3]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) e[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = cabs(a[0][1]) + cabs(a[0][2]) + cabs(a[1][2]);
    double diag = fabs(creal(a[0][0])) + fabs(creal(a[1][1])) + fabs(creal(a[2][2]));
    if (off <= 1e-18 * diag) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        double r = cabs(a[p][q]);
        if (r == 0.0) continue;
        double complex ph = a[p][q] / r;  /* e^{i phi} */
        double tau = (creal(a[q][q]) - creal(a[p][p])) / (2 * r);
        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
        double c = 1 / sqrt(1 + t * t), s = t * c;
        /* J = diag(1, e^{-i phi}) [[c, s], [-s, c]] in the (p, q) plane */
        double complex Jpp = c, Jpq = s, Jqp = -s * conj(ph), Jqq = c * conj(ph);
        /* a <- a J */
        for (int k = 0; k < 3; k++) {
          double complex akp = a[k][p], akq = a[k][q];
          a[k][p] = akp * Jpp + akq * Jqp;
          a[k][q] = akp * Jpq + akq * Jqq;
        }
        /* a <- J^dag a */
        for (int k = 0; k < 3; k++) {
          double complex apk = a[p][k], aqk = a[q][k];
          a[p][k] = conj(Jpp) * apk + conj(Jqp) * aqk;
          a[q][k] = conj(Jpq) * apk + conj(Jqq) * aqk;
        }
        /* e <- e J */
        for (int k = 0; k < 3; k++) {
          double complex ekp = e[k][p], ekq = e[k][q];
          e[k][p] = ekp * Jpp + ekq * Jqp;
          e[k][q] = ekp * Jpq + ekq * Jqq;
        }
      }
  }
  for (int i = 0; i < 3; i++) l[i] = creal(a[i][i]);
}

void kso_u3_adjoint(const double *Vm, const double *FW, double *FV) {
  double complex v[3][3], f[3][3], m[3][3], e[3][3], q[3][3], k[3][3], h[3][3], t[3][3], lm[3][3];
  double l[3], sq[3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      v[i][j] = RE(Vm, i, j) + I * IM(Vm, i, j);
      f[i][j] = RE(FW, i, j) + I * IM(FW, i, j);
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      m[i][j] = 0;
      for (int c = 0; c < 3; c++) m[i][j] += conj(v[c][i]) * v[c][j];
    }
  jacobi_herm3(m, e, l);
  double lmin = fmin(l[0], fmin(l[1], l[2]));
  if (lmin < 5.0e-5)
    for (int i = 0; i < 3; i++) l[i] += 5.0e-5;
  for (int i = 0; i < 3; i++) sq[i] = sqrt(l[i]);
  /* Q = E diag(1/sqrt l) E^dag */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      q[i][j] = 0;
      for (int c = 0; c < 3; c++) q[i][j] += e[i][c] * (1.0 / sq[c]) * conj(e[j][c]);
    }
  /* K = F V, H = K + K^dag */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      k[i][j] = 0;
      for (int c = 0; c < 3; c++) k[i][j] += f[i][c] * v[c][j];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) h[i][j] = k[i][j] + conj(k[j][i]);
  /* H' = E^dag H E, scaled by f(l_i, l_j), back: L = E H' E^dag */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      t[i][j] = 0;
      for (int c = 0; c < 3; c++) t[i][j] += h[i][c] * e[c][j];
    }
  double complex hp[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double complex s2 = 0;
      for (int c = 0; c < 3; c++) s2 += conj(e[c][i]) * t[c][j];
      hp[i][j] = s2 * (-1.0 / (sq[i] * sq[j] * (sq[i] + sq[j])));
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      t[i][j] = 0;
      for (int c = 0; c < 3; c++) t[i][j] += e[i][c] * hp[c][j];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      lm[i][j] = 0;
      for (int c = 0; c < 3; c++) lm[i][j] += t[i][c] * conj(e[j][c]);
    }
  /* F_V = Q F + L V^dag */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double complex s2 = 0;
      for (int c = 0; c < 3; c++) s2 += q[i][c] * f[c][j] + lm[i][c] * conj(v[j][c]);
      RE(FV, i, j) = creal(s2);
      IM(FV, i, j) = cimag(s2);
    }
}

/* HISQ thin-link force (fn_fermion_force_multi_hisq_wrapper_mx_cpu,
 * generic_ks/fermion_force_hisq_multi.c:1182-1475, one Naik set):
 * O1/O3 (level 0) -> level-2 path adjoint on W -> U(3) projection adjoint at
 * V -> level-1 (fat7) path adjoint on U -> mom = TA(2 eps U F_U).
 * U has the KS phases in; V and W are rebuilt here as kso_hisq_links does. */
void kso_hisq_force(const kso_geom *g, const double *U, const double *const *multi_x, const double *residues,
                    int nterms, double eps, double *mom) {
  kso_hisq_force_naik(g, U, multi_x, residues, nterms, 0, 0.0, eps, mom);
}

/* With a second Naik set (fn_fermion_force_multi_hisq_wrapper_mx,
 * generic_ks/fermion_force_hisq_multi.c:1284-1374): every term goes through
 * the level-2 adjoint (ap->p2); the last n_naik terms, whose links carry
 * eps_naik times the "1-link + Naik" set (ap->p3: 1/8, -1/24,
 * imp_actions/hisq/hisq_u3_action.h:86-101), also through that table's
 * adjoint, scaled by eps_naik, before the U(3) derivative. */
void kso_hisq_force_naik(const kso_geom *g, const double *U, const double *const *multi_x, const double *residues,
                         int nterms, int n_naik, double eps_naik, double eps, double *mom) {
  const size_t n = 4 * g->V * 18;
  const double c3[6] = {1.0 / 8.0, -1.0 / 24.0, 0.0, 0.0, 0.0, 0.0};
  const double c1[6] = {1.0 / 8.0, 0.0, -1.0 / 8.0 * 0.5, 1.0 / 8.0 * 0.25 * 0.5,
                        -1.0 / 8.0 * 0.125 * (1.0 / 6.0), 0.0};
  const double c2[6] = {(1.0 / 8.0) + (2.0 * 6.0 / 16.0) + (1.0 / 8.0), -1.0 / 24.0,
                        -1.0 / 8.0 * 0.5, 1.0 / 8.0 * 0.25 * 0.5,
                        -1.0 / 8.0 * 0.125 * (1.0 / 6.0), -2.0 / 16.0};
  double *Vl = (double *)malloc(sizeof(double) * n);
  double *W = (double *)malloc(sizeof(double) * n);
  double *o1 = (double *)malloc(sizeof(double) * n);
  double *o3 = (double *)malloc(sizeof(double) * n);
  double *FW = (double *)calloc(n, sizeof(double));
  double *FV = (double *)malloc(sizeof(double) * n);
  double *FU = (double *)calloc(n, sizeof(double));
  kso_fat_links(g, c1, U, Vl);
#pragma omp parallel for
  for (size_t l = 0; l < 4 * g->V; l++) kso_u3_unitarize(Vl + 18 * l, W + 18 * l);
  kso_force_oprod(g, multi_x, residues, nterms, o1, o3);
  force_adjoint(g, c2, W, o1, o3, 1, FW);
  if (n_naik > 0) {
    double *FE = (double *)calloc(n, sizeof(double));
    kso_force_oprod(g, multi_x + (nterms - n_naik), residues + (nterms - n_naik), n_naik, o1, o3);
    force_adjoint(g, c3, W, o1, o3, 1, FE);
    for (size_t k = 0; k < n; k++) FW[k] += eps_naik * FE[k];
    free(FE);
  }
#pragma omp parallel for
  for (size_t l = 0; l < 4 * g->V; l++) kso_u3_adjoint(Vl + 18 * l, FW + 18 * l, FV + 18 * l);
  force_adjoint(g, c1, U, FV, FV, 0, FU);
  ta_force(g, U, FU, 2.0 * eps, mom);
  free(Vl); free(W); free(o1); free(o3); free(FW); free(FV); free(FU);
}

/* ------------------------------------------------------------------------ */
/* Symanzik-improved gauge action (1x1 + 1x2 + 1x1x1), its force on the      */
/* momenta, the momentum action and the link update.                         */
/* ------------------------------------------------------------------------ */

/* Loop coefficients of generic/imp_actions/symanzik_1loop_hisq_action.h
 * (hisq = 1, HISQ quark loops, nf = total dynamical flavours) and
 * symanzik_1loop_action.h (hisq = 0). */
void kso_gauge_coeffs(double u0, int nf, int hisq, double coeff[3]) {
  double l = log(u0);
  coeff[0] = 1.0;
  if (hisq) {
    coeff[1] = -1.0 / (20.0 * u0 * u0) * (1.0 - (0.6264 - 1.1746 * nf) * l);
    coeff[2] = 1.0 / (u0 * u0) * (0.0433 - 0.0156 * nf) * l;
  } else {
    coeff[1] = -1.0 / (20.0 * u0 * u0) * (1.0 - 0.6264 * l);
    coeff[2] = 1.0 / (u0 * u0) * 0.04335 * l;
  }
}

typedef struct { int n, shape; int d[6]; } kso_loop;

/* smallest rotation of d[0..n-1] as a base-8 number */
static long loop_key1(const int *d, int n) {
  long best = -1;
  for (int r = 0; r < n; r++) {
    long k = 0;
    for (int j = 0; j < n; j++) k = 8 * k + d[(r + j) % n];
    if (best < 0 || k < best) best = k;
  }
  return best;
}

/* Every rotation and reflection of the three loop shapes, one loop per class
 * of cyclic shifts and reversal (generic/gauge_stuff.c:68-181): 6 plaquettes,
 * 12 rectangles, 16 parallelograms.  Directions in MILC's codes. */
int kso_gauge_loops(int *dirs, int *len, int *shape) {
  static const int shapes[3][6] = {{0, 1, 7, 6, 0, 0}, {0, 0, 1, 7, 7, 6}, {0, 1, 2, 7, 6, 5}};
  static const int slen[3] = {4, 6, 6};
  kso_loop out[64];
  long keys[64];
  int count = 0;
  int perm[4];
  for (perm[0] = 0; perm[0] < 4; perm[0]++)
    for (perm[1] = 0; perm[1] < 4; perm[1]++)
      for (perm[2] = 0; perm[2] < 4; perm[2]++)
        for (perm[3] = 0; perm[3] < 4; perm[3]++) {
          if (perm[0] == perm[1] || perm[0] == perm[2] || perm[0] == perm[3] || perm[1] == perm[2] ||
              perm[1] == perm[3] || perm[2] == perm[3])
            continue;
          for (int refl = 0; refl < 16; refl++) {
            int map[8];
            for (int a = 0; a < 4; a++) {
              map[a] = ((refl >> a) & 1) ? 7 - perm[a] : perm[a];
              map[7 - a] = 7 - map[a];
            }
            for (int s = 0; s < 3; s++) {
              int d[6], rv[6], n = slen[s];
              for (int j = 0; j < n; j++) d[j] = map[shapes[s][j]];
              for (int j = 0; j < n; j++) rv[j] = 7 - d[n - 1 - j];
              long k1 = loop_key1(d, n), k2 = loop_key1(rv, n), key = (k1 < k2 ? k1 : k2) * 4 + s;
              int seen = 0;
              for (int q = 0; q < count; q++) seen |= keys[q] == key;
              if (seen) continue;
              keys[count] = key;
              out[count].n = n;
              out[count].shape = s;
              memcpy(out[count].d, d, sizeof d);
              count++;
            }
          }
        }
  for (int q = 0; q < count; q++) {
    if (len) len[q] = out[q].n;
    if (shape) shape[q] = out[q].shape;
    if (dirs)
      for (int j = 0; j < 6; j++) dirs[6 * q + j] = j < out[q].n ? out[q].d[j] : -1;
  }
  return count;
}

/* sum over loops and base sites of coeff[shape] (3 - Re Tr P) (imp_gauge_action,
 * generic/gauge_stuff.c:248-322); the gauge action is beta/3 times this
 * (ks_imp_rhmc/d_action_rhmc.c:27) */
double kso_gauge_action(const kso_geom *g, const double coeff[3], const double *U) {
  int dirs[64 * 6], len[64], shp[64];
  int nl = kso_gauge_loops(dirs, len, shp);
  double total = 0.0;
  for (int q = 0; q < nl; q++) {
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (size_t x = 0; x < g->V; x++) {
      double P[18], T[18], L[18];
      memset(P, 0, sizeof P);
      RE(P, 0, 0) = RE(P, 1, 1) = RE(P, 2, 2) = 1.0;
      size_t y = x;
      for (int j = 0; j < len[q]; j++) {
        int d = dirs[6 * q + j];
        size_t yn = NB(g, d, y);
        if (d < 4) memcpy(L, U + (4 * y + d) * 18, sizeof L);
        else m_adj(U + (4 * yn + (7 - d)) * 18, L);
        m_nn(P, L, T);
        memcpy(P, T, sizeof P);
        y = yn;
      }
      sum += coeff[shp[q]] * (3.0 - (RE(P, 0, 0) + RE(P, 1, 1) + RE(P, 2, 2)));
    }
    total += sum;
  }
  return total;
}

/* mom[4V][18] (full anti-Hermitian matrices) += -eb3 TA(U G), G = sum over loop
 * occurrences of the link of coeff (rest of the loop, in the loop's order from
 * the link's end), adjoint for a link run backwards -- the update of
 * imp_gauge_force_cpu (generic/gauge_force_imp.c:37-219), eb3 = eps beta / 3 */
void kso_gauge_force(const kso_geom *g, const double coeff[3], double eb3, const double *U, double *mom) {
  int dirs[64 * 6], len[64], shp[64];
  int nl = kso_gauge_loops(dirs, len, shp);
  const size_t V = g->V;
  double *G = (double *)calloc(18 * 4 * V, sizeof(double));
  double *F = (double *)malloc(sizeof(double) * 18 * 4 * V);
  for (int q = 0; q < nl; q++) {
    const int n = len[q];
    for (size_t x = 0; x < V; x++) {
      size_t y[7];
      double Lk[6][18], pre[7][18], suf[6][18];
      y[0] = x;
      for (int j = 0; j < n; j++) {
        int d = dirs[6 * q + j];
        y[j + 1] = NB(g, d, y[j]);
        if (d < 4) memcpy(Lk[j], U + (4 * y[j] + d) * 18, sizeof Lk[j]);
        else m_adj(U + (4 * y[j + 1] + (7 - d)) * 18, Lk[j]);
      }
      memset(pre[0], 0, sizeof pre[0]);
      RE(pre[0], 0, 0) = RE(pre[0], 1, 1) = RE(pre[0], 2, 2) = 1.0;
      for (int j = 0; j < n; j++) m_nn(pre[j], Lk[j], pre[j + 1]);
      memset(suf[n - 1], 0, sizeof suf[0]);
      RE(suf[n - 1], 0, 0) = RE(suf[n - 1], 1, 1) = RE(suf[n - 1], 2, 2) = 1.0;
      for (int j = n - 2; j >= 0; j--) m_nn(Lk[j + 1], suf[j + 1], suf[j]);
      for (int j = 0; j < n; j++) {
        double M[18];
        m_nn(suf[j], pre[j], M);
        int d = dirs[6 * q + j];
        if (d < 4) m_axpy(G + (4 * y[j] + d) * 18, coeff[shp[q]], M);
        else {
          double Md[18];
          m_adj(M, Md);
          m_axpy(G + (4 * y[j + 1] + (7 - d)) * 18, coeff[shp[q]], Md);
        }
      }
    }
  }
  ta_force(g, U, G, -eb3, F);
  for (size_t i = 0; i < 18 * 4 * V; i++) mom[i] += F[i];
  free(G);
  free(F);
}

/* sum over links of |P|^2 / 2 - 4 (hmom_action, ks_imp_rhmc/d_action_rhmc.c:86-101),
 * P as full anti-Hermitian matrices */
double kso_mom_action(const kso_geom *g, const double *mom) {
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
  for (size_t i = 0; i < 4 * g->V; i++) {
    double s = 0.0;
    for (int k = 0; k < 18; k++) s += mom[18 * i + k] * mom[18 * i + k];
    sum += 0.5 * s - 4.0;
  }
  return sum;
}

/* U <- sum_{k=0}^{8} (eps P)^k / k! U, Horner form (update_u,
 * ks_imp_rhmc/update_u.c:66-111) */
void kso_update_u(const kso_geom *g, double eps, double *U, const double *mom) {
#pragma omp parallel for
  for (size_t i = 0; i < 4 * g->V; i++) {
    const double *P = mom + 18 * i;
    double *L = U + 18 * i, t1[18], t2[18];
    m_nn(P, L, t1);
    for (int k = 0; k < 18; k++) t2[k] = L[k] + (eps / 8.0) * t1[k];
    for (int m = 7; m >= 1; m--) {
      m_nn(P, t2, t1);
      for (int k = 0; k < 18; k++) t2[k] = L[k] + (eps / m) * t1[k];
    }
    memcpy(L, t2, sizeof t2);
  }
}

/* ------------------------------------------------------------------------- */
/* MILC's per-site random numbers (SITERAND builds), so that a trajectory can
 * be reproduced draw for draw.
 *
 * Generator: generic/ranstuff.c:66-115 (the default variant, 64-bit
 * congruential part).  State per site, 9 words: r[0..6] (24-bit values of a
 * feedback shift register), mult, ic.  Seeding from (seed, index): a 32-bit
 * linear congruential sequence s <- (69607 + 8 index) s + 12345 (int
 * arithmetic, wrapping) gives r[k] = bits 8..31 of the k-th value, and the 8th
 * value, sign-extended to 64 bits, starts ic; mult = 100000005 + 8 index.
 * Draw: t = ((r5 >> 7 | r6 << 17) ^ (r4 >> 1 | r5 << 23)) & 0xffffff, the
 * register shifts (r0 <- t), ic <- ic mult + 12345 (mod 2^64), and the value
 * is (t ^ bits 40..63 of ic) / 2^24.  index = the site's lexicographic index
 * (generic/make_lattice.c:49,55).
 * Gaussian: polar Box-Muller with weight exp(-x^2) (<x^2> = 1/2), the complex
 * draw being (v2 fac, v1 fac) (libraries/gaussrand.c, default variant);
 * momenta: random_anti_hermitian (libraries/rand_ahmat.c, default variant).
 * single = 1 rounds as a MILC_PRECISION=1 build (Real = float) does: the
 * rejection radius, fac and the products in float, fac from a double sqrt/log.
 * Pinned by the GRSOURCE sums and the starting momentum action MILC printed in
 * ks_imp_rhmc/test/su3_rhmc_hisq.1.sample-out (tests/test_cpu_rng.py). */
static void prn_init(uint64_t *st, int seed, int index) {
  uint32_t a = 69607u + 8u * (uint32_t)index, s = (uint32_t)seed;
  for (int k = 0; k < 7; k++) {
    s = a * s + 12345u;
    st[k] = (s >> 8) & 0xffffffu;
  }
  s = a * s + 12345u;
  st[7] = (uint64_t)(100000005u + 8u * (uint32_t)index);
  st[8] = (uint64_t)(int64_t)(int32_t)s;
}

static uint32_t prn_next(uint64_t *st) {
  uint64_t t = (((st[5] >> 7) | (st[6] << 17)) ^ ((st[4] >> 1) | (st[5] << 23))) & 0xffffffu;
  for (int k = 6; k > 0; k--) st[k] = st[k - 1];
  st[0] = t;
  st[8] = st[8] * st[7] + 12345u;
  return (uint32_t)(t ^ ((st[8] >> 40) & 0xffffffu));
}

static void prn_cgauss(uint64_t *st, int single, double *re, double *im) {
  if (single) {
    volatile float v1, v2, a, b, r;
    do {
      v1 = (float)(2.0 * (prn_next(st) * (1.0 / 16777216.0)) - 1.0);
      v2 = (float)(2.0 * (prn_next(st) * (1.0 / 16777216.0)) - 1.0);
      a = v1 * v1;
      b = v2 * v2;
      r = a + b;
    } while (r >= 1.0f);
    volatile float fac = (float)sqrt(-log((double)r) / (double)r);
    volatile float x = v2 * fac, y = v1 * fac;
    *re = x;
    *im = y;
  } else {
    double v1, v2, r;
    do {
      v1 = 2.0 * (prn_next(st) * (1.0 / 16777216.0)) - 1.0;
      v2 = 2.0 * (prn_next(st) * (1.0 / 16777216.0)) - 1.0;
      r = v1 * v1 + v2 * v2;
    } while (r >= 1.0);
    double fac = sqrt(-log(r) / r);
    *re = v2 * fac;
    *im = v1 * fac;
  }
}

void kso_prn_seed(const kso_geom *g, int iseed, uint64_t *st) {
  for (size_t i = 0; i < g->V; i++) {
    const int *c = g->coords + 4 * i;
    int index = c[0] + g->dims[0] * (c[1] + g->dims[1] * (c[2] + g->dims[2] * c[3]));
    prn_init(st + 9 * i, iseed, index);
  }
}

void kso_prn_seed_one(int seed, int index, uint64_t *st) { prn_init(st, seed, index); }
uint32_t kso_prn_next(uint64_t *st) { return prn_next(st); }

/* ranmom(): every site, directions XUP..TUP in turn; packed anti_hermitmat
 * [V][4][10] = m01, m02, m12 (re, im), m00im, m11im, m22im, 0 */
void kso_ranmom_prn(const kso_geom *g, uint64_t *st, int single, double *mom) {
  for (size_t i = 0; i < g->V; i++)
    for (int dir = 0; dir < 4; dir++) {
      double *m = mom + 10 * (4 * i + dir), a, b;
      prn_cgauss(st + 9 * i, single, &a, &b);
      if (single) {
        volatile float s3 = (float)sqrt(1.0 / 3.0), fa = (float)a, fb = (float)b;
        volatile float p = s3 * fb, d0 = fa + p, d1 = -fa + p;
        m[6] = d0;
        m[7] = d1;
        m[8] = (float)(-2.0 * (double)s3 * (double)fb);
      } else {
        double s3 = sqrt(1.0 / 3.0);
        m[6] = a + s3 * b;
        m[7] = -a + s3 * b;
        m[8] = -2.0 * s3 * b;
      }
      for (int k = 0; k < 3; k++) prn_cgauss(st + 9 * i, single, &m[2 * k], &m[2 * k + 1]);
      m[9] = 0.0;
    }
}

/* three complex Gaussians on each site of `parity` (grsource_imp_rhmc,
 * ks_imp_rhmc/grsource_rhmc.c:30-37); other sites zeroed; returns the sum of
 * |v|^2 over the drawn sites (MILC prints it as GRSOURCE: sum) */
double kso_gaussian_prn(const kso_geom *g, uint64_t *st, int single, int parity, double *v) {
  double sum = 0.0;
  memset(v, 0, 6 * g->V * sizeof(double));
  size_t lo = parity == KSO_ODD ? g->Vh : 0, hi = parity == KSO_EVEN ? g->Vh : g->V;
  for (size_t i = lo; i < hi; i++)
    for (int j = 0; j < 3; j++) {
      prn_cgauss(st + 9 * i, single, &v[6 * i + 2 * j], &v[6 * i + 2 * j + 1]);
      sum += v[6 * i + 2 * j] * v[6 * i + 2 * j] + v[6 * i + 2 * j + 1] * v[6 * i + 2 * j + 1];
    }
  return sum;
}

/* reunit_su3 (generic/reunitarize2.c:133-205) on one matrix: normalise row 0,
 * take row 0's component out of row 1 and normalise it, row 2 = conj(row0 x
 * row1).  Returns the number of the four deviations (|row0|^2 - 1, |<row0,
 * row1>|^2, |row1'|^2 - 1, |row2 change|^2) at or above tol; *maxdev the
 * largest.  Pinned by tests/test_cpu_gauge.py: numpy's QR of U^T gives the
 * same rows 0 and 1, det = 1, unitarity, SU(3) inputs unchanged. */
int kso_reunit_su3(double *M, double tol, double *maxdev) {
  double complex r0[3], r1[3], r2[3], old[3];
  for (int k = 0; k < 3; k++) {
    r0[k] = M[2 * k] + I * M[2 * k + 1];
    r1[k] = M[6 + 2 * k] + I * M[6 + 2 * k + 1];
    old[k] = M[12 + 2 * k] + I * M[12 + 2 * k + 1];
  }
  double dev[4], n0 = 0, n1 = 0;
  for (int k = 0; k < 3; k++) n0 += creal(r0[k] * conj(r0[k]));
  dev[0] = fabs(n0 - 1.0);
  for (int k = 0; k < 3; k++) r0[k] /= sqrt(n0);
  double complex ov = 0;
  for (int k = 0; k < 3; k++) ov += conj(r0[k]) * r1[k];
  dev[1] = creal(ov * conj(ov));
  for (int k = 0; k < 3; k++) r1[k] -= ov * r0[k];
  for (int k = 0; k < 3; k++) n1 += creal(r1[k] * conj(r1[k]));
  dev[2] = fabs(n1 - 1.0);
  for (int k = 0; k < 3; k++) r1[k] /= sqrt(n1);
  r2[0] = conj(r0[1] * r1[2] - r0[2] * r1[1]);
  r2[1] = conj(r0[2] * r1[0] - r0[0] * r1[2]);
  r2[2] = conj(r0[0] * r1[1] - r0[1] * r1[0]);
  dev[3] = 0;
  for (int k = 0; k < 3; k++) dev[3] += creal((old[k] - r2[k]) * conj(old[k] - r2[k]));
  int errors = 0;
  for (int k = 0; k < 4; k++) {
    if (dev[k] > *maxdev) *maxdev = dev[k];
    errors += dev[k] >= tol;
  }
  double complex *rows[3] = {r0, r1, r2};
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) {
      M[6 * r + 2 * k] = creal(rows[r][k]);
      M[6 * r + 2 * k + 1] = cimag(rows[r][k]);
    }
  return errors;
}

/* reunitarize every link (links without KS phases); returns the error count */
int kso_reunitarize(const kso_geom *g, double tol, double *U, double *maxdev) {
  int errors = 0;
  *maxdev = 0.0;
  for (size_t i = 0; i < 4 * g->V; i++) errors += kso_reunit_su3(U + 18 * i, tol, maxdev);
  return errors;
}
