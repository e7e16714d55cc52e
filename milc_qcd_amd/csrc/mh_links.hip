// mh_links.hip -- gauge-field generation and HISQ link construction on the device.
//
// Restates (MILC 7.8.1):
//   KS phases           generic_ks/rephase.c:64-143
//   fat links           generic_ks/fermion_links_fn_load_milc.c:134-266 (ASQ_OPTIMIZED_FATTENING)
//   generalised staple  generic/general_staple.c:42-124
//   Naik links          generic_ks/fermion_links_fn_load_milc.c:51-131
//   U(3) projection     generic_ks/su3_mat_op.c:828-1211 (u3_unitarize_analytic)
//   HISQ chain          generic_ks/fermion_links_hisq_load_milc.c:546-702,
//                       coefficients generic_ks/imp_actions/hisq/hisq_u3_action.h
// All arithmetic is FP64.  Link fields use the device LinkField layout; the
// intermediate "mu-link" fields of the staple recursion are MatFields.
#include <cmath>
#include <vector>

#include "mh_internal.h"

namespace mh {

// One matrix per site, both parities: element e of site (pb, cb) at
// data[(pb*9 + e)*stride + cb].
struct MatField {
  double2* data;
  int stride;
};

__device__ __forceinline__ void ld(const double2* base, int stride, int cb, Mat3<double>& m) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    double2 a = base[(size_t)e * stride + cb];
    m.e[e / 3][e % 3].re = a.x;
    m.e[e / 3][e % 3].im = a.y;
  }
}
__device__ __forceinline__ void st(double2* base, int stride, int cb, const Mat3<double>& m) {
#pragma unroll
  for (int e = 0; e < 9; e++) base[(size_t)e * stride + cb] = make_double2(m.e[e / 3][e % 3].re, m.e[e / 3][e % 3].im);
}
__device__ __forceinline__ const double2* lptr(const LinkField& f, int pb, int mu) {
  return (const double2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride;
}
__device__ __forceinline__ double2* lptr_w(const LinkField& f, int pb, int mu) {
  return (double2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride;
}
__device__ __forceinline__ const double2* mptr(const MatField& f, int pb) {
  return f.data + (size_t)pb * 9 * f.stride;
}

// c = a b, c = a b^dag, c = a^dag b
__device__ __forceinline__ void mm_nn(const Mat3<double>& a, const Mat3<double>& b, Mat3<double>& c) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += a.e[i][k].re * b.e[k][j].re - a.e[i][k].im * b.e[k][j].im;
        im += a.e[i][k].re * b.e[k][j].im + a.e[i][k].im * b.e[k][j].re;
      }
      c.e[i][j].re = re;
      c.e[i][j].im = im;
    }
}
__device__ __forceinline__ void mm_na(const Mat3<double>& a, const Mat3<double>& b, Mat3<double>& c) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += a.e[i][k].re * b.e[j][k].re + a.e[i][k].im * b.e[j][k].im;
        im += a.e[i][k].im * b.e[j][k].re - a.e[i][k].re * b.e[j][k].im;
      }
      c.e[i][j].re = re;
      c.e[i][j].im = im;
    }
}
__device__ __forceinline__ void mm_an(const Mat3<double>& a, const Mat3<double>& b, Mat3<double>& c) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
      for (int k = 0; k < 3; k++) {
        re += a.e[k][i].re * b.e[k][j].re + a.e[k][i].im * b.e[k][j].im;
        im += a.e[k][i].re * b.e[k][j].im - a.e[k][i].im * b.e[k][j].re;
      }
      c.e[i][j].re = re;
      c.e[i][j].im = im;
    }
}

// ---------------------------------------------------------------------------
// counter-based generator (identical specification to the oracle's)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t sm_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uni(uint64_t key, uint64_t ctr) {
  uint64_t r = sm_mix(key + ctr);
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ void gauss_pair(uint64_t key, uint64_t p, double& g0, double& g1) {
  double u1 = uni(key, 2 * p), u2 = uni(key, 2 * p + 1);
  double r = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586476925286766559 * u2;
  g0 = r * cos(th);
  g1 = r * sin(th);
}

__device__ __forceinline__ uint64_t lex_of(const Geom& g, const Coord& c) {
  return (uint64_t)c.x + (uint64_t)g.X[0] * ((uint64_t)c.y + (uint64_t)g.X[1] * ((uint64_t)c.z + (uint64_t)g.X[2] * (uint64_t)c.t));
}

// U = exp(i*scale*H): 4 halvings, degree-14 Horner Taylor, 4 squarings.
__global__ void __launch_bounds__(256) k_gauge_random(Geom g, uint64_t key, double scale, LinkField U) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Coord c = site_coords(g, cb, pb);
  uint64_t s = lex_of(g, c);
  Mat3<double> A, X, E, T;
  for (int n = 0; n < 9; n++) gauss_pair(key, (4 * s + mu) * 9 + n, A.e[n / 3][n % 3].re, A.e[n / 3][n % 3].im);
  Mat3<double> H;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      H.e[a][b].re = 0.5 * (A.e[a][b].re + A.e[b][a].re);
      H.e[a][b].im = 0.5 * (A.e[a][b].im - A.e[b][a].im);
    }
  double tr = (H.e[0][0].re + H.e[1][1].re + H.e[2][2].re) / 3.0;
  for (int a = 0; a < 3; a++) H.e[a][a].re -= tr;
  // X = (i*scale*H) / 16, in the oracle's order of operations
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      X.e[a][b].re = (-scale * H.e[a][b].im) * 0.0625;
      X.e[a][b].im = (scale * H.e[a][b].re) * 0.0625;
    }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      E.e[a][b].re = (a == b) ? 1.0 : 0.0;
      E.e[a][b].im = 0.0;
    }
  for (int n = 14; n >= 1; n--) {
    mm_nn(X, E, T);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        E.e[a][b].re = T.e[a][b].re / (double)n;
        E.e[a][b].im = T.e[a][b].im / (double)n;
      }
    for (int a = 0; a < 3; a++) E.e[a][a].re += 1.0;
  }
  for (int k = 0; k < 4; k++) {
    mm_nn(E, E, T);
    E = T;
  }
  st(lptr_w(U, pb, mu), U.stride, cb, E);
}

__global__ void k_gaussian_vec(Geom g, uint64_t key, int pb, VecField v) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= g.Vh) return;
  Coord c = site_coords(g, cb, pb);
  uint64_t s = lex_of(g, c);
  const double s2 = 0.70710678118654752440;
  Vec3<double> a;
  for (int k = 0; k < 3; k++) {
    double x, y;
    gauss_pair(key, 3 * s + k, x, y);
    a.c[k].re = s2 * x;
    a.c[k].im = s2 * y;
  }
  store_vec<P_F64>(v, cb, a);
}

__global__ void k_rephase(Geom g, LinkField U) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Coord c = site_coords(g, cb, pb);
  double ph = ks_phase(g, c, mu);
  double2* p = lptr_w(U, pb, mu);
  for (int e = 0; e < 9; e++) {
    double2 a = p[(size_t)e * U.stride + cb];
    p[(size_t)e * U.stride + cb] = make_double2(a.x * ph, a.y * ph);
  }
}

// fat_mu = one_link * U_mu
__global__ void k_fat_init(Geom g, LinkField U, LinkField fat, int mu, double one_link) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pb = blockIdx.y;
  if (cb >= g.Vh) return;
  const double2* s = lptr(U, pb, mu);
  double2* d = lptr_w(fat, pb, mu);
  for (int e = 0; e < 9; e++) {
    double2 a = s[(size_t)e * U.stride + cb];
    d[(size_t)e * fat.stride + cb] = make_double2(one_link * a.x, one_link * a.y);
  }
}

// Generalised staple of the "mu link" field Lf (a direction of a LinkField when
// lmu >= 0, else a MatField):
//   upper(x) = U_nu(x) [L(x+nu) U_nu(x+mu)^dag]
//   lower(x) = [U_nu(x-nu)^dag L(x-nu)] U_nu(x-nu+mu)
// staple != null: staple = upper + lower, fat_mu += coef*staple;
// else fat_mu += coef*upper, then += coef*lower.
__global__ void __launch_bounds__(256) k_gen_staple(Geom g, LinkField U, LinkField Lf, int lmu, MatField Lm, MatField staple,
                             int save, LinkField fat, int mu, int nu, double coef) {
  int bx, pb;
  if (!xcd_block((g.Vh + blockDim.x - 1) / blockDim.x, 2, bx, pb)) return;
  int cb = bx * blockDim.x + threadIdx.x;
  if (cb >= g.Vh) return;
  int po = pb ^ 1;
  Coord c = site_coords(g, cb, pb);
  int xnu = neighbour(g, c, cb, nu, 1);   // x+nu, other parity
  int xmu = neighbour(g, c, cb, mu, 1);   // x+mu, other parity
  int xmn = neighbour(g, c, cb, nu, -1);  // x-nu, other parity
  // x - nu + mu has the parity of x
  Coord cm = site_coords(g, xmn, po);
  int xmnpm = neighbour(g, cm, xmn, mu, 1);
  Mat3<double> A, B, C, t1, up, lo;
  // upper
  if (lmu >= 0) ld(lptr(Lf, po, lmu), Lf.stride, xnu, A);
  else ld(mptr(Lm, po), Lm.stride, xnu, A);
  ld(lptr(U, po, nu), U.stride, xmu, B);
  mm_na(A, B, t1);
  ld(lptr(U, pb, nu), U.stride, cb, C);
  mm_nn(C, t1, up);
  // lower
  ld(lptr(U, po, nu), U.stride, xmn, A);
  if (lmu >= 0) ld(lptr(Lf, po, lmu), Lf.stride, xmn, B);
  else ld(mptr(Lm, po), Lm.stride, xmn, B);
  mm_an(A, B, t1);
  ld(lptr(U, pb, nu), U.stride, xmnpm, C);
  mm_nn(t1, C, lo);
  double2* f = lptr_w(fat, pb, mu);
  Mat3<double> F;
  ld(f, fat.stride, cb, F);
  if (save) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        up.e[i][j].re += lo.e[i][j].re;
        up.e[i][j].im += lo.e[i][j].im;
        F.e[i][j].re += coef * up.e[i][j].re;
        F.e[i][j].im += coef * up.e[i][j].im;
      }
    st(staple.data + (size_t)pb * 9 * staple.stride, staple.stride, cb, up);
  } else {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        F.e[i][j].re += coef * up.e[i][j].re;
        F.e[i][j].im += coef * up.e[i][j].im;
      }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        F.e[i][j].re += coef * lo.e[i][j].re;
        F.e[i][j].im += coef * lo.e[i][j].im;
      }
  }
  st(f, fat.stride, cb, F);
}

// Up to three generalised staples of stored matrix fields for the mu link in
// one pass: for job k, S = S_nu(in)(x) = U_nu(x) in(x+nu) U_nu(x+mu)^dag +
// U_nu(x-nu)^dag in(x-nu) U_nu(x-nu+mu); fat_mu += coef S; when out is given,
// out = S (acc == 0) or out += S (acc == 1).  The fat link is read and written
// once for all jobs.
struct StapleJob {
  MatField in;
  int nu;
  double coef;
  MatField out;  // data == nullptr: not stored
  int acc;
};
struct StapleJobs {
  StapleJob j[3];
  int n;
};

__global__ void __launch_bounds__(256) k_gen_staples(Geom g, LinkField U, StapleJobs jobs, LinkField fat, int mu) {
  int bx, pb;
  if (!xcd_block((g.Vh + blockDim.x - 1) / blockDim.x, 2, bx, pb)) return;
  int cb = bx * blockDim.x + threadIdx.x;
  if (cb >= g.Vh) return;
  int po = pb ^ 1;
  Coord c = site_coords(g, cb, pb);
  int xmu = neighbour(g, c, cb, mu, 1);  // x+mu, other parity
  double2* f = lptr_w(fat, pb, mu);
  Mat3<double> F;
  ld(f, fat.stride, cb, F);
  for (int k = 0; k < jobs.n; k++) {
    const StapleJob& J = jobs.j[k];
    const int nu = J.nu;
    int xnu = neighbour(g, c, cb, nu, 1);   // x+nu, other parity
    int xmn = neighbour(g, c, cb, nu, -1);  // x-nu, other parity
    Coord cm = site_coords(g, xmn, po);
    int xmnpm = neighbour(g, cm, xmn, mu, 1);  // x-nu+mu, parity of x
    Mat3<double> A, B, C, t1, up, lo;
    ld(mptr(J.in, po), J.in.stride, xnu, A);
    ld(lptr(U, po, nu), U.stride, xmu, B);
    mm_na(A, B, t1);
    ld(lptr(U, pb, nu), U.stride, cb, C);
    mm_nn(C, t1, up);
    ld(lptr(U, po, nu), U.stride, xmn, A);
    ld(mptr(J.in, po), J.in.stride, xmn, B);
    mm_an(A, B, t1);
    ld(lptr(U, pb, nu), U.stride, xmnpm, C);
    mm_nn(t1, C, lo);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        up.e[i][j].re += lo.e[i][j].re;
        up.e[i][j].im += lo.e[i][j].im;
        F.e[i][j].re += J.coef * up.e[i][j].re;
        F.e[i][j].im += J.coef * up.e[i][j].im;
      }
    if (J.out.data) {
      double2* o = J.out.data + (size_t)pb * 9 * J.out.stride;
      if (J.acc) {
        ld(o, J.out.stride, cb, A);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            up.e[i][j].re += A.e[i][j].re;
            up.e[i][j].im += A.e[i][j].im;
          }
      }
      st(o, J.out.stride, cb, up);
    }
  }
  st(f, fat.stride, cb, F);
}

// lng_mu(x) = naik * U_mu(x) [U_mu(x+mu) U_mu(x+2mu)]
__global__ void __launch_bounds__(256) k_long_links(Geom g, LinkField U, LinkField lng, double naik) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3, po = pb ^ 1;
  Coord c = site_coords(g, cb, pb);
  int x1 = neighbour(g, c, cb, mu, 1);
  Coord c1 = site_coords(g, x1, po);
  int x2 = neighbour(g, c1, x1, mu, 1);
  Mat3<double> A, B, C, t;
  ld(lptr(U, po, mu), U.stride, x1, A);
  ld(lptr(U, pb, mu), U.stride, x2, B);
  mm_nn(A, B, t);
  ld(lptr(U, pb, mu), U.stride, cb, C);
  mm_nn(C, t, A);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      A.e[i][j].re *= naik;
      A.e[i][j].im *= naik;
    }
  st(lptr_w(lng, pb, mu), lng.stride, cb, A);
}

// W = V (V^dag V)^-1/2 with the KS phase taken out of V and put back into W;
// counts links that meet MILC's SVD-fallback condition.
__global__ void __launch_bounds__(256) k_u3_unitarize(Geom g, LinkField Vf, LinkField Wf, int* nsvd) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Coord c = site_coords(g, cb, pb);
  double ph = ks_phase(g, c, mu);
  Mat3<double> V, Q, Q2, S2, W;
  ld(lptr(Vf, pb, mu), Vf.stride, cb, V);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      V.e[i][j].re *= ph;
      V.e[i][j].im *= ph;
    }
  mm_an(V, V, Q);
  mm_nn(Q, Q, Q2);
  double c0 = Q.e[0][0].re + Q.e[1][1].re + Q.e[2][2].re;
  double c1 = (Q2.e[0][0].re + Q2.e[1][1].re + Q2.e[2][2].re) / 2;
  double q3 = 0;
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) q3 += Q2.e[i][k].re * Q.e[k][i].re - Q2.e[i][k].im * Q.e[k][i].im;
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
  // MILC's SVD-fallback test on |det V|^2 against g0 g1 g2
  {
    const Mat3<double>& M = V;
    double a1re = M.e[1][1].re * M.e[2][2].re - M.e[1][1].im * M.e[2][2].im - M.e[1][2].re * M.e[2][1].re + M.e[1][2].im * M.e[2][1].im;
    double a1im = M.e[1][1].re * M.e[2][2].im + M.e[1][1].im * M.e[2][2].re - M.e[1][2].re * M.e[2][1].im - M.e[1][2].im * M.e[2][1].re;
    double a2re = M.e[1][0].re * M.e[2][2].re - M.e[1][0].im * M.e[2][2].im - M.e[1][2].re * M.e[2][0].re + M.e[1][2].im * M.e[2][0].im;
    double a2im = M.e[1][0].re * M.e[2][2].im + M.e[1][0].im * M.e[2][2].re - M.e[1][2].re * M.e[2][0].im - M.e[1][2].im * M.e[2][0].re;
    double a3re = M.e[1][0].re * M.e[2][1].re - M.e[1][0].im * M.e[2][1].im - M.e[1][1].re * M.e[2][0].re + M.e[1][1].im * M.e[2][0].im;
    double a3im = M.e[1][0].re * M.e[2][1].im + M.e[1][0].im * M.e[2][1].re - M.e[1][1].re * M.e[2][0].im - M.e[1][1].im * M.e[2][0].re;
    double dre = M.e[0][0].re * a1re - M.e[0][0].im * a1im - M.e[0][1].re * a2re + M.e[0][1].im * a2im + M.e[0][2].re * a3re - M.e[0][2].im * a3im;
    double dim = M.e[0][0].im * a1re + M.e[0][0].re * a1im - M.e[0][1].im * a2re - M.e[0][1].re * a2im + M.e[0][2].im * a3re + M.e[0][2].re * a3im;
    double det_check = dre * dre + dim * dim;
    int svd = 0;
    if (det_check != 0 && fabs(det_check - g0 * g1 * g2) / fabs(det_check) > 1e-8) svd = 1;
    if (det_check < 1e-8) svd = 1;
    if (svd) atomicAdd(nsvd, 1);
  }
  double g0s = sqrt(g0), g1s = sqrt(g1), g2s = sqrt(g2);
  double us = g1s + g2s, ws = g1s * g2s;
  double vs = g0s * us + ws;
  us += g0s;
  ws *= g0s;
  double denom = ws * (us * vs - ws);
  double f0 = (us * vs * vs - ws * (us * us + vs)) / denom;
  double f1 = (2 * us * vs - ws - us * us * us) / denom;
  double f2 = us / denom;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      S2.e[i][j].re = f1 * Q.e[i][j].re + f2 * Q2.e[i][j].re;
      S2.e[i][j].im = f1 * Q.e[i][j].im + f2 * Q2.e[i][j].im;
    }
  S2.e[0][0].re += f0;
  S2.e[1][1].re += f0;
  S2.e[2][2].re += f0;
  mm_nn(V, S2, W);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      W.e[i][j].re *= ph;
      W.e[i][j].im *= ph;
    }
  st(lptr_w(Wf, pb, mu), Wf.stride, cb, W);
}

// ===========================================================================
static int nb_for(int n) { return (n + 255) / 256; }

int gauge_random_dev(uint64_t seed, double scale, LinkField& U) {
  Context& c = ctx();
  uint64_t key = sm_mix(seed ^ 0x5851F42D4C957F2DULL);
  hipLaunchKernelGGL(k_gauge_random, dim3(nb_for(c.Vh), 8), dim3(256), 0, c.stream, c.geom, key, scale, U);
  MH_CHECK(hipGetLastError());
  return 0;
}

// Gaussian anti-Hermitian momenta, packed (m01, m02, m12, m00im, m11im,
// m22im, space), weight exp(-|P|^2/2): every real component of the
// off-diagonal entries and r3, r8 with <x^2> = 1/2, diagonal r3 + r8/sqrt(3),
// -r3 + r8/sqrt(3), -2 r8/sqrt(3) (random_anti_hermitian,
// libraries/rand_ahmat.c:17-33).  The random stream is this library's own
// counter-based generator, keyed by seed, lexicographic site and direction.
__global__ void k_ranmom(Geom g, uint64_t key, double* __restrict__ mom) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Coord c = site_coords(g, cb, pb);
  uint64_t s = lex_of(g, c);
  const double h = 0.70710678118654752440, st = 0.57735026918962576451;
  double r[8];
  for (int k = 0; k < 4; k++) gauss_pair(key, (4 * s + mu) * 4 + k, r[2 * k], r[2 * k + 1]);
  for (int k = 0; k < 8; k++) r[k] *= h;
  double* o = mom + ((size_t)(pb * g.Vh + cb) * 4 + mu) * 10;
  o[0] = r[2];
  o[1] = r[3];
  o[2] = r[4];
  o[3] = r[5];
  o[4] = r[6];
  o[5] = r[7];
  o[6] = r[0] + st * r[1];
  o[7] = -r[0] + st * r[1];
  o[8] = -2.0 * st * r[1];
  o[9] = 0.0;
}

int ranmom_dev(uint64_t seed, double* mom) {
  Context& c = ctx();
  uint64_t key = sm_mix(seed ^ 0x9E3779B97F4A7C15ULL);
  hipLaunchKernelGGL(k_ranmom, dim3(nb_for(c.Vh), 8), dim3(256), 0, c.stream, c.geom, key, mom);
  MH_CHECK(hipGetLastError());
  return 0;
}

int gaussian_vec_dev(uint64_t seed, int pb, VecField& v) {
  Context& c = ctx();
  uint64_t key = sm_mix(seed ^ 0x2545F4914F6CDD1DULL);
  hipLaunchKernelGGL(k_gaussian_vec, dim3(nb_for(c.Vh)), dim3(256), 0, c.stream, c.geom, key, pb, v);
  MH_CHECK(hipGetLastError());
  return 0;
}

int rephase_dev(LinkField& U) {
  Context& c = ctx();
  hipLaunchKernelGGL(k_rephase, dim3(nb_for(c.Vh), 8), dim3(256), 0, c.stream, c.geom, U);
  MH_CHECK(hipGetLastError());
  return 0;
}

static int alloc_mat(MatField& m, int stride) {
  m.stride = stride;
  return dev_alloc((void**)&m.data, 2ull * 9 * stride * sizeof(double2));
}

// Optimised fattening: coeff = {one_link, naik, 3-staple, 5-staple, 7-staple, Lepage}
// Per (mu, nu): the 3-staple S_nu(U_mu) (stored), then one pass for the Lepage
// term and the two 5-staples S_rho(S_nu(U_mu)).  The 7-staple S_sig(S_rho(S_nu))
// is linear in its 5-staple, so the two 5-staples that share the fourth
// direction sig (rho and nu swapped) are summed into one field per sig and the
// six 7-staples of a mu link take three staples in one final pass.  Terms with
// coefficient 0 are skipped.
static int fat_links_dev(const LinkField& U, const double coeff[6], LinkField& fat) {
  Context& c = ctx();
  double one_link = coeff[0] - 6.0 * coeff[5];
  MatField staple, five_sum[3], none{nullptr, 0};
  int rc;
  const bool five = coeff[3] != 0.0 || coeff[4] != 0.0, seven = coeff[4] != 0.0;
  struct Guard {
    MatField* m[4];
    ~Guard() {
      for (MatField* f : m) dev_free(f->data);
    }
  } guard{{&staple, &five_sum[0], &five_sum[1], &five_sum[2]}};
  for (MatField* f : guard.m) f->data = nullptr;
  if ((rc = alloc_mat(staple, c.stride))) return rc;
  for (int s = 0; seven && s < 3; s++)
    if ((rc = alloc_mat(five_sum[s], c.stride))) return rc;
  // k_fat_init streams: plain grid; the staple kernels read neighbours: XCD block order
  dim3 grid(nb_for(c.Vh), 2), xgrid = xcd_grid(nb_for(c.Vh), 2);
  for (int dir = 0; dir < 4; dir++) {
    hipLaunchKernelGGL(k_fat_init, grid, dim3(256), 0, c.stream, c.geom, U, fat, dir, one_link);
    if (coeff[2] == 0.0 && coeff[5] == 0.0 && !five) continue;
    int od[3], n = 0;
    for (int r = 0; r < 4; r++)
      if (r != dir) od[n++] = r;
    bool filled[3] = {false, false, false};
    for (int i = 0; i < 3; i++) {
      const int nu = od[i];
      hipLaunchKernelGGL(k_gen_staple, xgrid, dim3(256), 0, c.stream, c.geom, U, U, dir, none, staple, 1, fat, dir,
                         nu, coeff[2]);
      StapleJobs a{};
      if (coeff[5] != 0.0) a.j[a.n++] = StapleJob{staple, nu, coeff[5], none, 0};
      if (five)
        for (int k = 1; k <= 2; k++) {
          // S_rho(S_nu(U_mu)) belongs to the sum of the fourth direction sig
          const int rho = (k == 1) ? od[(i + 1) % 3] : od[(i + 2) % 3];
          const int sl = (k == 1) ? (i + 2) % 3 : (i + 1) % 3;
          a.j[a.n++] = StapleJob{staple, rho, coeff[3], seven ? five_sum[sl] : none, filled[sl] ? 1 : 0};
          filled[sl] = true;
        }
      if (a.n) hipLaunchKernelGGL(k_gen_staples, xgrid, dim3(256), 0, c.stream, c.geom, U, a, fat, dir);
    }
    if (seven) {
      StapleJobs b{};
      for (int s = 0; s < 3; s++) b.j[b.n++] = StapleJob{five_sum[s], od[s], coeff[4], none, 0};
      hipLaunchKernelGGL(k_gen_staples, xgrid, dim3(256), 0, c.stream, c.geom, U, b, fat, dir);
    }
  }
  MH_CHECK(hipGetLastError());
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

int hisq_links_dev(const LinkField& U, LinkField& fat, LinkField& lng, LinkField* W_out, int* nsvd,
                   LinkField* V_out) {
  Context& c = ctx();
  const double c1[6] = {1.0 / 8.0, 0.0, -1.0 / 8.0 * 0.5, 1.0 / 8.0 * 0.25 * 0.5, -1.0 / 8.0 * 0.125 * (1.0 / 6.0),
                        0.0};
  const double c2[6] = {(1.0 / 8.0) + (2.0 * 6.0 / 16.0) + (1.0 / 8.0), -1.0 / 24.0, -1.0 / 8.0 * 0.5,
                        1.0 / 8.0 * 0.25 * 0.5, -1.0 / 8.0 * 0.125 * (1.0 / 6.0), -2.0 / 16.0};
  LinkField V, W;
  int rc;
  if (V_out) V = *V_out;
  else if ((rc = alloc_links(V, P_F64, c.stride))) return rc;
  if (W_out) W = *W_out;
  else if ((rc = alloc_links(W, P_F64, c.stride))) return rc;
  if ((rc = fat_links_dev(U, c1, V))) return rc;
  int* d_nsvd = nullptr;
  if ((rc = dev_alloc((void**)&d_nsvd, sizeof(int)))) return rc;
  MH_CHECK(hipMemsetAsync(d_nsvd, 0, sizeof(int), c.stream));
  hipLaunchKernelGGL(k_u3_unitarize, dim3(nb_for(c.Vh), 8), dim3(256), 0, c.stream, c.geom, V, W, d_nsvd);
  MH_CHECK(hipGetLastError());
  if (!V_out) free_links(V);
  if ((rc = fat_links_dev(W, c2, fat))) return rc;
  hipLaunchKernelGGL(k_long_links, dim3(nb_for(c.Vh), 8), dim3(256), 0, c.stream, c.geom, W, lng, c2[1]);
  MH_CHECK(hipGetLastError());
  int h = 0;
  MH_CHECK(hipMemcpyAsync(&h, d_nsvd, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  dev_free(d_nsvd);
  if (nsvd) *nsvd = h;
  if (!W_out) free_links(W);
  return 0;
}

// V = fat7(U) and W = U(3) projection of V (the auxiliary links the HISQ force
// needs, hisq_auxiliary_t of generic_ks/fermion_links_hisq_load_milc.c)
int hisq_aux_dev(const LinkField& U, LinkField& V, LinkField& W) {
  Context& c = ctx();
  const double c1[6] = {1.0 / 8.0, 0.0, -1.0 / 8.0 * 0.5, 1.0 / 8.0 * 0.25 * 0.5, -1.0 / 8.0 * 0.125 * (1.0 / 6.0),
                        0.0};
  int rc;
  if ((rc = fat_links_dev(U, c1, V))) return rc;
  int* d_nsvd = nullptr;
  if ((rc = dev_alloc((void**)&d_nsvd, sizeof(int)))) return rc;
  MH_CHECK(hipMemsetAsync(d_nsvd, 0, sizeof(int), c.stream));
  hipLaunchKernelGGL(k_u3_unitarize, dim3(nb_for(c.Vh), 8), dim3(256), 0, c.stream, c.geom, V, W, d_nsvd);
  MH_CHECK(hipGetLastError());
  MH_CHECK(hipStreamSynchronize(c.stream));
  dev_free(d_nsvd);
  return 0;
}

int asqtad_links_dev(const LinkField& U, double u0, LinkField& fat, LinkField& lng) {
  Context& c = ctx();
  double u2 = u0 * u0, u4 = u2 * u2, u6 = u4 * u2;
  const double co[6] = {5.0 / 8.0, -1.0 / 24.0 / u2, -1.0 / 16.0 / u2, 1.0 / 64.0 / u4, -1.0 / 384.0 / u6,
                        -1.0 / 16.0 / u4};
  int rc = fat_links_dev(U, co, fat);
  if (rc) return rc;
  hipLaunchKernelGGL(k_long_links, dim3(nb_for(c.Vh), 8), dim3(256), 0, c.stream, c.geom, U, lng, co[1]);
  MH_CHECK(hipGetLastError());
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

}  // namespace mh
