// mh_fnforce.hip -- fat + Naik (FN) fermion force on the thin links.
//
// Restates generic_ks/fermion_force_fn_multi.c:81-313 (fermion_force_fn_multi)
// for the asqtad-type path table (generic_ks/imp_actions/asqtad_action.h:27-45):
// with dS = sum 2 Re Tr[dF_mu(x) eps(x) O1_mu(x) + dL_mu(x) eps(x) O3_mu(x)],
// F_mu(x) = sum_p c_p P_p(x) and L_mu(x) = c_N U_mu(x)U_mu(x+mu)U_mu(x+2mu),
// a link at position j of path p collects
//   M_j = (L_{j+1} ... L_{n-1}) eps(x) O_mu(x) (L_0 ... L_{j-1})
// (its adjoint for a backward link) into G, and the momentum gets
// TA(2 eps_md U G), TA = traceless anti-Hermitian part (make_ahmat.c).
//
// MI355X formulation: gather, not scatter.  One thread owns one target link
// (y, nu) and walks every (path, position) whose link is that link, with the
// path start x = y - offset; no atomics, deterministic order.  The path table
// (344 paths, ~508 entries per target direction) is read uniformly by a
// wavefront and stays in the scalar/L1 caches.
#include <algorithm>
#include <array>
#include <cstring>
#include <initializer_list>
#include <vector>

#include "mh_internal.h"

namespace mh {

struct FnPath {
  int n, naik;
  int d[7];
  double c;
};
struct FnEntry {
  int path, mu, j, adj;
  int off[4];
};

// path table: for direction mu the paths of coefficients
// {one_link, naik, 3-staple, 5-staple, 7-staple, Lepage}; directions use MILC's
// codes, XUP..TUP = 0..3 and OPP_DIR(d) = 7 - d (include/dirs.h)
static void build_paths(const double coeff[6], std::vector<FnPath>& paths, std::vector<int>& path_mu) {
  auto sd = [](int nu, int s) { return s > 0 ? nu : 7 - nu; };
  auto add = [&](int mu, std::initializer_list<int> ds, double c, int naik) {
    FnPath p{};
    p.n = 0;
    for (int d : ds) p.d[p.n++] = d;
    p.c = c;
    p.naik = naik;
    paths.push_back(p);
    path_mu.push_back(mu);
  };
  for (int mu = 0; mu < 4; mu++) {
    add(mu, {mu}, coeff[0], 0);
    add(mu, {mu, mu, mu}, coeff[1], 1);
    for (int nu = 0; nu < 4; nu++) {
      if (nu == mu) continue;
      for (int s = -1; s <= 1; s += 2) {
        add(mu, {sd(nu, s), mu, sd(nu, -s)}, coeff[2], 0);
        add(mu, {sd(nu, s), sd(nu, s), mu, sd(nu, -s), sd(nu, -s)}, coeff[5], 0);
        for (int rho = 0; rho < 4; rho++) {
          if (rho == mu || rho == nu) continue;
          for (int s2 = -1; s2 <= 1; s2 += 2) {
            add(mu, {sd(nu, s), sd(rho, s2), mu, sd(rho, -s2), sd(nu, -s)}, coeff[3], 0);
            for (int sig = 0; sig < 4; sig++) {
              if (sig == mu || sig == nu || sig == rho) continue;
              for (int s3 = -1; s3 <= 1; s3 += 2)
                add(mu, {sd(nu, s), sd(rho, s2), sd(sig, s3), mu, sd(sig, -s3), sd(rho, -s2), sd(nu, -s)},
                    coeff[4], 0);
            }
          }
        }
      }
    }
  }
}

// per target direction: every (path, position) whose link is along that direction
static void build_entries(const std::vector<FnPath>& paths, const std::vector<int>& path_mu,
                          std::vector<FnEntry>& ent, int start[5]) {
  std::vector<FnEntry> per[4];
  for (size_t p = 0; p < paths.size(); p++) {
    int pos[4] = {0, 0, 0, 0};
    for (int j = 0; j < paths[p].n; j++) {
      int d = paths[p].d[j];
      FnEntry e{};
      e.path = (int)p;
      e.mu = path_mu[p];
      e.j = j;
      if (d < 4) {
        e.adj = 0;
        for (int k = 0; k < 4; k++) e.off[k] = pos[k];
        per[d].push_back(e);
        pos[d] += 1;
      } else {
        int dp = 7 - d;
        pos[dp] -= 1;
        e.adj = 1;
        for (int k = 0; k < 4; k++) e.off[k] = pos[k];
        per[dp].push_back(e);
      }
    }
  }
  ent.clear();
  for (int nu = 0; nu < 4; nu++) {
    start[nu] = (int)ent.size();
    ent.insert(ent.end(), per[nu].begin(), per[nu].end());
  }
  start[4] = (int)ent.size();
}

__device__ __forceinline__ void mload(const LinkField& f, int pb, int mu, int cb, Mat3<double>& m) {
  const double2* p = (const double2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    double2 a = p[(size_t)e * f.stride];
    m.e[e / 3][e % 3].re = a.x;
    m.e[e / 3][e % 3].im = a.y;
  }
}
__device__ __forceinline__ void mmul(const Mat3<double>& a, const Mat3<double>& b, Mat3<double>& c) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        re += a.e[i][k].re * b.e[k][j].re - a.e[i][k].im * b.e[k][j].im;
        im += a.e[i][k].re * b.e[k][j].im + a.e[i][k].im * b.e[k][j].re;
      }
      c.e[i][j].re = re;
      c.e[i][j].im = im;
    }
}
__device__ __forceinline__ void madj(Mat3<double>& a) {
  Mat3<double> t = a;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a.e[i][j].re = t.e[j][i].re;
      a.e[i][j].im = -t.e[j][i].im;
    }
}
__device__ __forceinline__ void munit(Mat3<double>& a) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a.e[i][j].re = i == j ? 1.0 : 0.0;
      a.e[i][j].im = 0.0;
    }
}

// (pb, cb) of lattice point (x, y, z, t), coordinates already in range
__device__ __forceinline__ int cb_of(const Geom& g, const int* p, int& pb) {
  pb = (p[0] + p[1] + p[2] + p[3]) & 1;
  return (p[0] >> 1) + g.Xh * (p[1] + g.X[1] * (p[2] + g.X[2] * p[3]));
}
__device__ __forceinline__ int wrap(int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); }

// Gather form of the path-table adjoint.  Thread = target link (y, nu):
//   G_nu(y) = sum over (path p, position j) whose link is (y, nu) of
//             c_p M_j (forward link) or c_p M_j^dag (backward link),
//   M_j = (L_j+1 ... L_n-1) F(x) (L_0 ... L_j-1),  F = F3 on Naik paths else F1,
// F(x) multiplied by eps(x) when with_eps (the level-0 force is eps O).
// MODE 0 writes G; MODE 1 adds TA(scale U_nu(y) G_nu(y)) to MILC's packed
// anti_hermitmat (m01, m02, m12 complex; m00im, m11im, m22im; space) at MILC
// index 4 i + nu, i = node_index (even sites first).
template <int MODE>
__global__ void __launch_bounds__(128) k_path_adjoint(Geom g, LinkField Lk, LinkField F1, LinkField F3, int with_eps,
                                                      const FnPath* __restrict__ paths,
                                                      const FnEntry* __restrict__ ent, int4 e_lo, int4 e_hi,
                                                      LinkField Gout, LinkField U, double scale,
                                                      double* __restrict__ mom) {
  int bx, pm;
  if (!xcd_block((g.Vh + blockDim.x - 1) / blockDim.x, 8, bx, pm)) return;
  int cb = bx * blockDim.x + threadIdx.x;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, nu = pm & 3;
  Coord c = site_coords(g, cb, pb);
  int y[4] = {c.x, c.y, c.z, c.t};
  int lo = nu == 0 ? e_lo.x : (nu == 1 ? e_lo.y : (nu == 2 ? e_lo.z : e_lo.w));
  int hi = nu == 0 ? e_hi.x : (nu == 1 ? e_hi.y : (nu == 2 ? e_hi.z : e_hi.w));
  Mat3<double> G;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) G.e[i][j].re = G.e[i][j].im = 0.0;
  for (int e = lo; e < hi; e++) {
    const FnEntry* E = ent + e;
    const FnPath* P = paths + E->path;
    const int n = P->n, jt = E->j;
    int pos[4];
#pragma unroll
    for (int k = 0; k < 4; k++) pos[k] = wrap(y[k] - E->off[k], g.X[k]);
    int xpb;
    int xcb = cb_of(g, pos, xpb);
    Mat3<double> pre, suf, L, T;
    munit(pre);
    munit(suf);
    for (int k = 0; k < n; k++) {
      int d = P->d[k];
      bool back = d >= 4;
      int ld = back ? 7 - d : d;
      int step = back ? -1 : 1;
      int lpb, lcb;
      if (!back) lcb = cb_of(g, pos, lpb);
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q == ld) pos[q] = wrap(pos[q] + step, g.X[q]);
      if (back) lcb = cb_of(g, pos, lpb);
      if (k == jt) continue;
      mload(Lk, lpb, ld, lcb, L);
      if (back) madj(L);
      if (k < jt) {
        mmul(pre, L, T);
        pre = T;
      } else {
        mmul(suf, L, T);
        suf = T;
      }
    }
    Mat3<double> Ob, A, M;
    double w = ((with_eps && xpb) ? -1.0 : 1.0) * P->c;
    if (F1.data) {
      mload(P->naik ? F3 : F1, xpb, E->mu, xcb, Ob);
      mmul(suf, Ob, A);
      mmul(A, pre, M);
    } else {
      // closed loops (gauge force): the identity at the loop's base
      mmul(suf, pre, M);
    }
    if (E->adj) madj(M);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        G.e[i][j].re += w * M.e[i][j].re;
        G.e[i][j].im += w * M.e[i][j].im;
      }
  }
  if (MODE == 0) {
    double2* out = (double2*)Gout.data + (size_t)(pm * 9) * Gout.stride + cb;
#pragma unroll
    for (int e = 0; e < 9; e++) out[(size_t)e * Gout.stride] = make_double2(G.e[e / 3][e % 3].re, G.e[e / 3][e % 3].im);
    return;
  }
  Mat3<double> Uy, X;
  mload(U, pb, nu, cb, Uy);
  mmul(Uy, G, X);
  // traceless anti-Hermitian part (libraries/make_ahmat.c)
  double tr = (X.e[0][0].im + X.e[1][1].im + X.e[2][2].im) * (1.0 / 3.0);
  double h = 0.5 * scale;
  double* o = mom + ((size_t)(pb * g.Vh + cb) * 4 + nu) * 10;
  o[0] = h * (X.e[0][1].re - X.e[1][0].re);
  o[1] = h * (X.e[0][1].im + X.e[1][0].im);
  o[2] = h * (X.e[0][2].re - X.e[2][0].re);
  o[3] = h * (X.e[0][2].im + X.e[2][0].im);
  o[4] = h * (X.e[1][2].re - X.e[2][1].re);
  o[5] = h * (X.e[1][2].im + X.e[2][1].im);
  o[6] = scale * (X.e[0][0].im - tr);
  o[7] = scale * (X.e[1][1].im - tr);
  o[8] = scale * (X.e[2][2].im - tr);
  o[9] = 0.0;
}

// ---------------------------------------------------------------------------
// Force through the U(3) projection W = V (V^dag V)^-1/2 (the role of
// u3_unit_der_analytic, generic_ks/su3_mat_op.c:1379, in
// fn_fermion_force_multi_hisq_reunit, fermion_force_hisq_multi.c:1877-1998):
// with M = V^dag V = E diag(l) E^dag, Q = M^-1/2, K = F_W V,
//   F_V = Q F_W + E [ f(l_i, l_j) (E^dag (K + K^dag) E)_ij ] E^dag V^dag,
//   f(a, b) = -1 / (sqrt a sqrt b (sqrt a + sqrt b))   (divided difference of
//   x^-1/2, exact also at a = b).
// When min l < filter all l are raised by filter (MILC's HISQ_FORCE_FILTER).
// E from cyclic complex Jacobi rotations in FP64.
// ---------------------------------------------------------------------------
struct C2 {
  double re, im;
};
__device__ __forceinline__ C2 cm(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ C2 cmc(C2 a, C2 b) { return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }  // a conj(b)
__device__ __forceinline__ C2 cca(C2 a, C2 b) { return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }  // conj(a) b
__device__ __forceinline__ C2 cadd(C2 a, C2 b) { return {a.re + b.re, a.im + b.im}; }

__global__ void __launch_bounds__(128) k_u3_adjoint(Geom g, LinkField Vf, LinkField FW, LinkField FV, double filter) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Mat3<double> Vm, Fm;
  mload(Vf, pb, mu, cb, Vm);
  mload(FW, pb, mu, cb, Fm);
  C2 v[3][3], f[3][3], a[3][3], e[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      v[i][j] = {Vm.e[i][j].re, Vm.e[i][j].im};
      f[i][j] = {Fm.e[i][j].re, Fm.e[i][j].im};
      e[i][j] = {i == j ? 1.0 : 0.0, 0.0};
    }
  // a = V^dag V
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; k++) s = cadd(s, cca(v[k][i], v[k][j]));
      a[i][j] = s;
    }
  for (int sweep = 0; sweep < 8; sweep++) {
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
      const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
      double r = sqrt(a[p][q].re * a[p][q].re + a[p][q].im * a[p][q].im);
      if (r == 0.0) continue;
      C2 ph = {a[p][q].re / r, a[p][q].im / r};
      double tau = (a[q][q].re - a[p][p].re) / (2 * r);
      double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
      double c = 1 / sqrt(1 + t * t), s = t * c;
      // J = diag(1, e^{-i phi}) [[c, s], [-s, c]] in the (p, q) plane
      C2 Jpp = {c, 0}, Jpq = {s, 0}, Jqp = {-s * ph.re, s * ph.im}, Jqq = {c * ph.re, -c * ph.im};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        C2 akp = a[k][p], akq = a[k][q];
        a[k][p] = cadd(cm(akp, Jpp), cm(akq, Jqp));
        a[k][q] = cadd(cm(akp, Jpq), cm(akq, Jqq));
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        C2 apk = a[p][k], aqk = a[q][k];
        a[p][k] = cadd(cca(Jpp, apk), cca(Jqp, aqk));
        a[q][k] = cadd(cca(Jpq, apk), cca(Jqq, aqk));
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        C2 ekp = e[k][p], ekq = e[k][q];
        e[k][p] = cadd(cm(ekp, Jpp), cm(ekq, Jqp));
        e[k][q] = cadd(cm(ekp, Jpq), cm(ekq, Jqq));
      }
    }
  }
  double l[3] = {a[0][0].re, a[1][1].re, a[2][2].re};
  double lmin = fmin(l[0], fmin(l[1], l[2]));
  if (lmin < filter) {
    l[0] += filter;
    l[1] += filter;
    l[2] += filter;
  }
  double sq[3] = {sqrt(l[0]), sqrt(l[1]), sqrt(l[2])};
  // H = K + K^dag, K = F V; H' = E^dag H E scaled by f; then E H' E^dag
  C2 k[3][3], t[3][3], hp[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0};
#pragma unroll
      for (int c = 0; c < 3; c++) s = cadd(s, cm(f[i][c], v[c][j]));
      k[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      // (K + K^dag) E
      C2 s = {0, 0};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        C2 h = {k[i][c].re + k[c][i].re, k[i][c].im - k[c][i].im};
        s = cadd(s, cm(h, e[c][j]));
      }
      t[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0};
#pragma unroll
      for (int c = 0; c < 3; c++) s = cadd(s, cca(e[c][i], t[c][j]));
      double fij = -1.0 / (sq[i] * sq[j] * (sq[i] + sq[j]));
      hp[i][j] = {s.re * fij, s.im * fij};
    }
  // t = E H'
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0};
#pragma unroll
      for (int c = 0; c < 3; c++) s = cadd(s, cm(e[i][c], hp[c][j]));
      t[i][j] = s;
    }
  // L = t E^dag (into hp), Q = E diag(1/sq) E^dag (into k)
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0}, qv = {0, 0};
#pragma unroll
      for (int c = 0; c < 3; c++) {
        s = cadd(s, cmc(t[i][c], e[j][c]));
        C2 ec = cmc(e[i][c], e[j][c]);
        qv.re += ec.re / sq[c];
        qv.im += ec.im / sq[c];
      }
      hp[i][j] = s;
      k[i][j] = qv;
    }
  // F_V = Q F + L V^dag
  double2* out = (double2*)FV.data + (size_t)(pm * 9) * FV.stride + cb;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      C2 s = {0, 0};
#pragma unroll
      for (int c = 0; c < 3; c++) s = cadd(s, cadd(cm(k[i][c], f[c][j]), cmc(hp[i][c], v[j][c])));
      out[(size_t)(3 * i + j) * FV.stride] = make_double2(s.re, s.im);
    }
}

// ---------------------------------------------------------------------------
// Staple form of the smearing adjoint: reverse mode through the optimised
// fattening X = (c1 - 6 c_lep) U + sum_nu [c3 S1 + c_lep S_nu(S1) + sum_rho
// (c5 S2 + sum_sig c7 S_sig(S2))], S1 = S_nu(U_mu), S2 = S_rho(S1),
//   S_nu(A)(x) = U_nu(x) A(x+nu) U_nu(x+mu)^dag + U_nu(x-nu)^dag A(x-nu) U_nu(x-nu+mu),
// plus the Naik product.  Every step is a gather over one lattice of 3x3
// matrices (no atomics), 38 launches per direction mu instead of the path
// walk per (path, position).  For unitary links the result differs from the
// path-table gather only by components the traceless anti-Hermitian part (or
// the U(3) projection derivative) removes: the retracing Lepage terms are
// constant on the group.
// ---------------------------------------------------------------------------
// a matrix per site and parity: element e of site (pb, cb) at
// data[(pb * pstride + off + e) * stride + cb]  (pstride 36 / off 9 mu for one
// direction of a link field, pstride 9 / off 0 for a site field)
struct MF {
  double2* data;
  int stride, pstride, off;
};
__host__ inline MF mf_link(const LinkField& f, int mu) { return MF{(double2*)f.data, f.stride, 36, 9 * mu}; }
__host__ inline MF mf_site(double2* d, int stride) { return MF{d, stride, 9, 0}; }

__device__ __forceinline__ void mfl(const MF& f, int pb, int cb, Mat3<double>& m) {
  const double2* p = f.data + (size_t)(pb * f.pstride + f.off) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    double2 a = p[(size_t)e * f.stride];
    m.e[e / 3][e % 3].re = a.x;
    m.e[e / 3][e % 3].im = a.y;
  }
}
__device__ __forceinline__ void mfs(const MF& f, int pb, int cb, const Mat3<double>& m) {
  double2* p = f.data + (size_t)(pb * f.pstride + f.off) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) p[(size_t)e * f.stride] = make_double2(m.e[e / 3][e % 3].re, m.e[e / 3][e % 3].im);
}
// f += c m
__device__ __forceinline__ void mfa(const MF& f, int pb, int cb, double c, const Mat3<double>& m) {
  double2* p = f.data + (size_t)(pb * f.pstride + f.off) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    double2 a = p[(size_t)e * f.stride];
    a.x += c * m.e[e / 3][e % 3].re;
    a.y += c * m.e[e / 3][e % 3].im;
    p[(size_t)e * f.stride] = a;
  }
}
// Non-temporal forms for the own-site accumulators of the staple kernels,
// which no other site reads in the same pass: their lines are not kept in L2,
// leaving it to the neighbour-read links and forces.
__device__ __forceinline__ void mfl_nt(const MF& f, int pb, int cb, Mat3<double>& m) {
  const double2* p = f.data + (size_t)(pb * f.pstride + f.off) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    const nat_d2 a = ldnt(p + (size_t)e * f.stride);
    m.e[e / 3][e % 3].re = a.x;
    m.e[e / 3][e % 3].im = a.y;
  }
}
__device__ __forceinline__ void mfs_nt(const MF& f, int pb, int cb, const Mat3<double>& m) {
  double2* p = f.data + (size_t)(pb * f.pstride + f.off) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    nat_d2 a;
    a.x = m.e[e / 3][e % 3].re;
    a.y = m.e[e / 3][e % 3].im;
    __builtin_nontemporal_store(a, (nat_d2*)(p + (size_t)e * f.stride));
  }
}
// f += c m, non-temporal
__device__ __forceinline__ void mfa_nt(const MF& f, int pb, int cb, double c, const Mat3<double>& m) {
  double2* p = f.data + (size_t)(pb * f.pstride + f.off) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    nat_d2 a = ldnt(p + (size_t)e * f.stride);
    a.x += c * m.e[e / 3][e % 3].re;
    a.y += c * m.e[e / 3][e % 3].im;
    __builtin_nontemporal_store(a, (nat_d2*)(p + (size_t)e * f.stride));
  }
}
__device__ __forceinline__ void mmul_na(const Mat3<double>& a, const Mat3<double>& b, Mat3<double>& c) {
  Mat3<double> t = b;
  madj(t);
  mmul(a, t, c);
}
__device__ __forceinline__ void mmul_an(const Mat3<double>& a, const Mat3<double>& b, Mat3<double>& c) {
  Mat3<double> t = a;
  madj(t);
  mmul(t, b, c);
}
__device__ __forceinline__ void madd(Mat3<double>& a, const Mat3<double>& b) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a.e[i][j].re += b.e[i][j].re;
      a.e[i][j].im += b.e[i][j].im;
    }
}
__device__ __forceinline__ void madd_adj(Mat3<double>& a, const Mat3<double>& b) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a.e[i][j].re += b.e[j][i].re;
      a.e[i][j].im -= b.e[j][i].im;
    }
}

// F(x) = eps(x)^with_eps F1_mu(x);  G_mu(x) += c1 F(x)
__global__ void __launch_bounds__(128) k_sf_init(Geom g, MF F1, int with_eps, MF F, MF G, double c1) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pb = blockIdx.y;
  if (cb >= g.Vh) return;
  Mat3<double> m;
  mfl(F1, pb, cb, m);
  double e = (with_eps && pb) ? -1.0 : 1.0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      m.e[i][j].re *= e;
      m.e[i][j].im *= e;
    }
  mfs(F, pb, cb, m);
  if (c1 != 0.0) mfa(G, pb, cb, c1, m);
}

// Sites of the staple kernels, with the two directions (MU, NU) fixed at
// compile time: one coordinate decode, then every neighbour cb from
// coordinates.  Site order: TT = 0 is the XCD-contiguous cb order of
// xcd_block; TT = 1 (pairs with NU or MU = t) gives each wavefront 16
// consecutive cb of four consecutive t-slices, so that the +-t neighbours of
// three of every four sites are read by the same wavefront and the t-hop
// rows come through L1/L2 instead of being fetched again from HBM (needs
// st % 16 == 0 and nt % 4 == 0; the block is 128 threads, two wavefronts).
struct SfSites {
  int ypn, ymn, ypm, ymm, ymnpm, ymmpn;  // y+nu, y-nu, y+mu, y-mu, y-nu+mu, y-mu+nu
};
__device__ __forceinline__ int sf_wrap(int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); }
template <int MU, int NU>
__device__ __forceinline__ SfSites sf_sites(const Geom& g, int cb, int pb) {
  const Coord c = site_coords(g, cb, pb);
  const int x[4] = {c.x, c.y, c.z, c.t};
  auto at = [&](int dm, int dn) {
    int v[4] = {x[0], x[1], x[2], x[3]};
    v[MU] = sf_wrap(v[MU] + dm, g.X[MU]);
    v[NU] = sf_wrap(v[NU] + dn, g.X[NU]);
    return (v[0] + g.X[0] * (v[1] + g.X[1] * (v[2] + g.X[2] * v[3]))) >> 1;
  };
  SfSites n;
  n.ypn = at(0, 1);
  n.ymn = at(0, -1);
  n.ypm = at(1, 0);
  n.ymm = at(-1, 0);
  n.ymnpm = at(1, -1);
  n.ymmpn = at(-1, 1);
  return n;
}
template <int TT>
__device__ __forceinline__ bool sf_site(const Geom& g, int& cb, int& pb) {
  int bx;
  if (!xcd_block((g.Vh + 127) / 128, 2, bx, pb)) return false;
  if (TT == 0) {
    cb = bx * 128 + threadIdx.x;
    return cb < g.Vh;
  }
  const int lane = threadIdx.x & 63;
  const int w = bx * 2 + (threadIdx.x >> 6);  // wavefront index
  const int ns = g.st >> 4;                   // 16-site chunks per t-slice
  const int s = w % ns, t = (w / ns) * 4 + (lane >> 4);
  cb = t * g.st + s * 16 + (lane & 15);
  return t < g.X[3];
}
__host__ inline bool sf_tiled(const Geom& g, int mu, int nu) {
  return (mu == 3 || nu == 3) && g.st % 16 == 0 && g.X[3] % 4 == 0;
}

// out = S_nu(A) for the mu link
template <int MU, int NU, int TT>
__global__ void __launch_bounds__(128) k_sf_fwd(Geom g, LinkField U, MF A, MF out) {
  int cb, pb;
  if (!sf_site<TT>(g, cb, pb)) return;
  const int po = pb ^ 1;
  const SfSites n = sf_sites<MU, NU>(g, cb, pb);
  Mat3<double> a, b, t, up, lo;
  mfl(A, po, n.ypn, a);
  mload(U, po, NU, n.ypm, b);
  mmul_na(a, b, t);
  mload(U, pb, NU, cb, b);
  mmul(b, t, up);
  mload(U, po, NU, n.ymn, a);
  mfl(A, po, n.ymn, b);
  mmul_an(a, b, t);
  mload(U, pb, NU, n.ymnpm, b);
  mmul(t, b, lo);
  madd(up, lo);
  mfs(out, pb, cb, up);
}

// out = S_P(AP) + S_Q(AQ) for the mu link: the two 5-staples of one 7-staple
// direction in one pass
template <int MU, int P, int Q, int TT>
__global__ void __launch_bounds__(128) k_sf_fwd2(Geom g, LinkField U, MF AP, MF AQ, MF out) {
  int cb, pb;
  if (!sf_site<TT>(g, cb, pb)) return;
  const int po = pb ^ 1;
  Mat3<double> a, b, t, up, lo;
  {
    const SfSites n = sf_sites<MU, P>(g, cb, pb);
    mfl(AP, po, n.ypn, a);
    mload(U, po, P, n.ypm, b);
    mmul_na(a, b, t);
    mload(U, pb, P, cb, b);
    mmul(b, t, up);
    mload(U, po, P, n.ymn, a);
    mfl(AP, po, n.ymn, b);
    mmul_an(a, b, t);
    mload(U, pb, P, n.ymnpm, b);
    mmul(t, b, lo);
    madd(up, lo);
  }
  {
    const SfSites n = sf_sites<MU, Q>(g, cb, pb);
    mfl(AQ, po, n.ypn, a);
    mload(U, po, Q, n.ypm, b);
    mmul_na(a, b, t);
    mload(U, pb, Q, cb, b);
    mmul(b, t, lo);
    madd(up, lo);
    mload(U, po, Q, n.ymn, a);
    mfl(AQ, po, n.ymn, b);
    mmul_an(a, b, t);
    mload(U, pb, Q, n.ymnpm, b);
    mmul(t, b, lo);
    madd(up, lo);
  }
  mfs(out, pb, cb, up);
}
typedef void (*SfFwd2Fn)(Geom, LinkField, MF, MF, MF);
template <int MU, int P, int Q>
static SfFwd2Fn sf_fwd2_pick(bool tt) {
  return tt ? k_sf_fwd2<MU, P, Q, 1> : k_sf_fwd2<MU, P, Q, 0>;
}
// the fused kernel of mu and the pair p < q (both != mu)
static SfFwd2Fn sf_fwd2_kernel(const Geom& g, int mu, int p, int q) {
  const bool tt = (mu == 3 || q == 3) && g.st % 16 == 0 && g.X[3] % 4 == 0;
  switch (mu * 16 + p * 4 + q) {
    case 0 * 16 + 1 * 4 + 2: return sf_fwd2_pick<0, 1, 2>(tt);
    case 0 * 16 + 1 * 4 + 3: return sf_fwd2_pick<0, 1, 3>(tt);
    case 0 * 16 + 2 * 4 + 3: return sf_fwd2_pick<0, 2, 3>(tt);
    case 1 * 16 + 0 * 4 + 2: return sf_fwd2_pick<1, 0, 2>(tt);
    case 1 * 16 + 0 * 4 + 3: return sf_fwd2_pick<1, 0, 3>(tt);
    case 1 * 16 + 2 * 4 + 3: return sf_fwd2_pick<1, 2, 3>(tt);
    case 2 * 16 + 0 * 4 + 1: return sf_fwd2_pick<2, 0, 1>(tt);
    case 2 * 16 + 0 * 4 + 3: return sf_fwd2_pick<2, 0, 3>(tt);
    case 2 * 16 + 1 * 4 + 3: return sf_fwd2_pick<2, 1, 3>(tt);
    case 3 * 16 + 0 * 4 + 1: return sf_fwd2_pick<3, 0, 1>(tt);
    case 3 * 16 + 0 * 4 + 2: return sf_fwd2_pick<3, 0, 2>(tt);
    default: return sf_fwd2_pick<3, 1, 2>(tt);
  }
}

// reverse of O = S_nu(A) for the force Gs on O scaled by cg:
//   FA(y)    += cg [U_nu(y-nu+mu)^dag G(y-nu) U_nu(y-nu) + U_nu(y+mu) G(y+nu) U_nu(y)^dag]
//   FU_nu(y) += cg [A(y+nu) U_nu(y+mu)^dag G(y) + (G(y-mu) U_nu(y-mu) A(y-mu+nu))^dag
//                   + (A(y) U_nu(y+mu) G(y+nu))^dag + G(y-mu+nu) U_nu(y-mu)^dag A(y-mu)]
// With fa_init the FA term is not accumulated but initialised:
//   FA(y) = ci FI(y) + cg [...]
// which stands for a preceding FA = ci FI pass (the scaled staple force).
// MODE 0 computes both terms; MODE 1 only the FA term, MODE 2 only the FU term
// (FA depends on Gs alone, FU is linear in A: the 7-staple driver forms each
// FA once per direction pair and the FU of two 5-staples from their sum).
template <int MU, int NU, int TT, int MODE>
__global__ void __launch_bounds__(128) k_sf_adj(Geom g, LinkField U, MF A, MF Gs, double cg, MF FA, MF FU,
                                                int fa_init, MF FI, double ci) {
  int cb, pb;
  if (!sf_site<TT>(g, cb, pb)) return;
  const int po = pb ^ 1;
  const SfSites n = sf_sites<MU, NU>(g, cb, pb);
  Mat3<double> a, b, t, u, acc;
  if (MODE != 2) {
    mload(U, pb, NU, n.ymnpm, u);
    mfl(Gs, po, n.ymn, b);
    mmul_an(u, b, t);
    mload(U, po, NU, n.ymn, u);
    mmul(t, u, acc);
    mload(U, po, NU, n.ypm, u);
    mfl(Gs, po, n.ypn, b);
    mmul(u, b, t);
    mload(U, pb, NU, cb, u);
    mmul_na(t, u, a);
    madd(acc, a);
    if (fa_init) {
      mfl_nt(FI, pb, cb, a);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          double re = ci * a.e[i][j].re, im = ci * a.e[i][j].im;
          a.e[i][j].re = re + cg * acc.e[i][j].re;
          a.e[i][j].im = im + cg * acc.e[i][j].im;
        }
      mfs_nt(FA, pb, cb, a);
    } else {
      mfa_nt(FA, pb, cb, cg, acc);
    }
  }
  if (MODE == 1) return;
  // A(y+nu) U_nu(y+mu)^dag G(y)
  mfl(A, po, n.ypn, a);
  mload(U, po, NU, n.ypm, u);
  mmul_na(a, u, t);
  mfl(Gs, pb, cb, b);
  mmul(t, b, acc);
  // (G(y-mu) U_nu(y-mu) A(y-mu+nu))^dag
  mfl(Gs, po, n.ymm, b);
  mload(U, po, NU, n.ymm, u);
  mmul(b, u, t);
  mfl(A, pb, n.ymmpn, a);
  mmul(t, a, b);
  madd_adj(acc, b);
  // (A(y) U_nu(y+mu) G(y+nu))^dag
  mfl(A, pb, cb, a);
  mload(U, po, NU, n.ypm, u);
  mmul(a, u, t);
  mfl(Gs, po, n.ypn, b);
  mmul(t, b, a);
  madd_adj(acc, a);
  // G(y-mu+nu) U_nu(y-mu)^dag A(y-mu)
  mfl(Gs, pb, n.ymmpn, b);
  mload(U, po, NU, n.ymm, u);
  mmul_na(b, u, t);
  mfl(A, po, n.ymm, a);
  mmul(t, a, b);
  madd(acc, b);
  mfa_nt(FU, pb, cb, cg, acc);
}

typedef void (*SfFwdFn)(Geom, LinkField, MF, MF);
typedef void (*SfAdjFn)(Geom, LinkField, MF, MF, double, MF, MF, int, MF, double);
struct SfKernels {
  SfFwdFn fwd;
  SfAdjFn adj[3];  // by MODE
};
template <int MU, int NU, int TT>
static SfKernels sf_pick() {
  return SfKernels{k_sf_fwd<MU, NU, TT>, {k_sf_adj<MU, NU, TT, 0>, k_sf_adj<MU, NU, TT, 1>, k_sf_adj<MU, NU, TT, 2>}};
}
// kernels of the direction pair (mu, nu), mu != nu
static SfKernels sf_kernels(const Geom& g, int mu, int nu) {
  const bool tt = sf_tiled(g, mu, nu);
  switch (mu * 4 + nu) {
    case 1: return tt ? sf_pick<0, 1, 1>() : sf_pick<0, 1, 0>();
    case 2: return tt ? sf_pick<0, 2, 1>() : sf_pick<0, 2, 0>();
    case 3: return tt ? sf_pick<0, 3, 1>() : sf_pick<0, 3, 0>();
    case 4: return tt ? sf_pick<1, 0, 1>() : sf_pick<1, 0, 0>();
    case 6: return tt ? sf_pick<1, 2, 1>() : sf_pick<1, 2, 0>();
    case 7: return tt ? sf_pick<1, 3, 1>() : sf_pick<1, 3, 0>();
    case 8: return tt ? sf_pick<2, 0, 1>() : sf_pick<2, 0, 0>();
    case 9: return tt ? sf_pick<2, 1, 1>() : sf_pick<2, 1, 0>();
    case 11: return tt ? sf_pick<2, 3, 1>() : sf_pick<2, 3, 0>();
    case 12: return tt ? sf_pick<3, 0, 1>() : sf_pick<3, 0, 0>();
    case 13: return tt ? sf_pick<3, 1, 1>() : sf_pick<3, 1, 0>();
    default: return tt ? sf_pick<3, 2, 1>() : sf_pick<3, 2, 0>();
  }
}

// Naik: G_mu(y) += cN [U(y+mu) U(y+2mu) F3(y) + U(y+mu) F3(y-mu) U(y-mu) + F3(y-2mu) U(y-2mu) U(y-mu)],
// F3 multiplied by eps at its own site when with_eps
// (direction MU at compile time: one coordinate decode, the hops from it)
template <int MU>
__global__ void __launch_bounds__(128) k_sf_naik(Geom g, LinkField U, MF F3, int with_eps, double cN, MF G) {
  constexpr int mu = MU;
  int cb, pb;
  if (!sf_site<0>(g, cb, pb)) return;
  int po = pb ^ 1;
  const Coord c = site_coords(g, cb, pb);
  const int x[4] = {c.x, c.y, c.z, c.t};
  auto at = [&](int d) {
    int v[4] = {x[0], x[1], x[2], x[3]};
    v[MU] = sf_wrap(v[MU] + d, g.X[MU]);
    return (v[0] + g.X[0] * (v[1] + g.X[1] * (v[2] + g.X[2] * v[3]))) >> 1;
  };
  const int p1 = at(1), m1 = at(-1), p2 = at(2), m2 = at(-2);
  const double ey = (with_eps && pb) ? -1.0 : 1.0, eo = (with_eps && po) ? -1.0 : 1.0;
  Mat3<double> u1, u2, f, t, acc, b;
  mload(U, po, mu, p1, u1);
  mload(U, pb, mu, p2, u2);
  mmul(u1, u2, t);
  mfl(F3, pb, cb, f);
  mmul(t, f, acc);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      acc.e[i][j].re *= ey;
      acc.e[i][j].im *= ey;
    }
  mfl(F3, po, m1, f);
  mmul(u1, f, t);
  mload(U, po, mu, m1, u2);
  mmul(t, u2, b);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      acc.e[i][j].re += eo * b.e[i][j].re;
      acc.e[i][j].im += eo * b.e[i][j].im;
    }
  mfl(F3, pb, m2, f);
  mload(U, pb, mu, m2, u1);
  mmul(f, u1, t);
  mmul(t, u2, b);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      acc.e[i][j].re += ey * b.e[i][j].re;
      acc.e[i][j].im += ey * b.e[i][j].im;
    }
  mfa_nt(G, pb, cb, cN, acc);
}

// mom += TA(scale U G), packed (as k_path_adjoint<1>)
__global__ void __launch_bounds__(128) k_mom_ta(Geom g, LinkField U, LinkField G, double scale,
                                                double* __restrict__ mom, int accumulate) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, nu = pm & 3;
  Mat3<double> Uy, Gy, X;
  mload(U, pb, nu, cb, Uy);
  mload(G, pb, nu, cb, Gy);
  mmul(Uy, Gy, X);
  double tr = (X.e[0][0].im + X.e[1][1].im + X.e[2][2].im) * (1.0 / 3.0);
  double h = 0.5 * scale;
  double* o = mom + ((size_t)(pb * g.Vh + cb) * 4 + nu) * 10;
  o[0] = (accumulate ? o[0] : 0.0) + h * (X.e[0][1].re - X.e[1][0].re);
  o[1] = (accumulate ? o[1] : 0.0) + h * (X.e[0][1].im + X.e[1][0].im);
  o[2] = (accumulate ? o[2] : 0.0) + h * (X.e[0][2].re - X.e[2][0].re);
  o[3] = (accumulate ? o[3] : 0.0) + h * (X.e[0][2].im + X.e[2][0].im);
  o[4] = (accumulate ? o[4] : 0.0) + h * (X.e[1][2].re - X.e[2][1].re);
  o[5] = (accumulate ? o[5] : 0.0) + h * (X.e[1][2].im + X.e[2][1].im);
  o[6] = (accumulate ? o[6] : 0.0) + scale * (X.e[0][0].im - tr);
  o[7] = (accumulate ? o[7] : 0.0) + scale * (X.e[1][1].im - tr);
  o[8] = (accumulate ? o[8] : 0.0) + scale * (X.e[2][2].im - tr);
  o[9] = 0.0;
}

// Path tables, built once per coefficient set and kept on the device (a few
// sets: asqtad, HISQ level 2, HISQ level 1).
struct FnTable {
  double coeff[6] = {0, 0, 0, 0, 0, 0};
  FnPath* paths = nullptr;
  FnEntry* ent = nullptr;
  int start[5] = {0, 0, 0, 0, 0};
  bool valid = false;
};
static FnTable g_tables[4];
static int g_next_table = 0;

static int fn_table(const double coeff[6], FnTable*& out) {
  Context& c = ctx();
  for (auto& t : g_tables) {
    if (!t.valid) continue;
    bool same = true;
    for (int i = 0; i < 6 && same; i++) same = t.coeff[i] == coeff[i];
    if (same) {
      out = &t;
      return 0;
    }
  }
  FnTable& t = g_tables[g_next_table];
  g_next_table = (g_next_table + 1) % 4;
  if (t.valid) {
    dev_free(t.paths);
    dev_free(t.ent);
    t = FnTable{};
  }
  std::vector<FnPath> paths;
  std::vector<int> path_mu;
  std::vector<FnEntry> ent;
  build_paths(coeff, paths, path_mu);
  build_entries(paths, path_mu, ent, t.start);
  int rc;
  if ((rc = dev_alloc((void**)&t.paths, paths.size() * sizeof(FnPath))) ||
      (rc = dev_alloc((void**)&t.ent, ent.size() * sizeof(FnEntry))))
    return rc;
  MH_CHECK(hipMemcpyAsync(t.paths, paths.data(), paths.size() * sizeof(FnPath), hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipMemcpyAsync(t.ent, ent.data(), ent.size() * sizeof(FnEntry), hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  for (int i = 0; i < 6; i++) t.coeff[i] = coeff[i];
  t.valid = true;
  out = &t;
  return 0;
}

void fn_table_release() {
  for (auto& t : g_tables) {
    if (t.valid) {
      dev_free(t.paths);
      dev_free(t.ent);
    }
    t = FnTable{};
  }
  g_next_table = 0;
}

static int path_adjoint_launch(int mode, const LinkField& Lk, const LinkField& F1, const LinkField& F3, int with_eps,
                               const double coeff[6], const LinkField& Gout, const LinkField& U, double scale,
                               double* mom) {
  Context& c = ctx();
  FnTable* t;
  int rc = fn_table(coeff, t);
  if (rc) return rc;
  int4 lo = make_int4(t->start[0], t->start[1], t->start[2], t->start[3]);
  int4 hi = make_int4(t->start[1], t->start[2], t->start[3], t->start[4]);
  dim3 grid = xcd_grid((c.Vh + 127) / 128, 8);
  if (mode == 0)
    hipLaunchKernelGGL(k_path_adjoint<0>, grid, dim3(128), 0, c.stream, c.geom, Lk, F1, F3, with_eps, t->paths,
                       t->ent, lo, hi, Gout, U, scale, mom);
  else
    hipLaunchKernelGGL(k_path_adjoint<1>, grid, dim3(128), 0, c.stream, c.geom, Lk, F1, F3, with_eps, t->paths,
                       t->ent, lo, hi, Gout, U, scale, mom);
  MH_CHECK(hipGetLastError());
  return 0;
}

// Packed momentum increments (MILC AoS, 10 doubles per link) into mom_dev from
// the outer products O1/O3 (FP64), thin phased links U (FP64) and the path
// coefficients; scale multiplies U G (2 eps in fermion_force_fn_multi, 1 when
// the outer products already carry 2 eps, fermion_force_asqtad_gpu.c:53-56).
static int smear_adjoint_dev(const LinkField& L, const LinkField& F1, const LinkField& F3, int with_eps,
                             const double coeff[6], LinkField& G);
// Which device form the thin-link force adjoints use.  The staple form walks
// the optimised fattening, whose Lepage term includes retracing paths
// U_nu U_nu^+ that equal 1 only for unitary links; the path table holds the
// fattening's paths themselves.  0: path table everywhere; 1 (default): the
// staple form where it is exact for any links (a smearing without a Lepage
// term, or the U(3)-projected links W, unitary by construction), the path
// table otherwise; 2: the staple form everywhere (exact for unitary links).
static int g_force_form = 1;
void set_force_form(int form) { g_force_form = form; }
static bool use_staple_form(const double coeff[6], bool links_unitary) {
  return g_force_form == 2 || (g_force_form == 1 && (links_unitary || coeff[5] == 0.0));
}

int fn_force_from_oprod(const LinkField& U, const LinkField& o1, const LinkField& o3, const double coeff[6],
                        double scale, double* mom_dev) {
  LinkField none{nullptr, 1.0f, 0};
  if (!use_staple_form(coeff, false)) return path_adjoint_launch(1, U, o1, o3, 1, coeff, none, U, scale, mom_dev);
  Context& c = ctx();
  LinkField G{};
  int rc;
  if ((rc = alloc_links(G, P_F64, c.stride))) return rc;
  rc = smear_adjoint_dev(U, o1, o3, 1, coeff, G);
  if (!rc) {
    hipLaunchKernelGGL(k_mom_ta, dim3((c.Vh + 127) / 128, 8), dim3(128), 0, c.stream, c.geom, U, G, scale, mom_dev, 0);
    rc = hipGetLastError() == hipSuccess ? 0 : -1;
  }
  MH_CHECK(hipStreamSynchronize(c.stream));
  free_links(G);
  return rc;
}

// Gauge loops of the Symanzik 1x1 + 1x2 + 1x1x1 action: every rotation and
// reflection of the plaquette, the 1x2 rectangle and the 1x1x1 parallelogram,
// one loop per class of cyclic shifts and reversal (the loop set of
// make_loop_table, generic/gauge_stuff.c:68-181: 6 + 12 + 16 loops), as
// closed paths for the gather kernel.
static void build_gauge_paths(const double coeff[3], std::vector<FnPath>& paths, std::vector<int>& path_mu) {
  static const int shapes[3][6] = {{0, 1, 7, 6, 0, 0}, {0, 0, 1, 7, 7, 6}, {0, 1, 2, 7, 6, 5}};
  static const int slen[3] = {4, 6, 6};
  std::vector<std::vector<int>> keys;
  auto canon = [](const std::vector<int>& d) {
    std::vector<int> best;
    const int n = (int)d.size();
    std::vector<int> rv(n);
    for (int j = 0; j < n; j++) rv[j] = 7 - d[n - 1 - j];
    const std::vector<int>* both[2] = {&d, &rv};
    for (const std::vector<int>* v : both)
      for (int r = 0; r < n; r++) {
        std::vector<int> t(n);
        for (int j = 0; j < n; j++) t[j] = (*v)[(r + j) % n];
        if (best.empty() || t < best) best = t;
      }
    return best;
  };
  int perm[4] = {0, 1, 2, 3};
  std::vector<std::array<int, 4>> perms;
  do perms.push_back({perm[0], perm[1], perm[2], perm[3]});
  while (std::next_permutation(perm, perm + 4));
  for (int s = 0; s < 3; s++)
    for (const auto& pm : perms)
      for (int refl = 0; refl < 16; refl++) {
        int map[8];
        for (int a = 0; a < 4; a++) {
          map[a] = ((refl >> a) & 1) ? 7 - pm[a] : pm[a];
          map[7 - a] = 7 - map[a];
        }
        std::vector<int> d(slen[s]);
        for (int j = 0; j < slen[s]; j++) d[j] = map[shapes[s][j]];
        std::vector<int> k = canon(d);
        k.push_back(s);
        if (std::find(keys.begin(), keys.end(), k) != keys.end()) continue;
        keys.push_back(k);
        FnPath p{};
        p.n = slen[s];
        for (int j = 0; j < p.n; j++) p.d[j] = d[j];
        p.c = coeff[s];
        p.naik = 0;
        paths.push_back(p);
        path_mu.push_back(0);  // unused: the loop closes on its base, no force field is read
      }
}

// The gauge-loop table (one, rebuilt when the coefficients change)
struct GaugeTable {
  double coeff[3] = {0, 0, 0};
  FnPath* paths = nullptr;
  FnEntry* ent = nullptr;
  int start[5] = {0, 0, 0, 0, 0};
  int npaths = 0;
  bool valid = false;
};
static GaugeTable g_gauge_table;

static int gauge_table(const double coeff[3], GaugeTable*& out) {
  Context& c = ctx();
  GaugeTable& t = g_gauge_table;
  if (t.valid && t.coeff[0] == coeff[0] && t.coeff[1] == coeff[1] && t.coeff[2] == coeff[2]) {
    out = &t;
    return 0;
  }
  dev_free(t.paths);
  dev_free(t.ent);
  t = GaugeTable{};
  std::vector<FnPath> paths;
  std::vector<int> path_mu;
  std::vector<FnEntry> ent;
  build_gauge_paths(coeff, paths, path_mu);
  build_entries(paths, path_mu, ent, t.start);
  int rc;
  if ((rc = dev_alloc((void**)&t.paths, paths.size() * sizeof(FnPath))) ||
      (rc = dev_alloc((void**)&t.ent, ent.size() * sizeof(FnEntry))))
    return rc;
  MH_CHECK(hipMemcpyAsync(t.paths, paths.data(), paths.size() * sizeof(FnPath), hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipMemcpyAsync(t.ent, ent.data(), ent.size() * sizeof(FnEntry), hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  for (int k = 0; k < 3; k++) t.coeff[k] = coeff[k];
  t.npaths = (int)paths.size();
  t.valid = true;
  out = &t;
  return 0;
}

void gauge_table_release() {
  dev_free(g_gauge_table.paths);
  dev_free(g_gauge_table.ent);
  g_gauge_table = GaugeTable{};
}

// The 48 staples of a link in direction mu, as paths from x+mu back to x:
// signed directions (d < 4 forward, 7 - d backward) packed 3 bits per step,
// path length, and the loop class (0 plaquette, 1 rectangle, 2 parallelogram).
struct GaugeStaples {
  int code[4][48];
  int len[4][48];
  int cls[4][48];
};

static GaugeStaples make_gauge_staples() {
  GaugeStaples t{};
  for (int mu = 0; mu < 4; mu++) {
    const int f = mu, b = 7 - mu;
    int q = 0;
    auto add = [&](std::initializer_list<int> d, int cls) {
      int code = 0, k = 0;
      for (int x : d) code |= x << (3 * k++);
      t.code[mu][q] = code;
      t.len[mu][q] = k;
      t.cls[mu][q] = cls;
      q++;
    };
    for (int nu = 0; nu < 4; nu++) {
      if (nu == mu) continue;
      for (int s = 0; s < 2; s++) {
        const int dn = s ? 7 - nu : nu, dnb = 7 - dn;
        add({dn, b, dnb}, 0);
        add({f, dn, b, b, dnb}, 1);
        add({dn, b, b, dnb, f}, 1);
        add({dn, dn, b, dnb, dnb}, 1);
        for (int rho = 0; rho < 4; rho++) {
          if (rho == mu || rho == nu) continue;
          for (int r = 0; r < 2; r++) {
            const int dr = r ? 7 - rho : rho;
            add({dn, dr, b, dnb, 7 - dr}, 2);
          }
        }
      }
    }
  }
  return t;
}

// Gauge force of the Symanzik 1x1 + 1x2 + 1x1x1 action, one thread per link
// (x, mu): G = sum of coeff x staple over the 48 staples S (U_mu(x) S a closed
// loop) -- per nu != mu and sign s the plaquette [s nu, -mu, -s nu], the
// rectangles [mu, s nu, -mu, -mu, -s nu], [s nu, -mu, -mu, -s nu, mu],
// [s nu, s nu, -mu, -s nu, -s nu], per ordered pair nu != rho and signs the
// parallelogram [s nu, r rho, -mu, -s nu, -r rho]: every loop of
// imp_gauge_force_cpu (generic/gauge_force_imp.c:37-219) through the link,
// once -- then mom += -eb3 TA(U_mu(x) G).  U: links without KS phases.
//
// Block order: workgroups are handed to the 8 XCDs in turn, so block b runs on
// XCD b % 8; each XCD takes one contiguous eighth of the lattice (cbb_per
// blocks of 128 sites) and, for each block of sites, all eight (parity,
// direction) pairs in a row, so the links the staples share stay in its L2.
__global__ void __launch_bounds__(128) k_gauge_staples(Geom g, LinkField U, GaugeStaples st, double c0, double c1,
                                                       double c2, double eb3, double* __restrict__ mom,
                                                       int cbb_per) {
  const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
  const int pm = j & 7, cbb = xcd * cbb_per + (j >> 3);
  int cb = cbb * 128 + threadIdx.x;
  if (cb >= g.Vh) return;
  const int pb = pm >> 2, mu = pm & 3;
  Coord c = site_coords(g, cb, pb);
  int x1[4] = {c.x, c.y, c.z, c.t};
  x1[mu] = wrap(x1[mu] + 1, g.X[mu]);
  Mat3<double> G, acc, L, T, Ln;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) G.e[i][j].re = G.e[i][j].im = 0.0;
  // link k of a staple: its (parity, cb, direction) from the walk position,
  // which it advances
  auto link_of = [&](int code, int k, int (&pos)[4], int& lpb, int& lcb, int& ld, bool& back) {
    const int d = (code >> (3 * k)) & 7;
    back = d >= 4;
    ld = back ? 7 - d : d;
    if (!back) lcb = cb_of(g, pos, lpb);
#pragma unroll
    for (int a = 0; a < 4; a++)
      if (a == ld) pos[a] = wrap(pos[a] + (back ? -1 : 1), g.X[a]);
    if (back) lcb = cb_of(g, pos, lpb);
  };
#pragma unroll 1
  for (int q = 0; q < 48; q++) {
    const int code = st.code[mu][q], n = st.len[mu][q], cls = st.cls[mu][q];
    const double w = cls == 0 ? c0 : (cls == 1 ? c1 : c2);
    int pos[4] = {x1[0], x1[1], x1[2], x1[3]};
    int lpb, lcb, ld;
    bool back;
    link_of(code, 0, pos, lpb, lcb, ld, back);
    mload(U, lpb, ld, lcb, acc);
    if (back) madj(acc);
    link_of(code, 1, pos, lpb, lcb, ld, back);
    mload(U, lpb, ld, lcb, L);
    bool lback = back;
    // the next link is requested before the product with the current one, so
    // two loads of a staple are in flight (104 -> 87 ms per call at 48^3x96)
#pragma unroll 1
    for (int k = 1; k < n; k++) {
      bool nback = false;
      if (k + 1 < n) {
        link_of(code, k + 1, pos, lpb, lcb, ld, nback);
        mload(U, lpb, ld, lcb, Ln);
      }
      if (lback) madj(L);
      mmul(acc, L, T);
      acc = T;
      L = Ln;
      lback = nback;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        G.e[i][j].re += w * acc.e[i][j].re;
        G.e[i][j].im += w * acc.e[i][j].im;
      }
  }
  Mat3<double> X;
  mload(U, pb, mu, cb, L);
  mmul(L, G, X);
  // mom += -eb3 TA(X) (make_anti_hermitian, libraries/make_ahmat.c)
  double tr = (X.e[0][0].im + X.e[1][1].im + X.e[2][2].im) * (1.0 / 3.0);
  const double sc = -eb3, h = 0.5 * sc;
  double* o = mom + ((size_t)(pb * g.Vh + cb) * 4 + mu) * 10;
  o[0] += h * (X.e[0][1].re - X.e[1][0].re);
  o[1] += h * (X.e[0][1].im + X.e[1][0].im);
  o[2] += h * (X.e[0][2].re - X.e[2][0].re);
  o[3] += h * (X.e[0][2].im + X.e[2][0].im);
  o[4] += h * (X.e[1][2].re - X.e[2][1].re);
  o[5] += h * (X.e[1][2].im + X.e[2][1].im);
  o[6] += sc * (X.e[0][0].im - tr);
  o[7] += sc * (X.e[1][1].im - tr);
  o[8] += sc * (X.e[2][2].im - tr);
  o[9] = 0.0;
}

// mom_dev (packed, node_index order) += -eb3 TA(U G), G = sum over the loops
// through each link of coeff (rest of the loop from the link's end) --
// imp_gauge_force_cpu, generic/gauge_force_imp.c:37-219.  U: links without KS
// phases.
int gauge_force_dev(const LinkField& U, const double coeff[3], double eb3, double* mom_dev) {
  Context& c = ctx();
  static const GaugeStaples st = make_gauge_staples();
  const int nbx = (c.Vh + 127) / 128, cbb_per = (nbx + 7) / 8;
  hipLaunchKernelGGL(k_gauge_staples, dim3(64 * cbb_per), dim3(128), 0, c.stream, c.geom, U, st, coeff[0], coeff[1],
                     coeff[2], eb3, mom_dev, cbb_per);
  MH_CHECK(hipGetLastError());
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

// Per-block partial sums of coeff (3 - Re Tr P) over the loops based at each
// site (imp_gauge_action, generic/gauge_stuff.c:248-322)
__global__ void __launch_bounds__(128) k_gauge_action(Geom g, LinkField U, const FnPath* __restrict__ paths,
                                                      int npaths, double* __restrict__ partials) {
  const int nbx = (g.Vh + blockDim.x - 1) / blockDim.x;
  int bx, pb;
  if (!xcd_block(nbx, 2, bx, pb)) return;
  int cb = bx * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (cb < g.Vh) {
    Coord c = site_coords(g, cb, pb);
    for (int q = 0; q < npaths; q++) {
      const FnPath* P = paths + q;
      int pos[4] = {c.x, c.y, c.z, c.t};
      Mat3<double> acc, L, T;
      munit(acc);
      for (int k = 0; k < P->n; k++) {
        int d = P->d[k];
        bool back = d >= 4;
        int ld = back ? 7 - d : d;
        int lpb, lcb;
        if (!back) lcb = cb_of(g, pos, lpb);
        pos[ld] = wrap(pos[ld] + (back ? -1 : 1), g.X[ld]);
        if (back) lcb = cb_of(g, pos, lpb);
        mload(U, lpb, ld, lcb, L);
        if (back) madj(L);
        mmul(acc, L, T);
        acc = T;
      }
      s += P->c * (3.0 - (acc.e[0][0].re + acc.e[1][1].re + acc.e[2][2].re));
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double sm[2];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[pb * nbx + bx] = sm[0] + sm[1];
}

int gauge_action_dev(const LinkField& U, const double coeff[3], double* action) {
  Context& c = ctx();
  GaugeTable* t;
  int rc;
  if ((rc = gauge_table(coeff, t))) return rc;
  const int nb = (c.Vh + 127) / 128;
  double* part = nullptr;
  if ((rc = dev_alloc((void**)&part, (size_t)2 * nb * sizeof(double)))) return rc;
  hipLaunchKernelGGL(k_gauge_action, xcd_grid(nb, 2), dim3(128), 0, c.stream, c.geom, U, t->paths, t->npaths, part);
  std::vector<double> h((size_t)2 * nb);
  rc = hipGetLastError() == hipSuccess ? 0 : -1;
  if (!rc) rc = hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream) ==
                        hipSuccess ? 0 : -1;
  (void)hipStreamSynchronize(c.stream);
  dev_free(part);
  double sum = 0.0;
  for (double v : h) sum += v;
  *action = sum;
  return rc;
}

// packed anti_hermitmat (m01, m02, m12, m00im, m11im, m22im) -> matrix
__device__ __forceinline__ void unpack_ah(const double* o, Mat3<double>& P) {
  P.e[0][1].re = o[0];  P.e[0][1].im = o[1];
  P.e[0][2].re = o[2];  P.e[0][2].im = o[3];
  P.e[1][2].re = o[4];  P.e[1][2].im = o[5];
  P.e[1][0].re = -o[0]; P.e[1][0].im = o[1];
  P.e[2][0].re = -o[2]; P.e[2][0].im = o[3];
  P.e[2][1].re = -o[4]; P.e[2][1].im = o[5];
  P.e[0][0].re = 0.0;   P.e[0][0].im = o[6];
  P.e[1][1].re = 0.0;   P.e[1][1].im = o[7];
  P.e[2][2].re = 0.0;   P.e[2][2].im = o[8];
}

// U <- sum_{k=0}^{8} (eps P)^k / k! U, Horner form (update_u, ks_imp_rhmc/update_u.c:66-111)
__global__ void __launch_bounds__(128) k_update_u(Geom g, LinkField U, const double* __restrict__ mom, double eps) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Mat3<double> P, L, t1, t2;
  unpack_ah(mom + ((size_t)(pb * g.Vh + cb) * 4 + mu) * 10, P);
  mload(U, pb, mu, cb, L);
  mmul(P, L, t1);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      t2.e[i][j].re = L.e[i][j].re + (eps / 8.0) * t1.e[i][j].re;
      t2.e[i][j].im = L.e[i][j].im + (eps / 8.0) * t1.e[i][j].im;
    }
  for (int m = 7; m >= 1; m--) {
    mmul(P, t2, t1);
    const double f = eps / m;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        t2.e[i][j].re = L.e[i][j].re + f * t1.e[i][j].re;
        t2.e[i][j].im = L.e[i][j].im + f * t1.e[i][j].im;
      }
  }
  double2* out = (double2*)U.data + (size_t)((pb * 4 + mu) * 9) * U.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) out[(size_t)e * U.stride] = make_double2(t2.e[e / 3][e % 3].re, t2.e[e / 3][e % 3].im);
}

int update_u_dev(LinkField& U, const double* mom_dev, double eps) {
  Context& c = ctx();
  hipLaunchKernelGGL(k_update_u, dim3((c.Vh + 127) / 128, 8), dim3(128), 0, c.stream, c.geom, U, mom_dev, eps);
  MH_CHECK(hipGetLastError());
  return 0;
}

// reunit_su3 (generic/reunitarize2.c:133-205) on every link, with the KS phase
// taken out first when the links carry it (reunitarize_ks,
// generic_ks/reunitarize_ks.c:9-29): row 0 normalised, row 0's component taken
// out of row 1, row 1 normalised, row 2 = conj(row0 x row1).  The four
// deviations MILC checks (|row0|^2 - 1, |<row0,row1>|^2, |row1'|^2 - 1,
// |row2 change|^2) are counted against tol and their maximum kept.
__global__ void __launch_bounds__(128) k_reunitarize(Geom g, LinkField U, int phases_in, double tol,
                                                     unsigned int* __restrict__ errors,
                                                     unsigned long long* __restrict__ maxdev_bits) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= g.Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Mat3<double> M;
  mload(U, pb, mu, cb, M);
  double ph = phases_in ? ks_phase(g, site_coords(g, cb, pb), mu) : 1.0;
  double n0 = 0, n1 = 0, dev[4];
#pragma unroll
  for (int k = 0; k < 3; k++) n0 += ph * ph * (M.e[0][k].re * M.e[0][k].re + M.e[0][k].im * M.e[0][k].im);
  dev[0] = fabs(n0 - 1.0);
  const double s0 = ph / sqrt(n0);
  Cplx<double> r0[3], r1[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r0[k].re = s0 * M.e[0][k].re;
    r0[k].im = s0 * M.e[0][k].im;
    r1[k].re = ph * M.e[1][k].re;
    r1[k].im = ph * M.e[1][k].im;
  }
  double ore = 0, oim = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    ore += r0[k].re * r1[k].re + r0[k].im * r1[k].im;
    oim += r0[k].re * r1[k].im - r0[k].im * r1[k].re;
  }
  dev[1] = ore * ore + oim * oim;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r1[k].re -= ore * r0[k].re - oim * r0[k].im;
    r1[k].im -= ore * r0[k].im + oim * r0[k].re;
    n1 += r1[k].re * r1[k].re + r1[k].im * r1[k].im;
  }
  dev[2] = fabs(n1 - 1.0);
  const double s1 = 1.0 / sqrt(n1);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r1[k].re *= s1;
    r1[k].im *= s1;
  }
  Cplx<double> r2[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int a = (k + 1) % 3, b = (k + 2) % 3;
    // conj(r0[a] r1[b] - r0[b] r1[a])
    double re = r0[a].re * r1[b].re - r0[a].im * r1[b].im - (r0[b].re * r1[a].re - r0[b].im * r1[a].im);
    double im = r0[a].re * r1[b].im + r0[a].im * r1[b].re - (r0[b].re * r1[a].im + r0[b].im * r1[a].re);
    r2[k].re = re;
    r2[k].im = -im;
  }
  dev[3] = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double dr = ph * M.e[2][k].re - r2[k].re, di = ph * M.e[2][k].im - r2[k].im;
    dev[3] += dr * dr + di * di;
  }
  unsigned int nerr = 0;
  double dmax = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    nerr += dev[k] >= tol;
    dmax = fmax(dmax, dev[k]);
  }
  if (nerr) atomicAdd(errors, nerr);
  // non-negative doubles order as their bit patterns
  atomicMax(maxdev_bits, (unsigned long long)__double_as_longlong(dmax));
  double2* out = (double2*)U.data + (size_t)((pb * 4 + mu) * 9) * U.stride + cb;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    out[(size_t)k * U.stride] = make_double2(ph * r0[k].re, ph * r0[k].im);
    out[(size_t)(3 + k) * U.stride] = make_double2(ph * r1[k].re, ph * r1[k].im);
    out[(size_t)(6 + k) * U.stride] = make_double2(ph * r2[k].re, ph * r2[k].im);
  }
}

// the two flag words: error count and the bits of the largest deviation
__global__ void k_reunit_clear(unsigned long long* flags) {
  flags[0] = 0;
  flags[1] = 0;
}

int reunitarize_dev(LinkField& U, int phases_in, double tol, int* errors, double* maxdev) {
  Context& c = ctx();
  unsigned long long* flags = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&flags, 2 * sizeof(unsigned long long)))) return rc;
  unsigned int* err = (unsigned int*)flags;
  hipLaunchKernelGGL(k_reunit_clear, dim3(1), dim3(1), 0, c.stream, flags);
  hipLaunchKernelGGL(k_reunitarize, dim3((c.Vh + 127) / 128, 8), dim3(128), 0, c.stream, c.geom, U, phases_in, tol,
                     err, flags + 1);
  unsigned long long h[2] = {0, 0};
  rc = hipGetLastError() == hipSuccess ? 0 : -1;
  if (!rc)
    rc = hipMemcpyAsync(h, flags, sizeof h, hipMemcpyDeviceToHost, c.stream) == hipSuccess ? 0 : -1;
  (void)hipStreamSynchronize(c.stream);
  dev_free(flags);
  if (errors) *errors = (int)(unsigned int)(h[0] & 0xffffffffu);
  if (maxdev) {
    double d;
    memcpy(&d, &h[1], sizeof d);
    *maxdev = d;
  }
  return rc;
}

// per-block partial sums of ahmat_mag_sq - 4 (hmom_action, ks_imp_rhmc/d_action_rhmc.c:86-124)
__global__ void __launch_bounds__(256) k_mom_action(const double* __restrict__ mom, size_t nlinks,
                                                    double* __restrict__ partials) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nlinks; i += (size_t)gridDim.x * blockDim.x) {
    const double* o = mom + 10 * i;
    s += 0.5 * (o[6] * o[6] + o[7] * o[7] + o[8] * o[8]) + o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3] +
         o[4] * o[4] + o[5] * o[5] - 4.0;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

int mom_action_dev(const double* mom_dev, double* action) {
  Context& c = ctx();
  const int nb = 1024;
  double* part = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&part, nb * sizeof(double)))) return rc;
  hipLaunchKernelGGL(k_mom_action, dim3(nb), dim3(256), 0, c.stream, mom_dev, (size_t)2 * c.Vh * 4, part);
  std::vector<double> h(nb);
  rc = hipGetLastError() == hipSuccess ? 0 : -1;
  if (!rc) rc = hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream) == hipSuccess
                    ? 0 : -1;
  (void)hipStreamSynchronize(c.stream);
  dev_free(part);
  double sum = 0.0;
  for (double v : h) sum += v;
  *action = sum;
  return rc;
}

// Staple-form adjoint on the device: G (FP64 link field, zeroed here) = force on
// the links L of the smearing with path coefficients coeff, from the force
// F1 (non-Naik) / F3 (Naik) on its result (eps(x) applied when with_eps).
//
// Per direction mu, with F = eps F1_mu:
//   S1[nu] = S_nu(L_mu)                      (3-staples)
//   G2[sig] = c5 F + c7 FA_sig(F)            (the 7-staple force seen by a 5-staple)
//   G[sig] += c7 FU_sig(S_rho(S1[nu]) + S_nu(S1[rho]), F)   ({nu, rho, sig} the other three)
//   per nu: G1 = c3 F + cl FA_nu(F)  [Lepage] + sum_rho FA_rho(G2[sig]) (or c5 FA_rho(F)),
//           G[nu] += cl FU_nu(S1[nu], F), G[rho] += FU_rho(S1[nu], G2[sig]),
//           G_mu += FA_nu(G1), G[nu] += FU_nu(L_mu, G1)
// FA_nu(X) and FU_nu(A, X) are the two halves of the reverse of S_nu (k_sf_adj).
// A 7-staple force term depends on the 5-staple only through FU, linearly,
// and its FA part only on F and sig, so the two orders (nu, rho) and
// (rho, nu) of one sig share one FA pass and one FU pass.
static int smear_adjoint_dev(const LinkField& L, const LinkField& F1, const LinkField& F3, int with_eps,
                             const double coeff[6], LinkField& G) {
  Context& c = ctx();
  const int st = c.stride;
  const double c1 = coeff[0] - 6.0 * coeff[5], cN = coeff[1], c3 = coeff[2], c5 = coeff[3], c7 = coeff[4],
               cl = coeff[5];
  const bool staples = c3 != 0.0 || cl != 0.0 || c5 != 0.0 || c7 != 0.0;
  const bool five = c5 != 0.0 || c7 != 0.0, seven = c7 != 0.0;
  // F, G1, S1[4], G2[4], SS: the mu-th entries of S1 and G2 are not used
  double2 *F = nullptr, *G1 = nullptr, *SS = nullptr, *S1[4] = {}, *G2[4] = {};
  OnExit bufs([&] {
    dev_free(F);
    dev_free(G1);
    dev_free(SS);
    for (int d = 0; d < 4; d++) {
      dev_free(S1[d]);
      dev_free(G2[d]);
    }
  });
  const size_t bytes = (size_t)2 * 9 * st * sizeof(double2);
  int rc;
  if ((rc = dev_alloc((void**)&F, bytes))) return rc;
  if (staples) {
    if ((rc = dev_alloc((void**)&G1, bytes))) return rc;
    for (int d = 0; d < 3; d++)
      if ((rc = dev_alloc((void**)&S1[d], bytes))) return rc;
    if (seven) {
      if ((rc = dev_alloc((void**)&SS, bytes))) return rc;
      for (int d = 0; d < 3; d++)
        if ((rc = dev_alloc((void**)&G2[d], bytes))) return rc;
    }
  }
  MH_CHECK(hipMemsetAsync(G.data, 0, link_bytes(P_F64, st), c.stream));
  const MF mF = mf_site(F, st), mG1 = mf_site(G1, st), mSS = mf_site(SS, st);
  // init: plain grid; the neighbour-reading kernels: XCD block order
  dim3 grid((c.Vh + 127) / 128, 2), xgrid = xcd_grid((c.Vh + 127) / 128, 2), blk(128);
  const MF none{};
  auto adj = [&](int mode, int mu, int nu, const MF& A, const MF& Gs, double cg, const MF& FA, const MF& FU, int init,
                 double ci) {
    const SfKernels k = sf_kernels(c.geom, mu, nu);
    hipLaunchKernelGGL(k.adj[mode], xgrid, blk, 0, c.stream, c.geom, L, A, Gs, cg, FA, FU, init, mF, ci);
  };
  auto fwd = [&](int mu, int nu, const MF& A, const MF& out) {
    const SfKernels k = sf_kernels(c.geom, mu, nu);
    hipLaunchKernelGGL(k.fwd, xgrid, blk, 0, c.stream, c.geom, L, A, out);
  };
  for (int mu = 0; mu < 4; mu++) {
    const MF Gmu = mf_link(G, mu);
    hipLaunchKernelGGL(k_sf_init, grid, blk, 0, c.stream, c.geom, mf_link(F1, mu), with_eps, mF, Gmu, c1);
    if (staples) {
      // the three other directions and their buffers
      int od[3], k = 0;
      for (int d = 0; d < 4; d++)
        if (d != mu) od[k++] = d;
      MF mS1[4], mG2[4];
      for (int i = 0; i < 3; i++) {
        mS1[od[i]] = mf_site(S1[i], st);
        mG2[od[i]] = mf_site(G2[i], st);
      }
      for (int i = 0; i < 3; i++) fwd(mu, od[i], mf_link(L, mu), mS1[od[i]]);
      if (seven)
        for (int i = 0; i < 3; i++) {
          const int sig = od[i], nu = od[(i + 1) % 3], rho = od[(i + 2) % 3];
          adj(1, mu, sig, none, mF, c7, mG2[sig], none, 1, c5);
          const int p = std::min(nu, rho), q = std::max(nu, rho);
          hipLaunchKernelGGL(sf_fwd2_kernel(c.geom, mu, p, q), xgrid, blk, 0, c.stream, c.geom, L, mS1[q], mS1[p],
                             mSS);
          adj(2, mu, sig, mSS, mF, c7, none, mf_link(G, sig), 0, 0.0);
        }
      for (int i = 0; i < 3; i++) {
        const int nu = od[i];
        bool g1_set = false;
        if (cl != 0.0) {
          adj(0, mu, nu, mS1[nu], mF, cl, mG1, mf_link(G, nu), 1, c3);
          g1_set = true;
        }
        if (five)
          for (int j = 1; j < 3; j++) {
            const int rho = od[(i + j) % 3], sig = od[(i + 3 - j) % 3];
            adj(0, mu, rho, mS1[nu], seven ? mG2[sig] : mF, seven ? 1.0 : c5, mG1, mf_link(G, rho), g1_set ? 0 : 1,
                c3);
            g1_set = true;
          }
        adj(0, mu, nu, mf_link(L, mu), g1_set ? mG1 : mF, g1_set ? 1.0 : c3, Gmu, mf_link(G, nu), 0, 0.0);
      }
    }
    if (cN != 0.0)
      hipLaunchKernelGGL(mu == 0 ? k_sf_naik<0> : mu == 1 ? k_sf_naik<1> : mu == 2 ? k_sf_naik<2> : k_sf_naik<3>, xgrid,
                         blk, 0, c.stream, c.geom, L, mf_link(F3, mu), with_eps, cN, Gmu);
  }
  MH_CHECK(hipGetLastError());
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

// HISQ thin-link force from the level-0 outer products (the chain of
// fn_fermion_force_multi_hisq_wrapper_mx_cpu, fermion_force_hisq_multi.c:1182-1475):
// F_W = level-2 adjoint on W (coeff2, eps(x) O1 / O3), F_V = U(3) projection
// adjoint at V, then the fat7 adjoint on U (coeff1), mom += TA(scale U F_U).
// y = b y + a s x on FP64 link fields, s = +1 on even and -1 on odd sites
// when signed_odd (the parity factor of the force's outer products)
__global__ void __launch_bounds__(256) k_link_axpy(LinkField y, LinkField x, double a, double b, int signed_odd,
                                                   int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pb = blockIdx.y >> 2, mu = blockIdx.y & 3;
  if (cb >= Vh) return;
  Mat3<double> my, mx;
  load_link<P_F64>(y, pb, mu, cb, my);
  load_link<P_F64>(x, pb, mu, cb, mx);
  const double as = (signed_odd && pb == 1) ? -a : a;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    my.e[e / 3][e % 3].re = b * my.e[e / 3][e % 3].re + as * mx.e[e / 3][e % 3].re;
    my.e[e / 3][e % 3].im = b * my.e[e / 3][e % 3].im + as * mx.e[e / 3][e % 3].im;
  }
  store_link<P_F64, double>(y, pb, mu, cb, my, 1.0f);
}

int link_axpy_dev(LinkField& y, const LinkField& x, double a, double b, int signed_odd) {
  Context& c = ctx();
  hipLaunchKernelGGL(k_link_axpy, dim3((c.Vh + 255) / 256, 8), dim3(256), 0, c.stream, y, x, a, b, signed_odd,
                     c.Vh);
  MH_CHECK(hipGetLastError());
  return 0;
}

// one_link_extra: outer products of terms whose links carry a second Naik set,
// already weighted by eps_naik times that set's one-link coefficient; their
// one-link path adds them (with the parity factor) to the force on W
int hisq_force_from_oprod(const LinkField& U, const LinkField& V, const LinkField& W, const LinkField& o1,
                          const LinkField& o3, const double coeff2[6], const double coeff1[6], double scale,
                          double filter, double* mom_dev, const LinkField* one_link_extra) {
  Context& c = ctx();
  LinkField FW{}, FV{};
  int rc;
  if ((rc = alloc_links(FW, P_F64, c.stride)) || (rc = alloc_links(FV, P_F64, c.stride))) {
    free_links(FW);
    return rc;
  }
  LinkField none{nullptr, 1.0f, 0};
  if (use_staple_form(coeff2, true)) rc = smear_adjoint_dev(W, o1, o3, 1, coeff2, FW);
  else rc = path_adjoint_launch(0, W, o1, o3, 1, coeff2, FW, none, 0.0, nullptr);
  if (!rc && one_link_extra) rc = link_axpy_dev(FW, *one_link_extra, 1.0, 1.0, 1);
  if (!rc) {
    hipLaunchKernelGGL(k_u3_adjoint, dim3((c.Vh + 127) / 128, 8), dim3(128), 0, c.stream, c.geom, V, FW, FV, filter);
    rc = hipGetLastError() == hipSuccess ? 0 : -1;
  }
  if (!rc) {
    if (use_staple_form(coeff1, false)) {
      rc = smear_adjoint_dev(U, FV, FV, 0, coeff1, FW);  // FW reused as F_U
      if (!rc) {
        hipLaunchKernelGGL(k_mom_ta, dim3((c.Vh + 127) / 128, 8), dim3(128), 0, c.stream, c.geom, U, FW, scale,
                           mom_dev, 0);
        rc = hipGetLastError() == hipSuccess ? 0 : -1;
      }
    } else {
      rc = path_adjoint_launch(1, U, FV, FV, 0, coeff1, none, U, scale, mom_dev);
    }
  }
  MH_CHECK(hipStreamSynchronize(c.stream));
  free_links(FW);
  free_links(FV);
  return rc;
}

}  // namespace mh
