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

// mom(y, nu) += TA(scale U_nu(y) G_nu(y)), written as MILC's packed
// anti_hermitmat (m01, m02, m12 complex; m00im, m11im, m22im; space) at
// MILC index 4 i + nu, i = node_index (even sites first).
__global__ void __launch_bounds__(128) k_fn_force(Geom g, LinkField U, LinkField O1, LinkField O3,
                                                  const FnPath* __restrict__ paths, const FnEntry* __restrict__ ent,
                                                  int4 e_lo, int4 e_hi, double scale, double* __restrict__ mom) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
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
      mload(U, lpb, ld, lcb, L);
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
    mload(P->naik ? O3 : O1, xpb, E->mu, xcb, Ob);
    double w = (xpb ? -1.0 : 1.0) * P->c;
    mmul(suf, Ob, A);
    mmul(A, pre, M);
    if (E->adj) madj(M);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        G.e[i][j].re += w * M.e[i][j].re;
        G.e[i][j].im += w * M.e[i][j].im;
      }
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

// Path table of an action, built once per coefficient set and kept on the device.
struct FnTable {
  double coeff[6] = {0, 0, 0, 0, 0, 0};
  FnPath* paths = nullptr;
  FnEntry* ent = nullptr;
  int start[5] = {0, 0, 0, 0, 0};
  bool valid = false;
};
static FnTable g_table;

static int fn_table(const double coeff[6], FnTable*& out) {
  Context& c = ctx();
  bool same = g_table.valid;
  for (int i = 0; i < 6 && same; i++) same = g_table.coeff[i] == coeff[i];
  if (!same) {
    if (g_table.valid) {
      dev_free(g_table.paths);
      dev_free(g_table.ent);
      g_table.valid = false;
    }
    std::vector<FnPath> paths;
    std::vector<int> path_mu;
    std::vector<FnEntry> ent;
    build_paths(coeff, paths, path_mu);
    build_entries(paths, path_mu, ent, g_table.start);
    int rc;
    if ((rc = dev_alloc((void**)&g_table.paths, paths.size() * sizeof(FnPath))) ||
        (rc = dev_alloc((void**)&g_table.ent, ent.size() * sizeof(FnEntry))))
      return rc;
    MH_CHECK(hipMemcpyAsync(g_table.paths, paths.data(), paths.size() * sizeof(FnPath), hipMemcpyHostToDevice,
                            c.stream));
    MH_CHECK(hipMemcpyAsync(g_table.ent, ent.data(), ent.size() * sizeof(FnEntry), hipMemcpyHostToDevice, c.stream));
    MH_CHECK(hipStreamSynchronize(c.stream));
    for (int i = 0; i < 6; i++) g_table.coeff[i] = coeff[i];
    g_table.valid = true;
  }
  out = &g_table;
  return 0;
}

void fn_table_release() {
  if (g_table.valid) {
    dev_free(g_table.paths);
    dev_free(g_table.ent);
  }
  g_table = FnTable{};
}

// Packed momentum increments (MILC AoS, 10 doubles per link) into mom_dev from
// the outer products O1/O3 (FP64), thin phased links U (FP64) and the path
// coefficients; scale multiplies U G (2 eps in fermion_force_fn_multi, 1 when
// the outer products already carry 2 eps, fermion_force_asqtad_gpu.c:53-56).
int fn_force_from_oprod(const LinkField& U, const LinkField& o1, const LinkField& o3, const double coeff[6],
                        double scale, double* mom_dev) {
  Context& c = ctx();
  FnTable* t;
  int rc = fn_table(coeff, t);
  if (rc) return rc;
  int4 lo = make_int4(t->start[0], t->start[1], t->start[2], t->start[3]);
  int4 hi = make_int4(t->start[1], t->start[2], t->start[3], t->start[4]);
  hipLaunchKernelGGL(k_fn_force, dim3((c.Vh + 127) / 128, 8), dim3(128), 0, c.stream, c.geom, U, o1, o3, t->paths,
                     t->ent, lo, hi, scale, mom_dev);
  MH_CHECK(hipGetLastError());
  return 0;
}

}  // namespace mh
