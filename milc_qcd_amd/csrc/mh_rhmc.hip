// mh_rhmc.hip -- RHMC compositions of the solver, stencil and force kernels,
// all on device-resident fields.
//
// Restates (MILC 7.8.1):
//   ks_ratinv / ks_rateval   ks_imp_rhmc/ks_ratinv.c:33-138
//   fermion_action           ks_imp_rhmc/d_action_rhmc.c (sum over even sites of
//                            |a0 phi + sum_k a_k x_k|^2, FP64)
//   update_h_fermion         ks_imp_rhmc/update_h_rhmc.c:77-107: multi_x even =
//                            X_t, odd = D_oe X_t, then the force (level 0 here)
//   update (one trajectory)  ks_imp_rhmc/update_rhmc.c:157-345: ranmom,
//                            grsource_imp_rhmc, d_action_rhmc, the leapfrog
//                            and Omelyan integrators over update_u /
//                            update_h_rhmc, d_action_rhmc, Metropolis
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "mh_batch.h"
#include "mh_internal.h"

namespace mh {

int force_oprod_coeff(const std::vector<VecField>& xe, const std::vector<VecField>& xo, const std::vector<double>& c1,
                      const std::vector<double>& c3, LinkField& o1, LinkField& o3);

// dst = a0 * src + sum_k a_k x_k  (FP64)
#define RAT_MAX 32
struct RatSet {
  VecField x[RAT_MAX];
  double a[RAT_MAX];
};
__global__ void k_rateval(VecField src, double a0, RatSet rs, int n, VecField dst, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<double> s, acc;
  load_vec<P_F64>(src, cb, s);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    acc.c[k].re = a0 * s.c[k].re;
    acc.c[k].im = a0 * s.c[k].im;
  }
  for (int t = 0; t < n; t++) {
    Vec3<double> x;
    load_vec<P_F64>(rs.x[t], cb, x);
    double a = rs.a[t];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      acc.c[k].re += a * x.c[k].re;
      acc.c[k].im += a * x.c[k].im;
    }
  }
  store_vec<P_F64>(dst, cb, acc);
}

// partials of Re<a, b> (FP64)
__global__ void k_redot(VecField a, VecField b, int Vh, double* __restrict__ partials) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0;
  if (cb < Vh) {
    Vec3<double> x, y;
    load_vec<P_F64>(a, cb, x);
    load_vec<P_F64>(b, cb, y);
#pragma unroll
    for (int k = 0; k < 3; k++) s += x.c[k].re * y.c[k].re + x.c[k].im * y.c[k].im;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double sm[16];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sm[w];
    partials[blockIdx.x * 4] = t;
  }
}

static int redot(const VecField& a, const VecField& b, double* out) {
  Context& c = ctx();
  int nb = (c.Vh + 255) / 256;
  hipLaunchKernelGGL(k_redot, dim3(nb), dim3(256), 0, c.stream, a, b, c.Vh, c.ws.partials);
  MH_CHECK(hipGetLastError());
  std::vector<double> h((size_t)nb * 4);
  MH_CHECK(hipMemcpyAsync(h.data(), c.ws.partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double s = 0;
  for (int i = 0; i < nb; i++) s += h[(size_t)i * 4];
  *out = s;
  return 0;
}

// Multi-shift solve for the rational poles on parity `parity` (ks_ratinv):
// FP64 throughout (mixed = 0, MILC's CPU semantics) or FP32 bulk + per-shift
// mixed refinement (mixed = 1).
static int rational_solve(const VecField& phi, int n, const double* offsets, double resid, int mixed, int parity,
                          std::vector<VecField>& x, SolveStats& st) {
  Context& c = ctx();
  int rc;
  x.assign(n, VecField{});
  for (int t = 0; t < n; t++)
    if ((rc = alloc_vec(x[t], P_F64, c.stride))) return rc;
  std::vector<double> offs(offsets, offsets + n);
  MultiParams mp;
  mp.parity = parity;
  mp.max_iters = 20000;
  mp.resid = mixed ? std::max(resid, kFp32StageResid) : resid;
  mp.prec = mixed ? P_F32 : P_F64;
  mp.batch = 16;
  int bulk = mp.prec;
  if (!c.links[bulk].valid) return -3;
  if ((rc = multicg_solve(phi, x, offs, mp, st))) return rc;
  if (!mixed) return 0;
  // Refinement of every shift to the FP64 true-residual target from the
  // multi-shift solution: 16-bit sloppy operator, FP64 reliable updates, up to
  // four shifts per batched CG.
  std::vector<double> res(n, resid);
  long it = 0;
  if ((rc = refine_shifts(phi, x, offs.data(), res.data(), parity, 20000, P_H16, &it, nullptr, &st.converged)))
    return rc;
  st.iters += (int)it;
  return 0;
}

// Several sources' rational solves at once (every pseudofermion's MD, GR or
// FA rational): the batched multi-shift with a shift set per source, FP32
// iterations sharing every stencil launch, then every shift of every source
// refined to its FP64 true-residual target; x[s][t] allocated here.
static int rational_solve_all(const std::vector<const VecField*>& src, const std::vector<std::vector<double>>& offs,
                              const std::vector<double>& resid, std::vector<std::vector<VecField>>& x, long* iters) {
  const int ns = (int)src.size();
  if (ns < 1 || ns > BNR_MAX || (int)offs.size() != ns || (int)resid.size() != ns) return -2;
  const int nr = ns <= 4 ? 4 : (ns <= 8 ? 8 : 16);
  int n = 0;
  for (auto& o : offs) {
    if (o.empty() || (int)o.size() > BMS_MAX) return -2;
    n = std::max(n, (int)o.size());
  }
  // a source with fewer shifts repeats its last one; the padding sources are
  // zero and halt at once
  std::vector<double> flat((size_t)n * nr), stage(nr, 1.0), refine(nr, 1.0);
  for (int j = 0; j < n; j++)
    for (int r = 0; r < nr; r++) {
      const std::vector<double>& o = offs[r < ns ? r : 0];
      flat[(size_t)j * nr + r] = o[std::min(j, (int)o.size() - 1)];
    }
  for (int r = 0; r < ns; r++) {
    stage[r] = std::max(resid[r], kFp32StageResid);
    refine[r] = resid[r];
  }
  void* b = nullptr;
  std::vector<void*> xb(n, nullptr);
  auto release = [&]() {
    dev_free(b);
    for (void* p : xb) dev_free(p);
  };
  int rc = balloc(&b, P_F64, nr);
  for (int j = 0; j < n && !rc; j++) rc = balloc(&xb[j], P_F64, nr);
  for (int r = 0; r < ns && !rc; r++) rc = bpack(*src[r], b, P_F64, nr, r);
  BMSParams prm;
  prm.parity = EVEN;
  prm.max_iters = 20000;
  prm.resid = 0;
  prm.refine_resid = 0;
  prm.delta = kBatchReliableDelta;
  BMSStats st;
  if (!rc) rc = bmultishift_sets(b, nr, n, flat.data(), stage.data(), refine.data(), prm, xb, st);
  x.assign(ns, {});
  for (int r = 0; r < ns && !rc; r++) {
    x[r].assign(offs[r].size(), VecField{});
    for (int t = 0; t < (int)offs[r].size() && !rc; t++) {
      rc = alloc_vec(x[r][t], P_F64, ctx().stride);
      if (!rc) rc = bunpack(xb[t], P_F64, nr, r, x[r][t]);
    }
  }
  (void)hipStreamSynchronize(ctx().stream);
  release();
  if (iters) {
    long it = st.ms_iters;
    for (int j = 0; j < n; j++) it += st.refine_iters[j];
    *iters = it;
  }
  return rc;
}

// dst = res[0] src + sum_t res[t+1] x[t] (ks_rateval)
static int rateval(const VecField& src, const std::vector<VecField>& x, const double* res, VecField& dst) {
  Context& c = ctx();
  if (x.size() > RAT_MAX) return -2;
  RatSet rs{};
  for (size_t t = 0; t < x.size(); t++) {
    rs.x[t] = x[t];
    rs.a[t] = res[t + 1];
  }
  hipLaunchKernelGGL(k_rateval, dim3((c.Vh + 255) / 256), dim3(256), 0, c.stream, src, res[0], rs, (int)x.size(), dst,
                     c.Vh);
  MH_CHECK(hipGetLastError());
  return 0;
}

// dst = res[0] src + sum_k res[k] (M^dag M + off[k])^-1 src on even sites (ks_rateval)
static int rational_apply(const VecField& src, int order, const double* offs, const double* res, double resid,
                          int mixed, VecField& dst, int* iters) {
  Context& c = ctx();
  std::vector<VecField> x;
  SolveStats st;
  int rc = rational_solve(src, order, offs, resid, mixed, EVEN, x, st);
  if (!rc) {
    RatSet rs{};
    for (int t = 0; t < order; t++) {
      rs.x[t] = x[t];
      rs.a[t] = res[t + 1];
    }
    hipLaunchKernelGGL(k_rateval, dim3((c.Vh + 255) / 256), dim3(256), 0, c.stream, src, res[0], rs, order, dst,
                       c.Vh);
    rc = hipGetLastError() == hipSuccess ? 0 : -1;
  }
  (void)hipStreamSynchronize(c.stream);
  for (auto& v : x) free_vec(v);
  if (iters) *iters = st.iters;
  return rc;
}

// Multi-shift solve of one source on the current links (FP64, or FP32 bulk +
// mixed refinement); x is allocated here.
int multishift_dev(const VecField& src, int n, const double* offsets, double resid, int mixed, int parity,
                   std::vector<VecField>& x, SolveStats& st) {
  auto t0 = std::chrono::steady_clock::now();
  int rc = rational_solve(src, n, offsets, resid, mixed, parity, x, st);
  (void)hipStreamSynchronize(ctx().stream);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int rhmc_action(const VecField& phi, int order, const double* roots, const double* residues, double resid, int mixed,
                double* action, int* iters) {
  Context& c = ctx();
  std::vector<VecField> x;
  SolveStats st;
  int rc = rational_solve(phi, order, roots + 1, resid, mixed, EVEN, x, st);
  if (rc) return rc;
  RatSet rs{};
  for (int t = 0; t < order; t++) {
    rs.x[t] = x[t];
    rs.a[t] = residues[t + 1];
  }
  VecField sum;
  if ((rc = alloc_vec(sum, P_F64, c.stride))) return rc;
  hipLaunchKernelGGL(k_rateval, dim3((c.Vh + 255) / 256), dim3(256), 0, c.stream, phi, residues[0], rs, order, sum,
                     c.Vh);
  MH_CHECK(hipGetLastError());
  rc = vec_norm2(sum, P_F64, action);
  free_vec(sum);
  for (auto& v : x) free_vec(v);
  if (iters) *iters = st.iters;
  return rc;
}

int force_level0_from_solutions(const std::vector<VecField>& x, const double* residues, LinkField& o1, LinkField& o3);

// S = sum_t r_t Re<phi, X_t>, X_t = (M^dag M + sigma_t)^-1 phi on even sites,
// and the level-0 force O1/O3 with coefficients r_t from Z_t = (X_t, D_oe X_t).
int rhmc_force_level0(const VecField& phi, int n, const double* offsets, const double* residues, double resid,
                      int mixed, double* S, LinkField& o1, LinkField& o3, int* iters) {
  std::vector<VecField> x;
  SolveStats st;
  int rc = rational_solve(phi, n, offsets, resid, mixed, EVEN, x, st);
  if (rc) return rc;
  double s = 0;
  for (int t = 0; t < n; t++) {
    double d;
    if ((rc = redot(phi, x[t], &d))) return rc;
    s += residues[t] * d;
  }
  *S = s;
  rc = force_level0_from_solutions(x, residues, o1, o3);
  for (auto& v : x) free_vec(v);
  if (iters) *iters = st.iters;
  return rc;
}

// Level-0 force O1/O3 with coefficients r_t from the even-site solutions X_t:
// Z_t = (X_t, D_oe X_t) (update_h_rhmc.c:77-107), the odd halves in temporaries.
int force_level0_from_solutions(const std::vector<VecField>& x, const double* residues, LinkField& o1, LinkField& o3) {
  Context& c = ctx();
  const int n = (int)x.size();
  std::vector<VecField> y(n);
  std::vector<double> r(residues, residues + n);
  int rc = 0;
  for (int t = 0; t < n && !rc; t++) {
    rc = alloc_vec(y[t], P_F64, c.stride);
    if (!rc) rc = dslash(c.links[P_F64], x[t], y[t], 1, P_F64);  // odd sites: D_oe X_t
  }
  if (!rc) rc = force_oprod_coeff(x, y, r, r, o1, o3);
  (void)hipStreamSynchronize(c.stream);
  for (auto& v : y) free_vec(v);
  return rc;
}

// ---------------------------------------------------------------------------
// One RHMC trajectory on the device (update(), ks_imp_rhmc/update_rhmc.c:157-345)
// ---------------------------------------------------------------------------
__global__ void k_axpy_flat(double* __restrict__ y, const double* __restrict__ x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] += x[i];
}

namespace {

constexpr double kReunitTol = 1e-4;   // TOLERANCE, generic/reunitarize2.c:16
constexpr int kReunitMaxErrors = 100;  // MAXERRCOUNT, generic/reunitarize2.c:17

// One cycle of an integrator (update_rhmc.c:309-630): momentum kicks at times
// t (units of epsilon from the cycle start), the links carried from kick to
// kick by update_u, then to the cycle end, then reunitarize_ks.
enum KickKind { KICK_GAUGE = 0, KICK_FERMION = 1, KICK_BOTH = 2 };
struct Kick {
  double t;
  int kind;
  double size;  // the kick is update_h_*(size * epsilon)
};
struct Schedule {
  double length = 0;
  int steps = 0;  // MD steps (fermion forces) per cycle
  std::vector<Kick> kicks;
};

// MILC's int_alg_t values (ks_imp_rhmc/ks_imp_includes.h:58-60)
int make_schedule(int alg, Schedule& s) {
  s.kicks.clear();
  if (alg == 0) {  // INT_LEAPFROG: U(e/2) H(e) U(e/2)
    s.length = 1;
    s.steps = 1;
    s.kicks.push_back({0.5, KICK_BOTH, 1.0});
    return 0;
  }
  if (alg == 1) {  // INT_OMELYAN, lambda = 0.8: U(le/2) H(e) U((2-l)e) H(e) U(le/2)
    const double lam = 0.8;
    s.length = 2;
    s.steps = 2;
    s.kicks.push_back({0.5 * lam, KICK_BOTH, 1.0});
    s.kicks.push_back({0.5 * lam + (2.0 - lam), KICK_BOTH, 1.0});
    return 0;
  }
  int n = alg == 4 ? 2 : alg == 5 ? 3 : alg == 6 ? 5 : alg == 7 ? 6 : 0;
  if (!n) return -2;
  // INT_nG1F (alpha = beta = 0.1): n gauge kicks of epsilon/n per fermion
  // step, at ((2k+1)/(2n) -+ alpha/n) epsilon, fermion kicks of epsilon at
  // (1/2 - beta) epsilon and (3/2 + beta) epsilon
  const double alpha = 0.1, beta = 0.1;
  s.length = 2;
  s.steps = 2;
  for (int k = 0; k < 2 * n; k++)
    s.kicks.push_back({(2.0 * k + 1.0) / (2.0 * n) + ((k & 1) ? alpha : -alpha) / n, KICK_GAUGE, 1.0 / n});
  s.kicks.push_back({0.5 - beta, KICK_FERMION, 1.0});
  s.kicks.push_back({1.5 + beta, KICK_FERMION, 1.0});
  // in time order; at equal times the gauge kick first (INT_6G1F: 5/12 - alpha/6 = 1/2 - beta)
  std::stable_sort(s.kicks.begin(), s.kicks.end(), [](const Kick& a, const Kick& b) {
    if (std::fabs(a.t - b.t) > 1e-12) return a.t < b.t;
    return a.kind < b.kind;
  });
  return 0;
}

struct Pseudo {
  VecField phi{};
  double offs_md[33], offs_gr[33], offs_fa[33];
};

struct Traj {
  const RhmcParams& p;
  LinkField Ub{}, V{}, W{}, o1{}, o3{}, Uold{};
  // Naik sets (load_hisq_fn_links, fermion_links_hisq_load_milc.c:574-640):
  // the distinct naik_term_epsilon values, 0 first; with more than one, the
  // set-0 links X_2 are kept in fat0/lng0 and the context's links are made
  // X_2 + eps X_3 for the solves of that set's pseudofermions
  std::vector<double> sets{0.0};
  double cur_eps = 0;
  LinkField fat0{}, lng0{}, o1s{}, o3s{}, o1x{};
  double *mom = nullptr, *mtmp = nullptr;
  std::vector<Pseudo> pf;
  bool dirty = false;  // thin links changed since the fat and Naik links were built
  bool aux = false;    // V, W hold the fat7 links of the current thin links and their U(3) projection
  double t_ferm = 0, t_gauge = 0, t_links = 0, t_md_solve = 0;
  long md_iters = 0, fa_iters = 0;
  int n_ferm = 0, n_gauge = 0, n_rebuild = 0, reunit_errors = 0;
  double reunit_maxdev = 0;
  explicit Traj(const RhmcParams& prm) : p(prm) {}
  ~Traj() {
    free_links(Ub); free_links(V); free_links(W); free_links(o1); free_links(o3); free_links(Uold);
    free_links(fat0); free_links(lng0); free_links(o1s); free_links(o3s); free_links(o1x);
    dev_free(mom); dev_free(mtmp);
    for (auto& q : pf) free_vec(q.phi);
  }
  size_t nmom() const { return (size_t)2 * ctx().Vh * 4 * 10; }
  bool naik() const { return sets.size() > 1; }
  // the pseudofermions of the Naik set with correction eps, in their order
  std::vector<int> members(double eps) const {
    std::vector<int> v;
    for (int i = 0; i < p.n_pseudo; i++)
      if (p.pf[i].naik_eps == eps) v.push_back(i);
    return v;
  }
  // several pseudofermions with mixed solves share one batched multi-shift
  static bool batched(int mixed, size_t n) { return mixed && n >= 2; }
  std::vector<const VecField*> sources(const std::vector<int>& idx) const {
    std::vector<const VecField*> v;
    for (int i : idx) v.push_back(&pf[i].phi);
    return v;
  }
  std::vector<std::vector<double>> offsets(const std::vector<int>& idx, double (Pseudo::*offs)[33],
                                           RhmcRat RhmcPseudo::*rat) const {
    std::vector<std::vector<double>> v;
    for (int i : idx) v.emplace_back(pf[i].*offs + 1, pf[i].*offs + 1 + (p.pf[i].*rat).order);
    return v;
  }
  std::vector<double> resids(const std::vector<int>& idx, double RhmcPseudo::*r) const {
    std::vector<double> v;
    for (int i : idx) v.push_back(p.pf[i].*r);
    return v;
  }
  // the FP32 and 16-bit copies follow the FP64 links: drop them
  static void drop_derived() {
    Context& c = ctx();
    for (int q = P_F32; q <= P_H16; q++)
      if (c.links[q].valid) {
        free_links(c.links[q].fat);
        free_links(c.links[q].lng);
        c.links[q].valid = false;
      }
    c.links_version++;
  }
  // make the context's FP64 links those of the Naik set eps: X_2 + eps X_3,
  // X_3 = (1/8 W, -1/24 W W W), whose long link is X_2's times the ratio of
  // the two tables' Naik coefficients; W is the current thin links' (fresh())
  int use_set(double eps) {
    if (eps == cur_eps) return 0;
    Context& c = ctx();
    DevLinks& L = c.links[P_F64];
    const size_t nb = link_bytes(P_F64, c.stride);
    MH_CHECK(hipMemcpyAsync(L.fat.data, fat0.data, nb, hipMemcpyDeviceToDevice, c.stream));
    MH_CHECK(hipMemcpyAsync(L.lng.data, lng0.data, nb, hipMemcpyDeviceToDevice, c.stream));
    int rc = 0;
    if (eps != 0.0 && !(rc = link_axpy_dev(L.fat, W, eps * kHisqNaikSet[0], 1.0, 0)))
      rc = link_axpy_dev(L.lng, L.lng, eps * kHisqNaikSet[1] / kHisqLevel2[1], 1.0, 0);
    drop_derived();
    cur_eps = eps;
    return rc;
  }

  // fat and Naik links of the current thin links, rebuilt only when used
  // (MILC invalidates them in update_u and rebuilds on demand)
  int fresh() {
    if (!dirty) return 0;
    Context& c = ctx();
    auto t0 = std::chrono::steady_clock::now();
    DevLinks& L = c.links[P_F64];
    int nsvd = 0, rc;
    if (p.action == 1) {
      rc = asqtad_links_dev(c.thin, p.u0, L.fat, L.lng);
      aux = false;
    } else {
      rc = hisq_links_dev(c.thin, L.fat, L.lng, &W, &nsvd, &V);
      aux = rc == 0;
    }
    drop_derived();
    if (!rc && naik()) {
      // keep the set-0 links
      const size_t nb = link_bytes(P_F64, c.stride);
      MH_CHECK(hipMemcpyAsync(fat0.data, L.fat.data, nb, hipMemcpyDeviceToDevice, c.stream));
      MH_CHECK(hipMemcpyAsync(lng0.data, L.lng.data, nb, hipMemcpyDeviceToDevice, c.stream));
    }
    cur_eps = 0.0;
    (void)hipStreamSynchronize(c.stream);
    t_links += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    n_rebuild++;
    dirty = false;
    return rc;
  }
  int update_u(double e) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = update_u_dev(ctx().thin, mom, e);
    (void)hipStreamSynchronize(ctx().stream);
    t_links += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    dirty = true;
    aux = false;
    return rc;
  }
  // reunitarize_ks (generic_ks/reunitarize_ks.c): phases out, reunit_su3, phases in
  int reunitarize() {
    auto t0 = std::chrono::steady_clock::now();
    int err = 0;
    double md = 0;
    int rc = reunitarize_dev(ctx().thin, 1, kReunitTol, &err, &md);
    t_links += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    dirty = true;
    aux = false;
    reunit_errors += err;
    reunit_maxdev = std::max(reunit_maxdev, md);
    if (!rc && err > kReunitMaxErrors) rc = -6;  // MILC terminates: "Unitarity error count exceeded"
    return rc;
  }
  // links without the KS phases, for the gauge loops
  int bare() {
    Context& c = ctx();
    MH_CHECK(hipMemcpyAsync(Ub.data, c.thin.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice, c.stream));
    return rephase_dev(Ub);
  }
  // update_h_gauge: imp_gauge_force_ks(e)
  int gauge_kick(double e) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = bare();
    if (!rc) rc = gauge_force_dev(Ub, p.loop_coeff, e * p.beta / 3.0, mom);
    (void)hipStreamSynchronize(ctx().stream);
    t_gauge += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    n_gauge++;
    return rc;
  }
  // update_h_fermion (update_h_rhmc.c:38-111): every pseudofermion's MD
  // rational solve on the current links, then one force from all the terms
  int fermion_kick(double e) {
    Context& c = ctx();
    int rc;
    if ((rc = fresh())) return rc;
    auto t0 = std::chrono::steady_clock::now();
    if (naik()) {
      const size_t nb = link_bytes(P_F64, c.stride);
      MH_CHECK(hipMemsetAsync(o1.data, 0, nb, c.stream));
      MH_CHECK(hipMemsetAsync(o3.data, 0, nb, c.stream));
      MH_CHECK(hipMemsetAsync(o1x.data, 0, nb, c.stream));
    }
    // update_h_fermion: each Naik set's pseudofermions on that set's links,
    // X_t and D_oe X_t with them (update_h_rhmc.c:59-83)
    for (size_t s = 0; s < sets.size() && !rc; s++) {
      const double eps = sets[s];
      const std::vector<int> idx = members(eps);
      auto ts = std::chrono::steady_clock::now();
      if ((rc = use_set(eps))) break;
      // the sloppy links of the mixed solve were dropped by the last link update
      if (p.mixed_md && (rc = ensure_links(P_F32))) break;
      std::vector<VecField> all;
      std::vector<double> res;
      if (batched(p.mixed_md, idx.size())) {
        // the set's MD solves in one batched multi-shift
        std::vector<std::vector<VecField>> xs;
        long it = 0;
        rc = rational_solve_all(sources(idx), offsets(idx, &Pseudo::offs_md, &RhmcPseudo::md),
                                resids(idx, &RhmcPseudo::resid_md), xs, &it);
        md_iters += it;
        for (int j = 0; j < (int)xs.size(); j++)
          for (int t = 0; t < (int)xs[j].size(); t++) {
            all.push_back(xs[j][t]);
            res.push_back(p.pf[idx[j]].md.res[t + 1]);
          }
      } else {
        for (int i : idx) {
          if (rc) break;
          std::vector<VecField> x;
          SolveStats st;
          rc = rational_solve(pf[i].phi, p.pf[i].md.order, pf[i].offs_md + 1, p.pf[i].resid_md, p.mixed_md, EVEN,
                              x, st);
          md_iters += st.iters;
          for (int t = 0; t < (int)x.size(); t++) {
            all.push_back(x[t]);
            res.push_back(p.pf[i].md.res[t + 1]);
          }
        }
      }
      (void)hipStreamSynchronize(c.stream);
      t_md_solve += std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
      if (!rc) rc = force_level0_from_solutions(all, res.data(), naik() ? o1s : o1, naik() ? o3s : o3);
      for (auto& v : all) free_vec(v);
      if (!rc && naik()) {
        // every term takes the level-2 paths; a set with eps != 0 also eps
        // times the "1-link + Naik" paths: its three-hop products join the
        // Naik path, its one-hop products go onto W
        // (fermion_force_hisq_multi.c:1284-1374)
        rc = link_axpy_dev(o1, o1s, 1.0, 1.0, 0);
        if (!rc) rc = link_axpy_dev(o3, o3s, 1.0 + eps * kHisqNaikSet[1] / kHisqLevel2[1], 1.0, 0);
        if (!rc && eps != 0.0) rc = link_axpy_dev(o1x, o1s, eps * kHisqNaikSet[0], 1.0, 0);
      }
    }
    if (p.action == 1) {
      // asqtad: the FN force on the thin links (fermion_force_fn_multi)
      const double u2 = p.u0 * p.u0, u4 = u2 * u2, u6 = u4 * u2;
      const double co[6] = {5.0 / 8.0, -1.0 / 24.0 / u2, -1.0 / 16.0 / u2, 1.0 / 64.0 / u4, -1.0 / 384.0 / u6,
                            -1.0 / 16.0 / u4};
      if (!rc) rc = fn_force_from_oprod(c.thin, o1, o3, co, 2.0 * e, mtmp);
    } else {
      if (!rc && !aux) rc = hisq_aux_dev(c.thin, V, W);
      aux = rc == 0;
      if (!rc)
        rc = hisq_force_from_oprod(c.thin, V, W, o1, o3, kHisqLevel2, kHisqLevel1, 2.0 * e, kForceFilter, mtmp,
                                   naik() ? &o1x : nullptr);
    }
    if (!rc) {
      hipLaunchKernelGGL(k_axpy_flat, dim3(1024), dim3(256), 0, c.stream, mom, mtmp, nmom());
      rc = hipGetLastError() == hipSuccess ? 0 : -1;
    }
    (void)hipStreamSynchronize(c.stream);
    t_ferm += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    n_ferm++;
    return rc;
  }
  int kick(const Kick& k) {
    const double e = k.size * p.epsilon;
    if (k.kind == KICK_GAUGE) return gauge_kick(e);
    if (k.kind == KICK_FERMION) return fermion_kick(e);
    int rc = gauge_kick(e);  // update_h_rhmc: gauge force, then the fermion force
    return rc ? rc : fermion_kick(e);
  }
  // d_action_rhmc: beta/3 imp_gauge_action, hmom_action, fermion_action
  int actions(double* g, double* h, double* f) {
    int rc = bare();
    double ga = 0;
    if (!rc) rc = gauge_action_dev(Ub, p.loop_coeff, &ga);
    *g = p.beta / 3.0 * ga;
    if (!rc) rc = mom_action_dev(mom, h);
    if (!rc) rc = fresh();
    double fs = 0;
    // each Naik set's pseudofermions on that set's links
    for (size_t s = 0; s < sets.size() && !rc; s++) {
      const std::vector<int> idx = members(sets[s]);
      rc = use_set(sets[s]);
      if (!rc && p.mixed_action) rc = ensure_links(P_F32);
      if (!rc && !p.action_from_md && batched(p.mixed_action, idx.size())) {
        // sum over the set's pseudofermions of |R_FA phi|^2 from one batched multi-shift
        std::vector<std::vector<VecField>> xs;
        long it = 0;
        rc = rational_solve_all(sources(idx), offsets(idx, &Pseudo::offs_fa, &RhmcPseudo::fa),
                                resids(idx, &RhmcPseudo::resid_fa), xs, &it);
        fa_iters += it;
        VecField sum{};
        if (!rc) rc = alloc_vec(sum, P_F64, ctx().stride);
        for (int j = 0; j < (int)xs.size() && !rc; j++) {
          double a = 0;
          rc = rateval(pf[idx[j]].phi, xs[j], p.pf[idx[j]].fa.res, sum);
          if (!rc) rc = vec_norm2(sum, P_F64, &a);
          fs += a;
        }
        free_vec(sum);
        for (auto& v : xs)
          for (auto& y : v) free_vec(y);
        continue;
      }
      for (int i : idx) {
        if (rc) break;
        rc = pseudo_action(i, fs);
      }
    }
    *f = fs;
    return rc;
  }
  // one pseudofermion's action on the current links, added to fs
  int pseudo_action(int i, double& fs) {
    int rc = 0;
    {
      const RhmcPseudo& q = p.pf[i];
      if (p.action_from_md) {
        // phi^dag R_MD phi: the action whose derivative the MD force is
        std::vector<VecField> x;
        SolveStats st;
        rc = rational_solve(pf[i].phi, q.md.order, pf[i].offs_md + 1, q.resid_fa, p.mixed_action, EVEN, x, st);
        fa_iters += st.iters;
        double s = 0, n2 = 0;
        for (int t = 0; t < q.md.order && !rc; t++) {
          double d;
          rc = redot(pf[i].phi, x[t], &d);
          s += q.md.res[t + 1] * d;
        }
        if (!rc) rc = vec_norm2(pf[i].phi, P_F64, &n2);
        for (auto& v : x) free_vec(v);
        fs += q.md.res[0] * n2 + s;
      } else {
        double a = 0;
        int it = 0;
        rc = rhmc_action(pf[i].phi, q.fa.order, pf[i].offs_fa, q.fa.res, q.resid_fa, p.mixed_action, &a, &it);
        fa_iters += it;
        fs += a;
      }
    }
    return rc;
  }
};

// links_version of the context's links when a trajectory last left asqtad
// fat and Naik links there (every other builder makes HISQ links)
unsigned g_asqtad_links_version = ~0u;

uint64_t host_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

bool rat_ok(const RhmcRat& r) { return r.order > 0 && r.order <= 32; }

int g_fail_stage = 0;
// the injected failure of stage `stage`, once
int injected(int stage) {
  if (g_fail_stage != stage) return 0;
  g_fail_stage = 0;
  return -90;
}

}  // namespace

void rhmc_fail_at(int stage) { g_fail_stage = stage; }

int rhmc_trajectory(const RhmcParams& p, double* stats) {
  Context& c = ctx();
  if (!c.thin_valid || !c.links[P_F64].valid) return -3;
  Schedule sch;
  if (make_schedule(p.integrator, sch) || p.steps <= 0 || p.steps % sch.steps) return -2;
  if (p.n_pseudo <= 0 || p.n_pseudo > RHMC_MAX_PSEUDO) return -2;
  if (p.site_rng && !prn_valid()) return -3;
  auto t_start = std::chrono::steady_clock::now();
  Traj T(p);
  T.pf.resize(p.n_pseudo);
  for (int i = 0; i < p.n_pseudo; i++)
    if (std::find(T.sets.begin(), T.sets.end(), p.pf[i].naik_eps) == T.sets.end()) {
      if (p.action != 0) return -2;
      T.sets.push_back(p.pf[i].naik_eps);
    }
  // the fat and Naik links in place must be those of this quark action
  T.dirty = (p.action == 1) != (c.links_version == g_asqtad_links_version);
  for (int i = 0; i < p.n_pseudo; i++) {
    const RhmcPseudo& q = p.pf[i];
    if (!rat_ok(q.md) || !rat_ok(q.gr) || !rat_ok(q.fa)) return -2;
    const double m4 = 4.0 * q.mass * q.mass;
    Pseudo& P = T.pf[i];
    P.offs_md[0] = P.offs_gr[0] = P.offs_fa[0] = 0.0;
    for (int k = 1; k <= 32; k++) {
      P.offs_md[k] = k <= q.md.order ? m4 + q.md.pole[k] : 0.0;
      P.offs_gr[k] = k <= q.gr.order ? m4 + q.gr.pole[k] : 0.0;
      P.offs_fa[k] = k <= q.fa.order ? m4 + q.fa.pole[k] : 0.0;
    }
    for (int k = 1; k <= q.md.order; k++)
      if (!(P.offs_md[k] > 0)) return -2;
    for (int k = 1; k <= q.gr.order; k++)
      if (!(P.offs_gr[k] > 0)) return -2;
    for (int k = 1; k <= q.fa.order; k++)
      if (!(P.offs_fa[k] > 0)) return -2;
  }
  int rc;
  if ((rc = alloc_links(T.Ub, P_F64, c.stride)) || (rc = alloc_links(T.V, P_F64, c.stride)) ||
      (rc = alloc_links(T.W, P_F64, c.stride)) || (rc = alloc_links(T.o1, P_F64, c.stride)) ||
      (rc = alloc_links(T.o3, P_F64, c.stride)) || (rc = alloc_links(T.Uold, P_F64, c.stride)) ||
      (rc = dev_alloc((void**)&T.mom, T.nmom() * sizeof(double))) ||
      (rc = dev_alloc((void**)&T.mtmp, T.nmom() * sizeof(double))))
    return rc;
  if (T.naik() &&
      ((rc = alloc_links(T.fat0, P_F64, c.stride)) || (rc = alloc_links(T.lng0, P_F64, c.stride)) ||
       (rc = alloc_links(T.o1s, P_F64, c.stride)) || (rc = alloc_links(T.o3s, P_F64, c.stride)) ||
       (rc = alloc_links(T.o1x, P_F64, c.stride))))
    return rc;
  // with Naik sets, W and the set-0 links of the links in place come from a rebuild
  if (T.naik()) T.dirty = true;
  for (auto& P : T.pf)
    if ((rc = alloc_vec(P.phi, P_F64, c.stride))) return rc;
  std::vector<VecField> chi(p.n_pseudo);
  VecScope chi_owned;
  for (auto& v : chi) chi_owned.add(&v);
  for (auto& v : chi)
    if ((rc = alloc_vec(v, P_F64, c.stride))) return rc;
  // the starting links, restored by every failure from here on
  MH_CHECK(hipMemcpyAsync(T.Uold.data, c.thin.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice, c.stream));
  (void)hipStreamSynchronize(c.stream);
  const double t_setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  // A failed trajectory leaves the starting thin links in place with their
  // fat and Naik links (set 0) of this quark action, and records which
  // action those links belong to.
  auto fail = [&](int code) {
    (void)hipMemcpyAsync(c.thin.data, T.Uold.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice, c.stream);
    T.dirty = true;
    T.aux = false;
    if (T.fresh() == 0) (void)T.use_set(0.0);
    g_asqtad_links_version = p.action == 1 ? c.links_version : ~0u;
    (void)hipStreamSynchronize(c.stream);
    return code;
  };
  // ranmom, then grsource_imp_rhmc for each pseudofermion: phi = R_GR(M^dag M) chi
  double grsum[RHMC_MAX_PSEUDO] = {0};
  int it_gr = 0;
  if ((rc = T.fresh())) return fail(rc);
  if (p.mixed_action && (rc = ensure_links(P_F32))) return fail(rc);
  rc = p.site_rng ? ranmom_prn_dev(T.mom) : ranmom_dev(host_mix(p.seed), T.mom);
  // the noise of every pseudofermion first, in MILC's order of draws
  for (int i = 0; i < p.n_pseudo && !rc; i++) {
    if (p.site_rng) {
      rc = gaussian_prn_dev(0, chi[i], &grsum[i]);
    } else {
      rc = gaussian_vec_dev(host_mix(p.seed + 1) + 0x632BE59BD9B4E019ULL * (uint64_t)i, 0, chi[i]);
      if (!rc) rc = vec_norm2(chi[i], P_F64, &grsum[i]);
    }
  }
  // each Naik set's pseudofermions on that set's links (grsource_imp_rhmc with fn[inaik])
  for (size_t s = 0; s < T.sets.size() && !rc; s++) {
    const std::vector<int> idx = T.members(T.sets[s]);
    rc = T.use_set(T.sets[s]);
    if (!rc && p.mixed_action) rc = ensure_links(P_F32);
    if (!rc && Traj::batched(p.mixed_action, idx.size())) {
      std::vector<const VecField*> src;
      for (int i : idx) src.push_back(&chi[i]);
      std::vector<std::vector<VecField>> xs;
      long it = 0;
      rc = rational_solve_all(src, T.offsets(idx, &Pseudo::offs_gr, &RhmcPseudo::gr),
                              T.resids(idx, &RhmcPseudo::resid_gr), xs, &it);
      it_gr += (int)it;
      for (int j = 0; j < (int)xs.size() && !rc; j++)
        rc = rateval(chi[idx[j]], xs[j], p.pf[idx[j]].gr.res, T.pf[idx[j]].phi);
      (void)hipStreamSynchronize(c.stream);
      for (auto& v : xs)
        for (auto& y : v) free_vec(y);
    } else {
      for (int i : idx) {
        if (rc) break;
        int it = 0;
        rc = rational_apply(chi[i], p.pf[i].gr.order, T.pf[i].offs_gr + 1, p.pf[i].gr.res, p.pf[i].resid_gr,
                            p.mixed_action, T.pf[i].phi, &it);
        it_gr += it;
      }
    }
  }
  if (!rc) rc = injected(1);
  if (rc) return fail(rc);
  for (auto& v : chi) free_vec(v);
  (void)hipStreamSynchronize(c.stream);
  const double t_heatbath =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() - t_setup;
  double g0, h0, f0, g1, h1, f1;
  auto t_a0 = std::chrono::steady_clock::now();
  if ((rc = T.actions(&g0, &h0, &f0)) || (rc = injected(2))) return fail(rc);
  double t_act = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a0).count();
  const double e = p.epsilon;
  for (int cyc = 0; cyc < p.steps / sch.steps && !rc; cyc++) {
    double now = 0;
    for (const Kick& k : sch.kicks) {
      if (k.t - now > 1e-12 && (rc = T.update_u(e * (k.t - now)))) break;
      now = k.t;
      if ((rc = T.kick(k))) break;
    }
    if (!rc && sch.length - now > 1e-12) rc = T.update_u(e * (sch.length - now));
    if (!rc) rc = T.reunitarize();
  }
  if (!rc) rc = injected(3);
  if (rc) return fail(rc);
  auto t_a1 = std::chrono::steady_clock::now();
  if ((rc = T.actions(&g1, &h1, &f1))) return fail(rc);
  t_act += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a1).count();
  const double dH = (g1 + h1 + f1) - (g0 + h0 + f0);
  int accepted = 1;
  if (p.accept == 1) {
    // update(): reject when exp(-dS) < a uniform draw of the node generator
    double r = p.site_rng ? prn_node_uniform()
                          : (double)(host_mix(p.seed + 2) >> 11) * (1.0 / 9007199254740992.0);
    accepted = !(std::exp(-dH) < r);
  } else if (p.accept == 2) {
    accepted = 0;
  }
  if (!accepted) {
    if ((rc = hipMemcpyAsync(c.thin.data, T.Uold.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice,
                             c.stream)) != hipSuccess)
      return fail(-1000 - rc);
    T.dirty = true;
    T.aux = false;
  }
  // leave the fat and Naik links (set 0) of the final thin links in place
  if ((rc = T.fresh()) || (rc = T.use_set(0.0))) return fail(rc);
  g_asqtad_links_version = p.action == 1 ? c.links_version : ~0u;
  (void)hipStreamSynchronize(c.stream);
  double tot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  double out[32] = {g0, h0, f0, g1, h1, f1, dH, (double)accepted, tot, (double)T.md_iters, (double)T.n_ferm,
                    T.t_ferm, T.t_gauge, T.t_links, t_act, (double)it_gr, (double)T.n_gauge, T.reunit_maxdev,
                    (double)T.reunit_errors, (double)T.n_rebuild};
  for (int i = 0; i < p.n_pseudo && i < 8; i++) out[20 + i] = grsum[i];
  out[28] = (double)T.fa_iters;
  out[29] = t_setup;
  out[30] = t_heatbath;
  out[31] = T.t_md_solve;
  memcpy(stats, out, sizeof out);
  return 0;
}

}  // namespace mh
