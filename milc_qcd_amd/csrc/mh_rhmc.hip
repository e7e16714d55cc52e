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
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

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
  mp.resid = mixed ? std::max(resid, 1e-7) : resid;
  mp.prec = mixed ? P_F32 : P_F64;
  mp.batch = 16;
  int bulk = mp.prec;
  if (!c.links[bulk].valid) return -3;
  if ((rc = multicg_solve(phi, x, offs, mp, st))) return rc;
  if (!mixed) return 0;
  // Refinement of every shift to the FP64 true-residual target (the per-shift
  // CG of ks_multicg_hybrid_field, generic_ks/ks_multicg.c:170-236): mixed CG
  // from the multi-shift solution, 16-bit sloppy operator, FP64 reliable updates.
  if ((rc = ensure_links(P_H16))) return rc;
  for (int t = 0; t < n; t++) {
    SolveParams sp;
    sp.parity = parity;
    sp.max_iters = 20000;
    sp.max_restarts = 1;
    sp.resid = resid;
    sp.relresid = 0;
    sp.sloppy_prec = P_H16;
    sp.delta = kReliableDelta;
    sp.batch = 16;
    SolveStats s2;
    if ((rc = cg_solve(phi, x[t], 0.5 * std::sqrt(offs[t]), sp, s2))) return rc;
    st.iters += s2.iters;
    st.converged &= s2.converged;
  }
  return 0;
}

// dst = res[0] src + sum_k res[k] (M^dag M + off[k])^-1 src on even sites (ks_rateval)
static int rational_apply(const VecField& src, int order, const double* offs, const double* res, double resid,
                          VecField& dst, int* iters) {
  Context& c = ctx();
  std::vector<VecField> x;
  SolveStats st;
  int rc = rational_solve(src, order, offs, resid, 0, EVEN, x, st);
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

struct Traj {
  const RhmcParams& p;
  LinkField Ub{}, V{}, W{}, o1{}, o3{}, Uold{};
  double *mom = nullptr, *mtmp = nullptr;
  VecField phi{}, chi{};
  double offs_md[33], offs_gr[33], offs_fa[33];
  double t_ferm = 0, t_gauge = 0, t_links = 0;
  long md_iters = 0;
  int n_ferm = 0;
  explicit Traj(const RhmcParams& prm) : p(prm) {}
  ~Traj() {
    free_links(Ub); free_links(V); free_links(W); free_links(o1); free_links(o3); free_links(Uold);
    dev_free(mom); dev_free(mtmp);
    free_vec(phi); free_vec(chi);
  }
  size_t nmom() const { return (size_t)2 * ctx().Vh * 4 * 10; }

  // the thin links changed: rebuild the FP64 fat + Naik links, drop sloppy copies
  int links_changed() {
    Context& c = ctx();
    auto t0 = std::chrono::steady_clock::now();
    DevLinks& L = c.links[P_F64];
    int nsvd = 0;
    int rc = hisq_links_dev(c.thin, L.fat, L.lng, nullptr, &nsvd);
    for (int q = P_F32; q <= P_H16; q++)
      if (c.links[q].valid) {
        free_links(c.links[q].fat);
        free_links(c.links[q].lng);
        c.links[q].valid = false;
      }
    c.links_version++;
    (void)hipStreamSynchronize(c.stream);
    t_links += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
  }
  int update_u(double e) {
    int rc = update_u_dev(ctx().thin, mom, e);
    return rc ? rc : links_changed();
  }
  // links without the KS phases, for the gauge loops
  int bare() {
    Context& c = ctx();
    MH_CHECK(hipMemcpyAsync(Ub.data, c.thin.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice, c.stream));
    return rephase_dev(Ub);
  }
  // update_h_rhmc: gauge force, then the fermion force of the MD rational
  int force(double e) {
    Context& c = ctx();
    auto t0 = std::chrono::steady_clock::now();
    int rc = bare();
    if (!rc) rc = gauge_force_dev(Ub, p.loop_coeff, e * p.beta / 3.0, mom);
    (void)hipStreamSynchronize(c.stream);
    auto t1 = std::chrono::steady_clock::now();
    t_gauge += std::chrono::duration<double>(t1 - t0).count();
    if (rc) return rc;
    double S;
    int it = 0;
    // the sloppy links of the mixed solve were dropped by the last link update
    if (p.mixed_md && (rc = ensure_links(P_F32))) return rc;
    rc = rhmc_force_level0(phi, p.md.order, offs_md + 1, p.md.res + 1, p.resid_md, p.mixed_md, &S, o1, o3, &it);
    if (!rc) rc = hisq_aux_dev(c.thin, V, W);
    if (!rc) rc = hisq_force_from_oprod(c.thin, V, W, o1, o3, kHisqLevel2, kHisqLevel1, 2.0 * e, kForceFilter, mtmp);
    if (!rc) {
      hipLaunchKernelGGL(k_axpy_flat, dim3(1024), dim3(256), 0, c.stream, mom, mtmp, nmom());
      rc = hipGetLastError() == hipSuccess ? 0 : -1;
    }
    (void)hipStreamSynchronize(c.stream);
    t_ferm += std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    md_iters += it;
    n_ferm++;
    return rc;
  }
  // gauge (beta/3 imp_gauge_action), momentum (hmom_action) and fermion actions
  int actions(double* g, double* h, double* f) {
    int rc = bare();
    double ga = 0;
    if (!rc) rc = gauge_action_dev(Ub, p.loop_coeff, &ga);
    *g = p.beta / 3.0 * ga;
    if (!rc) rc = mom_action_dev(mom, h);
    if (rc) return rc;
    int it;
    if (p.action_from_md) {
      // phi^dag R_MD phi: the action whose derivative the MD force is
      std::vector<VecField> x;
      SolveStats st;
      rc = rational_solve(phi, p.md.order, offs_md + 1, p.resid_fa, 0, EVEN, x, st);
      double s = 0, n2 = 0;
      for (int t = 0; t < p.md.order && !rc; t++) {
        double d;
        rc = redot(phi, x[t], &d);
        s += p.md.res[t + 1] * d;
      }
      if (!rc) rc = vec_norm2(phi, P_F64, &n2);
      for (auto& v : x) free_vec(v);
      *f = p.md.res[0] * n2 + s;
      return rc;
    }
    return rhmc_action(phi, p.fa.order, offs_fa, p.fa.res, p.resid_fa, 0, f, &it);
  }
};

uint64_t host_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

int rhmc_trajectory(const RhmcParams& p, double* stats) {
  Context& c = ctx();
  if (!c.thin_valid || !c.links[P_F64].valid) return -3;
  if (p.steps <= 0 || (p.integrator == 1 && p.steps % 2) || p.md.order <= 0 || p.md.order > 32 ||
      p.gr.order <= 0 || p.gr.order > 32 || p.fa.order <= 0 || p.fa.order > 32)
    return -2;
  auto t_start = std::chrono::steady_clock::now();
  Traj T(p);
  const double m4 = 4.0 * p.mass * p.mass;
  T.offs_md[0] = T.offs_gr[0] = T.offs_fa[0] = 0.0;
  for (int k = 1; k <= 32; k++) {
    T.offs_md[k] = k <= p.md.order ? m4 + p.md.pole[k] : 0.0;
    T.offs_gr[k] = k <= p.gr.order ? m4 + p.gr.pole[k] : 0.0;
    T.offs_fa[k] = k <= p.fa.order ? m4 + p.fa.pole[k] : 0.0;
  }
  for (int k = 1; k <= p.md.order; k++)
    if (!(T.offs_md[k] > 0)) return -2;
  for (int k = 1; k <= p.gr.order; k++)
    if (!(T.offs_gr[k] > 0)) return -2;
  for (int k = 1; k <= p.fa.order; k++)
    if (!(T.offs_fa[k] > 0)) return -2;
  int rc;
  if ((rc = alloc_links(T.Ub, P_F64, c.stride)) || (rc = alloc_links(T.V, P_F64, c.stride)) ||
      (rc = alloc_links(T.W, P_F64, c.stride)) || (rc = alloc_links(T.o1, P_F64, c.stride)) ||
      (rc = alloc_links(T.o3, P_F64, c.stride)) || (rc = alloc_links(T.Uold, P_F64, c.stride)) ||
      (rc = dev_alloc((void**)&T.mom, T.nmom() * sizeof(double))) ||
      (rc = dev_alloc((void**)&T.mtmp, T.nmom() * sizeof(double))) || (rc = alloc_vec(T.phi, P_F64, c.stride)) ||
      (rc = alloc_vec(T.chi, P_F64, c.stride)))
    return rc;
  // momenta and pseudofermion: phi = R_GR(M^dag M) chi (grsource_imp_rhmc)
  if ((rc = ranmom_dev(host_mix(p.seed), T.mom)) || (rc = gaussian_vec_dev(host_mix(p.seed + 1), 0, T.chi)))
    return rc;
  int it_gr = 0;
  if ((rc = rational_apply(T.chi, p.gr.order, T.offs_gr + 1, p.gr.res, p.resid_gr, T.phi, &it_gr))) return rc;
  double g0, h0, f0, g1, h1, f1;
  auto t_a0 = std::chrono::steady_clock::now();
  if ((rc = T.actions(&g0, &h0, &f0))) return rc;
  double t_act = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a0).count();
  MH_CHECK(hipMemcpyAsync(T.Uold.data, c.thin.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice, c.stream));
  const double e = p.epsilon;
  if (p.integrator == 0) {
    for (int s = 1; s <= p.steps && !rc; s++) {
      if (!(rc = T.update_u(0.5 * e)) && !(rc = T.force(e))) rc = T.update_u(0.5 * e);
    }
  } else {
    const double lam = 0.8;
    for (int s = 2; s <= p.steps && !rc; s += 2) {
      if ((rc = T.update_u(0.5 * e * lam)) || (rc = T.force(e)) || (rc = T.update_u(e * (2.0 - lam))) ||
          (rc = T.force(e)))
        break;
      rc = T.update_u(0.5 * e * lam);
    }
  }
  if (rc) return rc;
  auto t_a1 = std::chrono::steady_clock::now();
  if ((rc = T.actions(&g1, &h1, &f1))) return rc;
  t_act += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_a1).count();
  const double dH = (g1 + h1 + f1) - (g0 + h0 + f0);
  int accepted = 1;
  if (p.accept == 1) {
    double r = (double)(host_mix(p.seed + 2) >> 11) * (1.0 / 9007199254740992.0);
    accepted = std::exp(-dH) >= r;
  } else if (p.accept == 2) {
    accepted = 0;
  }
  if (!accepted) {
    MH_CHECK(hipMemcpyAsync(c.thin.data, T.Uold.data, link_bytes(P_F64, c.stride), hipMemcpyDeviceToDevice, c.stream));
    if ((rc = T.links_changed())) return rc;
  }
  (void)hipStreamSynchronize(c.stream);
  double tot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  double out[16] = {g0, h0, f0, g1, h1, f1, dH, (double)accepted, tot, (double)T.md_iters, (double)T.n_ferm,
                    T.t_ferm, T.t_gauge, T.t_links, t_act, (double)it_gr};
  memcpy(stats, out, sizeof out);
  return 0;
}

}  // namespace mh
