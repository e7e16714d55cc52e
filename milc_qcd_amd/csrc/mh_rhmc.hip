// mh_rhmc.hip -- RHMC compositions of the solver, stencil and force kernels,
// all on device-resident fields.
//
// Restates (MILC 7.8.1):
//   ks_ratinv / ks_rateval   ks_imp_rhmc/ks_ratinv.c:33-138
//   fermion_action           ks_imp_rhmc/d_action_rhmc.c (sum over even sites of
//                            |a0 phi + sum_k a_k x_k|^2, FP64)
//   update_h_fermion         ks_imp_rhmc/update_h_rhmc.c:77-107: multi_x even =
//                            X_t, odd = D_oe X_t, then the force (level 0 here)
#include <chrono>
#include <cmath>
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

}  // namespace mh
