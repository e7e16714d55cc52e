// mh_multicg.hip -- multi-shift CG on the device.
//
// Restates generic_ks/ks_multicg_offset.c:63-505 (ks_multicg_offset_field_cpu):
// Jegerlehner multi-mass CG for (-D_eo D_oe + offset_j) x_j = b with the
// smallest offset as the base system, zeta/beta/alpha recurrences in FP64 and
// retirement of shifts whose zeta underflows (:324-355, :429-444).
// The recurrences run in single-block kernels on device-resident scalars.
// The updates of psim_j and pm_j (:390-396, :447-453) are deferred over
// windows of MS_KW iterations: the window's residuals are kept in a ring, the
// per-iteration beta/zeta/alpha are folded into FP64 window coefficients
//   pm_j^(s+m) = wa_m pm_j^(s) + sum_l wc_{m,l} r^(s+l)
//   psim_j^(s+K) = psim_j^(s) + wx pm_j^(s) + sum_l wb_l r^(s+l)
// and one flush kernel per window streams every shift's pair once.  The scalar
// recurrences (and so the iteration count) are those of MILC; only the
// rounding of the vector arithmetic differs.  One window (reset, K
// iterations, flush) is captured as a hipGraph and replayed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "mh_dslash.h"
#include "mh_internal.h"

namespace mh {

#define MS_MAX 32
#define MS_KW 8

struct MSScalars {
  double zeta_i[MS_MAX], zeta_im1[MS_MAX], zeta_ip1[MS_MAX];
  double beta_i[MS_MAX], beta_im1[MS_MAX], alpha[MS_MAX];
  double shifts[MS_MAX];
  double wa[MS_MAX], wx[MS_MAX], wc[MS_MAX][MS_KW], wb[MS_MAX][MS_KW];
  double rsq, oldrsq, pkp, source_norm, rsqstop;
  int num_now, j_low, n, halt, iters, converged, m, now_ws;
  int max_iter, pad;
};

template <int N>
__device__ __forceinline__ void block_sum4(double (&v)[N], double* __restrict__ partials) {
#pragma unroll
  for (int k = 0; k < N; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  __shared__ double sm[N][16];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; k++) sm[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < N; k++) {
      double s = 0;
      for (int w = 0; w < nw; w++) s += sm[k][w];
      partials[blockIdx.x * 4 + k] = s;
    }
  }
}

__device__ __forceinline__ double reduce_one(const double* __restrict__ parts, int nblocks) {
  double t[1] = {0.0};
  reduce_records<1>(parts, nblocks, t);
  return t[0];
}

template <int P>
__global__ void __launch_bounds__(256) k_ms_dslash1(Geom g, LinkField fat, LinkField lng, VecField src, VecField dst,
                                                    int pb, const MSScalars* __restrict__ sc) {
  if (sc->halt) return;
  int bid = xcd_block(blockIdx.x, gridDim.x);
  int cb = bid * blockDim.x + threadIdx.x;
  if (cb >= g.Vh) return;
  Vec3<typename PrecTraits<P>::Real> acc;
  dslash_site<P>(g, fat, lng, src, cb, pb, acc);
  store_vec_nt<P>(dst, cb, acc);
}

// ttt = D tmp + shift0 * cg_p ; pkp = <cg_p, ttt>   (ks_multicg_offset.c:296-318)
template <int P>
__global__ void __launch_bounds__(256) k_ms_dslash2(Geom g, LinkField fat, LinkField lng, VecField tmp, VecField p,
                                                    VecField ttt, int pb, const MSScalars* __restrict__ sc,
                                                    double* __restrict__ partials) {
  if (sc->halt) return;
  typedef typename PrecTraits<P>::Real R;
  int bid = xcd_block(blockIdx.x, gridDim.x);
  int cb = bid * blockDim.x + threadIdx.x;
  double s[1] = {0};
  if (cb < g.Vh) {
    R shift0 = (R)(-sc->shifts[sc->j_low]);
    Vec3<R> acc, pv;
    dslash_site<P>(g, fat, lng, tmp, cb, pb, acc);
    load_vec<P>(p, cb, pv);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      acc.c[k].re = acc.c[k].re + shift0 * pv.c[k].re;
      acc.c[k].im = acc.c[k].im + shift0 * pv.c[k].im;
      s[0] += (double)pv.c[k].re * (double)acc.c[k].re + (double)pv.c[k].im * (double)acc.c[k].im;
    }
    store_vec_nt<P>(ttt, cb, acc);
  }
  block_sum4<1>(s, partials);
}

// beta/zeta recurrences (ks_multicg_offset.c:322-356)
__global__ void k_ms_beta(const double* __restrict__ partials, int nblocks, MSScalars* sc) {
  if (sc->halt) return;
  double pkp = reduce_one(partials, nblocks);
  if (threadIdx.x != 0) return;
  int jl = sc->j_low;
  sc->pkp = pkp;
  sc->iters += 1;
  sc->beta_i[jl] = -sc->rsq / pkp;
  sc->zeta_ip1[jl] = 1.0;
  int now = sc->num_now;
  for (int j = 0; j < now; j++) {
    if (j == jl) continue;
    sc->zeta_ip1[j] = sc->zeta_i[j] * sc->zeta_im1[j] * sc->beta_im1[jl];
    double c1 = sc->beta_i[jl] * sc->alpha[jl] * (sc->zeta_im1[j] - sc->zeta_i[j]);
    double c2 = sc->zeta_im1[j] * sc->beta_im1[jl] * (1.0 + sc->shifts[j] * sc->beta_i[jl]);
    if (c1 + c2 != 0.0) sc->zeta_ip1[j] /= c1 + c2;
    else sc->zeta_ip1[j] = 0.0;
    if (sc->zeta_i[j] != 0.0) {
      sc->beta_i[j] = sc->beta_i[jl] * sc->zeta_ip1[j] / sc->zeta_i[j];
    } else {
      sc->zeta_ip1[j] = 0.0;
      sc->beta_i[j] = 0.0;
      if (j == now - 1 && j > jl) now--;
    }
  }
  sc->num_now = now;
  // psim_j += beta_j pm_j^(s+m-1), pm_j^(s+m-1) = wa pm_j^(s) + sum_{l<m-1} wc_l r^(s+1+l)
  int m = sc->m + 1;
  sc->m = m;
  for (int j = 0; j < now; j++) {
    double b = sc->beta_i[j];
    sc->wx[j] += b * sc->wa[j];
    for (int l = 0; l < m - 1; l++) sc->wb[j][l] += b * sc->wc[j][l];
  }
}

__global__ void k_ms_reset(MSScalars* sc) {
  if (threadIdx.x != 0) return;
  sc->m = 0;
  sc->now_ws = sc->num_now;
  for (int j = 0; j < sc->n; j++) {
    sc->wa[j] = 1.0;
    sc->wx[j] = 0.0;
    for (int l = 0; l < MS_KW; l++) sc->wc[j][l] = sc->wb[j][l] = 0.0;
  }
}

// r^(m) = r^(m-1) + beta_low ttt; |r^(m)|^2 partials (ks_multicg_offset.c:394-396)
template <int P>
__global__ void __launch_bounds__(256) k_ms_r(VecField rprev, VecField rcur, VecField ttt, int Vh,
                                              const MSScalars* __restrict__ sc, double* __restrict__ partials) {
  if (sc->halt) return;
  typedef typename PrecTraits<P>::Real R;
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  double s[1] = {0};
  if (cb < Vh) {
    Vec3<R> rv, tv;
    load_vec<P>(rprev, cb, rv);
    load_vec<P>(ttt, cb, tv);
    R bl = (R)sc->beta_i[sc->j_low];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      rv.c[k].re = rv.c[k].re + bl * tv.c[k].re;
      rv.c[k].im = rv.c[k].im + bl * tv.c[k].im;
      s[0] += (double)rv.c[k].re * (double)rv.c[k].re + (double)rv.c[k].im * (double)rv.c[k].im;
    }
    store_vec<P>(rcur, cb, rv);
  }
  block_sum4<1>(s, partials);
}

#define MS_RING (MS_KW + 1)
// Field table of one solve, in device memory: the residual ring, then psim_j,
// then pm_j (the flush kernel reads it instead of taking ~1.7 KB of by-value
// arguments).
#define MS_TAB_PSIM MS_RING
#define MS_TAB_PM (MS_RING + MS_MAX)
#define MS_TAB_N (MS_RING + 2 * MS_MAX)

// window flush: psim_j and pm_j brought up to date from the ring of residuals;
// ring[0] = the window's last residual
template <int P>
__global__ void __launch_bounds__(256) k_ms_flush(const VecField* __restrict__ tab, int Vh,
                                                  const MSScalars* __restrict__ sc) {
  typedef typename PrecTraits<P>::Real R;
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  const int md = sc->m;
  if (cb >= Vh || md == 0) return;
  Vec3<R> rv[MS_KW];
#pragma unroll
  for (int l = 0; l < MS_KW; l++) {
    if (l < md) load_vec<P>(tab[l + 1], cb, rv[l]);
    else zero(rv[l]);
  }
  const int nn = sc->now_ws;
  for (int j = 0; j < nn; j++) {
    Vec3<R> d, x;
    const VecField pmj = tab[MS_TAB_PM + j], psimj = tab[MS_TAB_PSIM + j];
    load_vec<P>(pmj, cb, d);
    load_vec<P>(psimj, cb, x);
    R cx = (R)sc->wx[j], ca = (R)sc->wa[j];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R xr = x.c[k].re + cx * d.c[k].re, xi = x.c[k].im + cx * d.c[k].im;
      R qr = ca * d.c[k].re, qi = ca * d.c[k].im;
#pragma unroll
      for (int l = 0; l < MS_KW; l++) {
        R b = (R)sc->wb[j][l], cc = (R)sc->wc[j][l];
        xr += b * rv[l].c[k].re;
        xi += b * rv[l].c[k].im;
        qr += cc * rv[l].c[k].re;
        qi += cc * rv[l].c[k].im;
      }
      x.c[k].re = xr;
      x.c[k].im = xi;
      d.c[k].re = qr;
      d.c[k].im = qi;
    }
    store_vec<P>(psimj, cb, x);
    store_vec<P>(pmj, cb, d);
  }
#pragma unroll
  for (int l = 0; l < MS_KW; l++)
    if (l + 1 == md) store_vec<P>(tab[0], cb, rv[l]);
}

// convergence test, alpha recurrences and scalar scroll (ks_multicg_offset.c:358-463)
__global__ void k_ms_alpha(const double* __restrict__ partials, int nblocks, MSScalars* sc) {
  if (sc->halt) return;
  double rsq = reduce_one(partials, nblocks);
  if (threadIdx.x != 0) return;
  double oldrsq = sc->rsq;
  sc->oldrsq = oldrsq;
  sc->rsq = rsq;
  if (sc->rsqstop > 0 && rsq <= sc->rsqstop) {
    sc->halt = 1;
    sc->converged = 1;
    return;
  }
  int jl = sc->j_low, now = sc->num_now;
  sc->alpha[jl] = rsq / oldrsq;
  for (int j = 0; j < now; j++) {
    if (j == jl) continue;
    if (sc->zeta_i[j] * sc->beta_i[jl] != 0.0)
      sc->alpha[j] = sc->alpha[jl] * sc->zeta_ip1[j] * sc->beta_i[j] / (sc->zeta_i[j] * sc->beta_i[jl]);
    else
      sc->alpha[j] = 0.0;
  }
  // pm_j = zeta_ip1 r + alpha pm_j (:447-453) as window coefficients
  const int m = sc->m;
  for (int j = 0; j < now; j++) {
    double a = sc->alpha[j];
    sc->wa[j] *= a;
    for (int l = 0; l < m - 1; l++) sc->wc[j][l] *= a;
    sc->wc[j][m - 1] = sc->zeta_ip1[j];
  }
  for (int j = 0; j < now; j++) {
    sc->beta_im1[j] = sc->beta_i[j];
    sc->zeta_im1[j] = sc->zeta_i[j];
    sc->zeta_i[j] = sc->zeta_ip1[j];
  }
  // the loop of ks_multicg_offset.c ends when iteration reaches niter
  if (1 + sc->iters >= sc->max_iter) sc->halt = 1;
}

// cg_p = r^(m) + alpha_low cg_p  (pm[j_low] with zeta_low = 1)
template <int P>
__global__ void k_ms_p(VecField p, VecField r, int Vh, const MSScalars* __restrict__ sc) {
  if (sc->halt) return;
  typedef typename PrecTraits<P>::Real R;
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  R a = (R)sc->alpha[sc->j_low];
  Vec3<R> pv, rv;
  load_vec<P>(p, cb, pv);
  load_vec<P>(r, cb, rv);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    Cplx<R> t;
    t.re = (R)1.0 * rv.c[k].re;
    t.im = (R)1.0 * rv.c[k].im;
    pv.c[k].re = t.re + a * pv.c[k].re;
    pv.c[k].im = t.im + a * pv.c[k].im;
  }
  store_vec<P>(p, cb, pv);
}

template <int P>
static int multicg_run(const VecField& src64, std::vector<VecField>& x64, const std::vector<double>& offsets,
                       const MultiParams& mp, SolveStats& st) {
  Context& c = ctx();
  const DevLinks& L = c.links[P];
  if (!L.valid) return -3;
  int n = (int)offsets.size();
  if (n <= 0 || n > MS_MAX) return -2;
  int pb = pbit(mp.parity);
  int rc;
  double source_norm;
  if ((rc = vec_norm2(src64, P_F64, &source_norm))) return rc;
  for (int j = 0; j < n; j++)
    if ((rc = vec_zero(x64[j], P_F64))) return rc;
  st = SolveStats();
  if (source_norm == 0.0) {
    st.converged = 1;
    return 0;
  }
  // tab[0] is r itself; tab[1..MS_KW] the rest of the ring
  VecField tab[MS_TAB_N]{};
  VecField p, r, ttt, tmp;
  MSScalars* d_sc = nullptr;
  VecField* d_tab = nullptr;
  VecScope owned{&p, &r, &ttt, &tmp};
  for (int l = 1; l < MS_RING; l++) owned.add(&tab[l]);
  for (int j = 0; j < n; j++) {
    owned.add(&tab[MS_TAB_PSIM + j]);
    owned.add(&tab[MS_TAB_PM + j]);
  }
  OnExit scalars([&] {
    (void)hipStreamSynchronize(c.stream);
    dev_free(d_sc);
    dev_free(d_tab);
  });
  for (int j = 0; j < n; j++) {
    if ((rc = alloc_vec(tab[MS_TAB_PSIM + j], P, c.stride))) return rc;
    if ((rc = alloc_vec(tab[MS_TAB_PM + j], P, c.stride))) return rc;
    if ((rc = convert_vec(src64, P_F64, tab[MS_TAB_PM + j], P))) return rc;
  }
  if ((rc = alloc_vec(p, P, c.stride)) || (rc = alloc_vec(r, P, c.stride)) || (rc = alloc_vec(ttt, P, c.stride)) ||
      (rc = alloc_vec(tmp, P, c.stride)))
    return rc;
  if ((rc = convert_vec(src64, P_F64, p, P)) || (rc = convert_vec(src64, P_F64, r, P))) return rc;
  tab[0] = r;
  for (int l = 1; l < MS_RING; l++)
    if ((rc = alloc_vec(tab[l], P, c.stride))) return rc;

  // scalars and the field table go up through pinned memory (and come back
  // through it once per window)
  const size_t tab_bytes = sizeof(VecField) * MS_TAB_N;
  char* hp = (char*)pinned_scratch(sizeof(MSScalars) + tab_bytes);
  if (!hp) return -5;
  MSScalars& h = *(MSScalars*)hp;
  memset(&h, 0, sizeof h);
  memcpy(hp + sizeof(MSScalars), tab, tab_bytes);
  double offset_low = 1.0e20;
  int j_low = -1;
  for (int j = 0; j < n; j++) {
    h.shifts[j] = offsets[j];
    if (offsets[j] < offset_low) {
      offset_low = offsets[j];
      j_low = j;
    }
  }
  for (int j = 0; j < n; j++)
    if (j != j_low) h.shifts[j] -= h.shifts[j_low];
  h.n = n;
  h.num_now = n;
  h.j_low = j_low;
  h.source_norm = source_norm;
  h.rsq = source_norm;
  h.rsqstop = mp.resid * mp.resid * source_norm;
  for (int j = 0; j < n; j++) {
    h.zeta_im1[j] = h.zeta_i[j] = 1.0;
    h.beta_im1[j] = -1.0;
    h.alpha[j] = 0.0;
  }
  h.max_iter = (int)std::min<long>(mp.max_iters, 1L << 30);
  // MILC stores shift0 = -shifts[j_low] with shifts[j_low] = offset_low
  if ((rc = dev_alloc((void**)&d_sc, sizeof(MSScalars))) || (rc = dev_alloc((void**)&d_tab, tab_bytes))) return rc;
  MH_CHECK(hipMemcpyAsync(d_sc, &h, sizeof h, hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipMemcpyAsync(d_tab, hp + sizeof(MSScalars), tab_bytes, hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  const int nb = (c.Vh + 255) / 256;
  // one window: reset, MS_KW iterations (ks_multicg_offset.c:291-460), flush
  auto window = [&]() -> int {
    hipLaunchKernelGGL(k_ms_reset, dim3(1), dim3(64), 0, c.stream, d_sc);
    for (int m = 1; m <= MS_KW; m++) {
      hipLaunchKernelGGL(k_ms_dslash1<P>, dim3(nb), dim3(256), 0, c.stream, c.geom, L.fat, L.lng, p, tmp, pb ^ 1,
                         d_sc);
      hipLaunchKernelGGL(k_ms_dslash2<P>, dim3(nb), dim3(256), 0, c.stream, c.geom, L.fat, L.lng, tmp, p, ttt, pb,
                         d_sc, c.ws.partials);
      hipLaunchKernelGGL(k_ms_beta, dim3(1), dim3(1024), 0, c.stream, c.ws.partials, nb, d_sc);
      hipLaunchKernelGGL(k_ms_r<P>, dim3(nb), dim3(256), 0, c.stream, tab[m - 1], tab[m], ttt, c.Vh, d_sc,
                         c.ws.partials);
      hipLaunchKernelGGL(k_ms_alpha, dim3(1), dim3(1024), 0, c.stream, c.ws.partials, nb, d_sc);
      hipLaunchKernelGGL(k_ms_p<P>, dim3(nb), dim3(256), 0, c.stream, p, tab[m], c.Vh, d_sc);
    }
    hipLaunchKernelGGL(k_ms_flush<P>, dim3(nb), dim3(256), 0, c.stream, (const VecField*)d_tab, c.Vh, d_sc);
    MH_CHECK(hipGetLastError());
    return 0;
  };
  HipGraph gr;
  if ((rc = capture_graph(gr, window))) return rc;
  long iteration = 1;  // the initialisation counts as one (ks_multicg_offset.c:246)
  while (iteration < mp.max_iters) {
    MH_CHECK(hipGraphLaunch(gr.e, c.stream));
    MH_CHECK(hipMemcpyAsync(&h, d_sc, sizeof h, hipMemcpyDeviceToHost, c.stream));
    MH_CHECK(hipStreamSynchronize(c.stream));
    iteration = 1 + h.iters;
    if (h.halt) break;
  }
  for (int j = 0; j < n; j++)
    if ((rc = convert_vec(tab[MS_TAB_PSIM + j], P, x64[j], P_F64))) return rc;
  MH_CHECK(hipStreamSynchronize(c.stream));
  st.iters = (int)iteration;
  st.converged = h.converged;
  st.final_rsq = h.rsq / source_norm;
  st.size_r = st.final_rsq;
  return 0;
}

int multicg_solve(const VecField& src, std::vector<VecField>& x, const std::vector<double>& offsets,
                  const MultiParams& mp, SolveStats& st) {
  auto t0 = std::chrono::steady_clock::now();
  int rc = mp.prec == P_F64 ? multicg_run<P_F64>(src, x, offsets, mp, st) : multicg_run<P_F32>(src, x, offsets, mp, st);
  (void)hipStreamSynchronize(ctx().stream);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // namespace mh
