// mh_bsolve.hip -- solvers on batched fields: multi-shift CG for NR sources at
// once and mixed-precision CG refinement, each iteration window replayed from
// a captured hipGraph.
//
// Multi-shift: the recurrences of generic_ks/ks_multicg_offset.c:63-505
// (ks_multicg_offset_field_cpu) run per right-hand side in FP64 on the device
// (zeta/beta/alpha, retirement of trailing shifts, :324-355, :429-444).  The
// vectors run in FP32.  Every shift j carries a solution psim_j and a search
// direction pm_j; updating both every iteration (:390-396, :447-453) streams
// 4 vectors per shift per iteration.  Here the updates are deferred over a
// window of K = BKW iterations: with r^(s+l) the residuals of the window,
//   pm_j^(s+m) = a_m pm_j^(s) + sum_{l<=m} c_{m,l} r^(s+l)
//     a_m = alpha_j a_{m-1},  c_{m,l} = alpha_j c_{m-1,l},  c_{m,m} = zeta_j^(i+1)
//   psim_j^(s+K) = psim_j^(s) + wx pm_j^(s) + sum_l wb_l r^(s+l)
//     wx += beta_j a_{m-1},  wb_l += beta_j c_{m-1,l}
// so psim_j and pm_j are read and written once per window, against the K
// residuals that all shifts share: per iteration and rhs-site the traffic
// of the shift updates falls from 96 n_shift bytes to (24 K + 96 n_shift) / K.
// The coefficients are the same products of the same FP64 scalars that the
// iteration-by-iteration update applies; only the rounding of the FP32
// vector arithmetic differs.
//
// Refinement (the role of ks_multicg_hybrid_field's per-shift CG,
// generic_ks/ks_multicg.c:170-236): per shift, a batched CG with FP32
// iterations (FEWSUMS recurrence, d_congrad5_fn_milc.c:339) and FP64
// reliable updates (true residual b - A x recomputed in FP64 whenever the
// sloppy residual has dropped by delta^2, the search direction corrected by
// r_true - r_sloppy) until the FP64 true residual meets the target.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "mh_batch.h"
#include "mh_bkern.h"

namespace mh {

// ---------------------------------------------------------------------------
// device scalars
// ---------------------------------------------------------------------------
struct BMSR {  // one right-hand side of the multi-shift solve
  double zeta_i[BMS_MAX], zeta_im1[BMS_MAX], zeta_ip1[BMS_MAX];
  double beta_i[BMS_MAX], beta_im1[BMS_MAX], alpha[BMS_MAX];
  double wa[BMS_MAX], wx[BMS_MAX], wc[BMS_MAX][BKW], wb[BMS_MAX][BKW];
  // this rhs's shifts: shifts[j_low] = its smallest offset, the others relative to it
  double shifts[BMS_MAX];
  double rsq, oldrsq, source_norm, rsqstop;
  int num_now, now_ws, m, halt, iters, converged, j_low, pad;
};
struct BMSG {  // shared by all right-hand sides
  int n, nr, halt_all, iters_max, max_iter, pad[3];
};
struct BCGR {
  double rsq, oldrsq, a, b, rsq_stop, pad;
  int halt, go, iters, pad2;
};
struct BCGG {
  int halt_all, iters_max, any_go, iter_cap;  // iter_cap: iterations left to this sloppy phase
};
struct FPtrs {
  float* v[BMS_MAX];
};

// ---------------------------------------------------------------------------
// multi-shift kernels
// ---------------------------------------------------------------------------
__global__ void k_bms_reset(BMSR* __restrict__ s, const BMSG* __restrict__ g) {
  int r = threadIdx.x;
  if (r >= g->nr) return;
  BMSR& S = s[r];
  S.m = 0;
  S.now_ws = S.num_now;
  for (int j = 0; j < g->n; j++) {
    S.wa[j] = 1.0;
    S.wx[j] = 0.0;
    for (int l = 0; l < BKW; l++) S.wc[j][l] = S.wb[j][l] = 0.0;
  }
}

// beta/zeta recurrences (ks_multicg_offset.c:322-356) and the solution-update
// coefficients of this iteration
__global__ void k_bms_beta(const double* __restrict__ sums, BMSR* __restrict__ s, BMSG* __restrict__ g) {
  int r = threadIdx.x;
  if (g->halt_all) return;
  if (r < g->nr && !s[r].halt) {
    BMSR& S = s[r];
    const int jl = S.j_low;
    double pkp = sums[r];
    S.iters += 1;
    int m = S.m + 1;
    S.m = m;
    S.beta_i[jl] = -S.rsq / pkp;
    S.zeta_ip1[jl] = 1.0;
    int now = S.num_now;
    for (int j = 0; j < now; j++) {
      if (j == jl) continue;
      S.zeta_ip1[j] = S.zeta_i[j] * S.zeta_im1[j] * S.beta_im1[jl];
      double c1 = S.beta_i[jl] * S.alpha[jl] * (S.zeta_im1[j] - S.zeta_i[j]);
      double c2 = S.zeta_im1[j] * S.beta_im1[jl] * (1.0 + S.shifts[j] * S.beta_i[jl]);
      if (c1 + c2 != 0.0) S.zeta_ip1[j] /= c1 + c2;
      else S.zeta_ip1[j] = 0.0;
      if (S.zeta_i[j] != 0.0) {
        S.beta_i[j] = S.beta_i[jl] * S.zeta_ip1[j] / S.zeta_i[j];
      } else {
        S.zeta_ip1[j] = 0.0;
        S.beta_i[j] = 0.0;
        if (j == now - 1 && j > jl) now--;
      }
    }
    S.num_now = now;
    // psim_j += beta_j pm_j^(s+m-1), pm_j^(s+m-1) = wa pm_j^(s) + sum_{l<m-1} wc_l r^(s+1+l)
    for (int j = 0; j < now; j++) {
      double b = S.beta_i[j];
      S.wx[j] += b * S.wa[j];
      for (int l = 0; l < m - 1; l++) S.wb[j][l] += b * S.wc[j][l];
    }
  }
}

// convergence, alpha recurrences (ks_multicg_offset.c:358-463), search-direction
// coefficients of this iteration, scalar scroll
__global__ void k_bms_alpha(const double* __restrict__ sums, BMSR* __restrict__ s, BMSG* __restrict__ g) {
  int r = threadIdx.x;
  if (g->halt_all) return;
  if (r < g->nr && !s[r].halt) {
    BMSR& S = s[r];
    const int jl = S.j_low;
    double rsq = sums[r];
    S.oldrsq = S.rsq;
    S.rsq = rsq;
    if (S.rsqstop > 0 && rsq <= S.rsqstop) {
      S.halt = 1;
      S.converged = 1;
    } else {
      int now = S.num_now, m = S.m;
      S.alpha[jl] = rsq / S.oldrsq;
      for (int j = 0; j < now; j++) {
        if (j == jl) continue;
        if (S.zeta_i[j] * S.beta_i[jl] != 0.0)
          S.alpha[j] = S.alpha[jl] * S.zeta_ip1[j] * S.beta_i[j] / (S.zeta_i[j] * S.beta_i[jl]);
        else
          S.alpha[j] = 0.0;
      }
      // pm_j = zeta_ip1 r + alpha pm_j
      for (int j = 0; j < now; j++) {
        double a = S.alpha[j];
        S.wa[j] *= a;
        for (int l = 0; l < m - 1; l++) S.wc[j][l] *= a;
        S.wc[j][m - 1] = S.zeta_ip1[j];
      }
      for (int j = 0; j < now; j++) {
        S.beta_im1[j] = S.beta_i[j];
        S.zeta_im1[j] = S.zeta_i[j];
        S.zeta_i[j] = S.zeta_ip1[j];
      }
      // the loop of ks_multicg_offset.c ends when iteration reaches niter
      if (1 + S.iters >= g->max_iter) S.halt = 1;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1, im = 0;
    for (int q = 0; q < g->nr; q++) {
      all &= s[q].halt;
      im = max(im, s[q].iters);
    }
    g->halt_all = all;
    g->iters_max = im;
  }
}

// r^(m) = r^(m-1) + beta_low ttt; |r^(m)|^2 per rhs
// Kernels over (site, rhs) records take `loop` records per thread, strided by
// the grid; the stride is a multiple of NR, so a thread keeps one rhs.
#define REC_LOOP(idx)                                                         \
  const int gid_ = blockIdx.x * blockDim.x + threadIdx.x;                     \
  const long step_ = (long)gridDim.x * blockDim.x;                            \
  for (long idx = gid_, it_ = 0; it_ < loop; it_++, idx += step_)

template <int NR>
__global__ void __launch_bounds__(256) k_bms_r(const float* __restrict__ rprev, float* __restrict__ rcur,
                                               const float* __restrict__ ttt, const BMSR* __restrict__ s,
                                               const BMSG* __restrict__ g, long nrec, int loop,
                                               double* __restrict__ partials, int pstride) {
  if (g->halt_all) return;
  const int rhs = (blockIdx.x * blockDim.x + threadIdx.x) % NR;
  const bool live = !s[rhs].halt;
  const float b = (float)s[rhs].beta_i[s[rhs].j_low];
  double d[1] = {0.0};
  REC_LOOP(idx) {
    if (idx >= nrec || !live) continue;
    size_t o = (size_t)idx * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      float v = rprev[o + k] + b * ttt[o + k];
      rcur[o + k] = v;
      d[0] += (double)v * v;
    }
  }
  bsum_lanes<1, NR>(d, partials, pstride);
}

// cg_p = r^(m) + alpha_low cg_p
template <int NR>
__global__ void __launch_bounds__(256) k_bms_p(float* __restrict__ p, const float* __restrict__ rcur,
                                               const BMSR* __restrict__ s, const BMSG* __restrict__ g, long nrec,
                                               int loop) {
  if (g->halt_all) return;
  const int rhs = (blockIdx.x * blockDim.x + threadIdx.x) % NR;
  if (s[rhs].halt) return;
  const float a = (float)s[rhs].alpha[s[rhs].j_low];
  REC_LOOP(idx) {
    if (idx >= nrec) continue;
    size_t o = (size_t)idx * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) p[o + k] = rcur[o + k] + a * p[o + k];
  }
}

// window coefficients as floats: cf[(r * n + j) * NC + {wx, wb[K], wa, wc[K]}]
#define BMS_NC (2 + 2 * BKW)
__global__ void k_bms_coef(const BMSR* __restrict__ s, const BMSG* __restrict__ g, float* __restrict__ cf,
                           int* __restrict__ meta) {
  const int n = g->n, nr = g->nr;
  for (int t = threadIdx.x; t < nr * n; t += blockDim.x) {
    int r = t / n, j = t % n;
    const BMSR& S = s[r];
    float* c = cf + (size_t)t * BMS_NC;
    c[0] = (float)S.wx[j];
    c[1 + BKW] = (float)S.wa[j];
    for (int l = 0; l < BKW; l++) {
      c[1 + l] = (float)S.wb[j][l];
      c[2 + BKW + l] = (float)S.wc[j][l];
    }
  }
  if (threadIdx.x < nr) {
    meta[2 * threadIdx.x] = s[threadIdx.x].m;
    meta[2 * threadIdx.x + 1] = s[threadIdx.x].now_ws;
  }
}

// Window flush: psim_j and pm_j brought up to date from the ring of residuals,
//   psim_j += wx p_j + sum_l wb_l r^(l+1),   pm_j = wa p_j + sum_l wc_l r^(l+1)
// over the md slots of the window (coefficients of the rhs from k_bms_coef;
// slots past md count as zero), then ring slot 0 takes the last slot.  The
// field is walked as a flat float array, one float4 per thread and pass, so
// every access is a whole 16-byte word and a wavefront's words are
// contiguous (48^3x96, 4 rhs, 9 shifts: 6.0 ms against 7.0 ms for one thread
// per 6-float record with 8-byte accesses; same results bit for bit).  A
// record is 6 floats, so a float4 holds parts of one or two records (rhs ra
// and rb, the next one); each component takes its own record's coefficients
// and window length.
template <int NR>
__global__ void __launch_bounds__(256) k_bms_flush4(float* __restrict__ ring, size_t ring_stride, FPtrs pm,
                                                    FPtrs psim, const float* __restrict__ cf,
                                                    const int* __restrict__ meta, int n, long nq, int loop) {
  REC_LOOP(q) {
    if (q >= nq) continue;
    const long f0 = q * 4;
    const long ra_rec = f0 / 6;
    const int ra = (int)(ra_rec % NR), rb = (int)((ra_rec + 1) % NR);
    const int na = (int)(ra_rec * 6 + 6 - f0);  // components u < na belong to record ra (na = 2, 4 or 6)
    const int mda = meta[2 * ra], nna = meta[2 * ra + 1];
    const int mdb = na < 4 ? meta[2 * rb] : 0, nnb = na < 4 ? meta[2 * rb + 1] : 0;
    int md[4];
#pragma unroll
    for (int u = 0; u < 4; u++) md[u] = u < na ? mda : mdb;
    const int mdmax = max(mda, mdb);
    if (mdmax == 0) continue;
    float rv[BKW][4];
#pragma unroll
    for (int l = 0; l < BKW; l++) {
      const float4 t = l < mdmax ? *(const float4*)(ring + (size_t)(l + 1) * ring_stride + f0)
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
      rv[l][0] = l < md[0] ? t.x : 0.f;
      rv[l][1] = l < md[1] ? t.y : 0.f;
      rv[l][2] = l < md[2] ? t.z : 0.f;
      rv[l][3] = l < md[3] ? t.w : 0.f;
    }
    const int nnmax = max(mda > 0 ? nna : 0, mdb > 0 ? nnb : 0);
    for (int j = 0; j < nnmax; j++) {
      const bool la = mda > 0 && j < nna, lb = mdb > 0 && j < nnb;
      float ca[BMS_NC], cb[BMS_NC];
      const float2* c2a = (const float2*)(cf + ((size_t)ra * n + j) * BMS_NC);
      const float2* c2b = (const float2*)(cf + ((size_t)rb * n + j) * BMS_NC);
#pragma unroll
      for (int l = 0; l < BMS_NC / 2; l++) {
        const float2 ta = la ? c2a[l] : make_float2(0.f, 0.f);
        const float2 tb = lb ? c2b[l] : make_float2(0.f, 0.f);
        ca[2 * l] = ta.x;
        ca[2 * l + 1] = ta.y;
        cb[2 * l] = tb.x;
        cb[2 * l + 1] = tb.y;
      }
      float4* pmj = (float4*)(pm.v[j] + f0);
      float4* xj = (float4*)(psim.v[j] + f0);
      const float4 p4 = *pmj, x4 = *xj;
      const float pv[4] = {p4.x, p4.y, p4.z, p4.w};
      float xv[4] = {x4.x, x4.y, x4.z, x4.w}, qv[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const bool a = u < na;
        if (!(a ? la : lb)) continue;
        const float* c = a ? ca : cb;
        float x = xv[u] + c[0] * pv[u];
        float qq = c[1 + BKW] * pv[u];
#pragma unroll
        for (int l = 0; l < BKW; l++) {
          x += c[1 + l] * rv[l][u];
          qq += c[2 + BKW + l] * rv[l][u];
        }
        xv[u] = x;
        qv[u] = qq;
      }
      *xj = make_float4(xv[0], xv[1], xv[2], xv[3]);
      *pmj = make_float4(qv[0], qv[1], qv[2], qv[3]);
    }
    // ring slot 0 takes the last slot of the window, per record
    float4* r0 = (float4*)(ring + f0);
    float4 o = (mda > 0 && (na >= 4 || mdb > 0)) ? make_float4(0.f, 0.f, 0.f, 0.f) : *r0;
    float ov[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int l = 0; l < BKW; l++)
        if (l + 1 == md[u]) ov[u] = rv[l][u];
    *r0 = make_float4(ov[0], ov[1], ov[2], ov[3]);
  }
}

// ---------------------------------------------------------------------------
// batched BLAS
// ---------------------------------------------------------------------------
template <typename R, int NR>
__global__ void __launch_bounds__(256) k_bnorm(const R* __restrict__ f, long nrec, int loop,
                                               double* __restrict__ partials, int pstride) {
  double d[1] = {0.0};
  REC_LOOP(idx) {
    if (idx >= nrec) continue;
    size_t o = (size_t)idx * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) d[0] += (double)f[o + k] * (double)f[o + k];
  }
  bsum_lanes<1, NR>(d, partials, pstride);
}

template <typename RI, typename RO>
__global__ void k_bconv(const RI* __restrict__ in, RO* __restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (RO)in[i];
}

// x64 += xs; xs = 0
__global__ void k_baccum(double* __restrict__ x, float* __restrict__ xs, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] += (double)xs[i];
    xs[i] = 0.f;
  }
}

// r64 = b + t (t = -A x); |r64|^2 per rhs
template <int NR>
__global__ void __launch_bounds__(256) k_bresid(const double* __restrict__ b, const double* __restrict__ t,
                                                double* __restrict__ r, long nrec, int loop,
                                                double* __restrict__ partials, int pstride) {
  double d[1] = {0.0};
  REC_LOOP(idx) {
    if (idx >= nrec) continue;
    size_t o = (size_t)idx * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      double v = b[o + k] + t[o + k];
      r[o + k] = v;
      d[0] += v * v;
    }
  }
  bsum_lanes<1, NR>(d, partials, pstride);
}

// reliable update of the sloppy vectors of unconverged rhs:
// first: r_s = r64, p_s = r64; later: p_s += r64 - r_s, r_s = r64.  |r_s|^2 per rhs.
template <int NR>
__global__ void __launch_bounds__(256) k_bru(const double* __restrict__ r64, float* __restrict__ rs,
                                             float* __restrict__ ps, const BCGR* __restrict__ s, int first, long nrec,
                                             int loop, double* __restrict__ partials, int pstride) {
  const int rhs = (blockIdx.x * blockDim.x + threadIdx.x) % NR;
  const bool live = !s[rhs].halt;
  double d[1] = {0.0};
  REC_LOOP(idx) {
    if (idx >= nrec || !live) continue;
    size_t o = (size_t)idx * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      double rh = r64[o + k];
      float rn = (float)rh;
      ps[o + k] = first ? rn : (float)((double)ps[o + k] + (rh - (double)rs[o + k]));
      rs[o + k] = rn;
      d[0] += (double)rn * rn;
    }
  }
  bsum_lanes<1, NR>(d, partials, pstride);
}

// FEWSUMS CG scalars per rhs (d_congrad5_fn_milc.c:318-339)
__global__ void k_bcg_scalars(const double* __restrict__ sums, const double* __restrict__ sums_rr, int nr,
                              BCGR* __restrict__ s, BCGG* __restrict__ g) {
  int r = threadIdx.x;
  if (g->halt_all) {
    // every rhs stopped in an earlier iteration: no update may run
    if (r < nr) s[r].go = 0;
    if (r == 0) g->any_go = 0;
    return;
  }
  if (r < nr) {
    BCGR& S = s[r];
    if (S.halt) {
      S.go = 0;
    } else {
      double pkp = sums[r], c_tr = sums[nr + r], c_tt = sums[2 * nr + r], oldrsq = sums_rr[r];
      double a = -S.rsq / pkp;
      double rsq = oldrsq + 2.0 * a * c_tr + a * a * c_tt;
      S.oldrsq = oldrsq;
      S.a = a;
      S.b = rsq / oldrsq;
      S.rsq = rsq;
      S.go = 1;
      S.iters += 1;
      if (rsq < S.rsq_stop || S.iters >= g->iter_cap) S.halt = 1;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1, im = 0, go = 0;
    for (int q = 0; q < nr; q++) {
      all &= s[q].halt;
      im = max(im, s[q].iters);
      go |= s[q].go;
    }
    g->halt_all = all;
    g->iters_max = im;
    g->any_go = go;
  }
}

// x_s += a p; r += a ttt; p = r + b p; |r|^2 per rhs
template <int NR>
__global__ void __launch_bounds__(256) k_bcg_update(float* __restrict__ x, float* __restrict__ p, float* __restrict__ r,
                                                    const float* __restrict__ ttt, const BCGR* __restrict__ s,
                                                    const BCGG* __restrict__ g, long nrec, int loop,
                                                    double* __restrict__ partials, int pstride) {
  if (!g->any_go) return;
  const int rhs = (blockIdx.x * blockDim.x + threadIdx.x) % NR;
  const bool live = s[rhs].go != 0;
  const float a = (float)s[rhs].a, b = (float)s[rhs].b;
  double d[1] = {0.0};
  REC_LOOP(idx) {
    if (idx >= nrec || !live) continue;
    size_t o = (size_t)idx * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      float pv = p[o + k];
      x[o + k] += a * pv;
      float rn = r[o + k] + a * ttt[o + k];
      r[o + k] = rn;
      p[o + k] = rn + b * pv;
      d[0] += (double)rn * rn;
    }
  }
  bsum_lanes<1, NR>(d, partials, pstride);
}

// The same with the search direction p in 16-bit records.
template <int NR>
__global__ void __launch_bounds__(256) k_bcg_update_h16(float* __restrict__ x, int4* __restrict__ p,
                                                        float* __restrict__ r, const float* __restrict__ ttt,
                                                        const BCGR* __restrict__ s, const BCGG* __restrict__ g,
                                                        long nrec, int loop, double* __restrict__ partials,
                                                        int pstride) {
  if (!g->any_go) return;
  const int rhs = (blockIdx.x * blockDim.x + threadIdx.x) % NR;
  const bool live = s[rhs].go != 0;
  const float a = (float)s[rhs].a, b = (float)s[rhs].b;
  double d[1] = {0.0};
  REC_LOOP(idx) {
    if (idx >= nrec || !live) continue;
    size_t o = (size_t)idx * 6;
    Vec3<float> pv, pn;
    h16_unpack(p[idx], pv);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      x[o + k] += a * pv.c[k].re;
      x[o + 3 + k] += a * pv.c[k].im;
      const float rre = r[o + k] + a * ttt[o + k], rim = r[o + 3 + k] + a * ttt[o + 3 + k];
      r[o + k] = rre;
      r[o + 3 + k] = rim;
      pn.c[k].re = rre + b * pv.c[k].re;
      pn.c[k].im = rim + b * pv.c[k].im;
      d[0] += (double)rre * rre + (double)rim * rim;
    }
    p[idx] = h16_pack(pn);
  }
  bsum_lanes<1, NR>(d, partials, pstride);
}

// Reliable update with a 16-bit search direction: r_s = r (FP64 true residual
// rounded), p = r_s on the first call, else p += r - r_s(old).
template <int NR>
__global__ void __launch_bounds__(256) k_bru_h16(const double* __restrict__ r64, float* __restrict__ rs,
                                                 int4* __restrict__ ps, const BCGR* __restrict__ s, int first,
                                                 long nrec, int loop, double* __restrict__ partials, int pstride) {
  const int rhs = (blockIdx.x * blockDim.x + threadIdx.x) % NR;
  const bool live = !s[rhs].halt;
  double d[1] = {0.0};
  REC_LOOP(idx) {
    if (idx >= nrec || !live) continue;
    size_t o = (size_t)idx * 6;
    Vec3<float> pv;
    if (!first) h16_unpack(ps[idx], pv);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double hre = r64[o + k], him = r64[o + 3 + k];
      const float nre = (float)hre, nim = (float)him;
      if (first) {
        pv.c[k].re = nre;
        pv.c[k].im = nim;
      } else {
        pv.c[k].re = (float)((double)pv.c[k].re + (hre - (double)rs[o + k]));
        pv.c[k].im = (float)((double)pv.c[k].im + (him - (double)rs[o + 3 + k]));
      }
      rs[o + k] = nre;
      rs[o + 3 + k] = nim;
      d[0] += (double)nre * nre + (double)nim * nim;
    }
    ps[idx] = h16_pack(pv);
  }
  bsum_lanes<1, NR>(d, partials, pstride);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { dev_free(p); }
};

// grid of a (site, rhs)-record kernel: at most `cap` blocks of 256 threads
struct RecGrid {
  int nb, loop;
  long nrec;
};
RecGrid rec_grid(int nr, int cap) {
  RecGrid g;
  g.nrec = (long)ctx().Vh * nr;
  long b1 = (g.nrec + 255) / 256;
  g.loop = 1;
  while ((b1 + g.loop - 1) / g.loop > cap) g.loop *= 2;
  g.nb = (int)((b1 + g.loop - 1) / g.loop);
  return g;
}

// per-rhs |v|^2 of a batched field (synchronous)
int bnorms(const void* f, int prec, int nr, double* sums_dev, double* out) {
  Context& c = ctx();
  double* part;
  int ps, rc;
  if ((rc = bpartials(&part, &ps))) return rc;
  RecGrid rg = rec_grid(nr, ps);
#define BN_CASE(R, N)                                                                                      \
  hipLaunchKernelGGL((k_bnorm<R, N>), dim3(rg.nb), dim3(256), 0, c.stream, (const R*)f, rg.nrec, rg.loop, part, \
                     ps)
  if (prec == P_F64) {
    if (nr == 4) BN_CASE(double, 4); else if (nr == 8) BN_CASE(double, 8); else BN_CASE(double, 16);
  } else {
    if (nr == 4) BN_CASE(float, 4); else if (nr == 8) BN_CASE(float, 8); else BN_CASE(float, 16);
  }
#undef BN_CASE
  MH_CHECK(hipGetLastError());
  if ((rc = breduce(nr, rg.nb, sums_dev))) return rc;
  MH_CHECK(hipMemcpyAsync(out, sums_dev, nr * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

template <typename RI, typename RO>
int bconv(const void* in, void* out, int nr) {
  Context& c = ctx();
  size_t n = (size_t)c.Vh * nr * 6;
  hipLaunchKernelGGL((k_bconv<RI, RO>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, (const RI*)in,
                     (RO*)out, n);
  MH_CHECK(hipGetLastError());
  return 0;
}

// offs[j * NR + r]: offset of shift j for rhs r; resid[r]: base-system target of rhs r
template <int NR>
int bmultishift_nr(const void* b64, const double* offs, int n, const double* resid, const BMSParams& prm,
                   std::vector<void*>& x64, BMSStats& st) {
  Context& c = ctx();
  if (n <= 0 || n > BMS_MAX || (int)x64.size() != n) return -2;
  const int pb = pbit(prm.parity);
  int rc;
  double* part;
  int pstride;
  if ((rc = bpartials(&part, &pstride))) return rc;
  const RecGrid rg = rec_grid(NR, pstride);
  // the flush walks float4 words: 6 NR / 4 of them per site
  const RecGrid rf = rec_grid(NR * 6 / 4, 2048);
  int variant;
  if ((rc = btune(P_F32, NR, &variant, nullptr))) return rc;

  // fields; the ring of residuals is one allocation of BKW + 1 slots
  std::vector<DevBuf> psim(n), pm(n);
  DevBuf ringbuf, p, tmp, ttt, sums, dscal, dglob, dcoef, dmeta;
  const size_t fb = bvec_bytes(P_F32, NR);
  if ((rc = dev_alloc(&ringbuf.p, fb * (BKW + 1)))) return rc;
  MH_CHECK(hipMemsetAsync(ringbuf.p, 0, fb * (BKW + 1), c.stream));
  float* ring0 = (float*)ringbuf.p;
  const size_t ring_stride = fb / sizeof(float);
  for (int j = 0; j < n; j++)
    if ((rc = balloc(&psim[j].p, P_F32, NR)) || (rc = balloc(&pm[j].p, P_F32, NR))) return rc;
  if ((rc = balloc(&p.p, P_F32, NR)) || (rc = balloc(&tmp.p, P_F32, NR)) || (rc = balloc(&ttt.p, P_F32, NR))) return rc;
  if ((rc = dev_alloc(&sums.p, 64 * sizeof(double))) || (rc = dev_alloc(&dscal.p, NR * sizeof(BMSR))) ||
      (rc = dev_alloc(&dglob.p, sizeof(BMSG))) ||
      (rc = dev_alloc(&dcoef.p, (size_t)NR * BMS_MAX * BMS_NC * sizeof(float))) ||
      (rc = dev_alloc(&dmeta.p, 2 * NR * sizeof(int))))
    return rc;

  // source norms; FP32 start vectors r = p = pm_j = b, psim_j = 0 (ks_multicg_offset.c:229-262)
  double norms[BNR_MAX];
  if ((rc = bnorms(b64, P_F64, NR, (double*)sums.p, norms))) return rc;
  if ((rc = bconv<double, float>(b64, ring0, NR)) || (rc = bconv<double, float>(b64, p.p, NR))) return rc;
  for (int j = 0; j < n; j++)
    if ((rc = bconv<double, float>(b64, pm[j].p, NR))) return rc;

  std::vector<BMSR> hs(NR);
  BMSG hg;
  memset(hs.data(), 0, NR * sizeof(BMSR));
  memset(&hg, 0, sizeof hg);
  hg.n = n;
  hg.nr = NR;
  hg.max_iter = (int)std::min<long>(prm.max_iters, 1L << 30);
  int all = 1;
  double base[BNR_MAX];
  for (int r = 0; r < NR; r++) {
    BMSR& S = hs[r];
    int jl = 0;
    for (int j = 1; j < n; j++)
      if (offs[j * NR + r] < offs[jl * NR + r]) jl = j;
    base[r] = offs[jl * NR + r];
    for (int j = 0; j < n; j++) S.shifts[j] = j == jl ? base[r] : offs[j * NR + r] - base[r];
    S.j_low = jl;
    S.num_now = n;
    S.source_norm = norms[r];
    S.rsq = norms[r];
    S.rsqstop = resid[r] * resid[r] * norms[r];
    for (int j = 0; j < n; j++) {
      S.zeta_im1[j] = S.zeta_i[j] = 1.0;
      S.beta_im1[j] = -1.0;
      S.alpha[j] = 0.0;
    }
    if (norms[r] == 0.0) S.halt = S.converged = 1;
    all &= S.halt;
  }
  hg.halt_all = all;
  MH_CHECK(hipMemcpyAsync(dscal.p, hs.data(), NR * sizeof(BMSR), hipMemcpyHostToDevice, c.stream));
  MH_CHECK(hipMemcpyAsync(dglob.p, &hg, sizeof hg, hipMemcpyHostToDevice, c.stream));
  BMSR* ds = (BMSR*)dscal.p;
  BMSG* dg = (BMSG*)dglob.p;
  double* dsum = (double*)sums.p;
  FPtrs fpm{}, fpsim{};
  for (int j = 0; j < n; j++) {
    fpm.v[j] = (float*)pm[j].p;
    fpsim.v[j] = (float*)psim[j].p;
  }
  // one window: reset, K iterations (ks_multicg_offset.c:291-460), flush
  auto window = [&]() -> int {
    int r2;
    hipLaunchKernelGGL(k_bms_reset, dim3(1), dim3(64), 0, c.stream, ds, dg);
    for (int m = 1; m <= BKW; m++) {
      float* rprev = ring0 + (size_t)(m - 1) * ring_stride;
      float* rcur = ring0 + (size_t)m * ring_stride;
      BDOp d1;
      d1.prec = P_F32;
      d1.nr = NR;
      d1.kind = BK_STORE;
      d1.pb = pb ^ 1;
      d1.src = p.p;
      d1.dst = tmp.p;
      d1.halt_all = &dg->halt_all;
      d1.variant = variant;
      if ((r2 = bdslash(d1, nullptr))) return r2;
      BDOp d2 = d1;
      d2.kind = BK_MS;
      d2.pb = pb;
      d2.src = tmp.p;
      d2.dst = ttt.p;
      d2.p = p.p;
      for (int r = 0; r < NR; r++) d2.coef[r] = -base[r];
      int nb2 = 0;
      if ((r2 = bdslash(d2, &nb2))) return r2;
      if ((r2 = breduce(NR, nb2, dsum))) return r2;
      hipLaunchKernelGGL(k_bms_beta, dim3(1), dim3(64), 0, c.stream, dsum, ds, dg);
      hipLaunchKernelGGL(k_bms_r<NR>, dim3(rg.nb), dim3(256), 0, c.stream, rprev, rcur, (const float*)ttt.p, ds, dg,
                         rg.nrec, rg.loop, part, pstride);
      if ((r2 = breduce(NR, rg.nb, dsum + 32))) return r2;
      hipLaunchKernelGGL(k_bms_alpha, dim3(1), dim3(64), 0, c.stream, dsum + 32, ds, dg);
      hipLaunchKernelGGL(k_bms_p<NR>, dim3(rg.nb), dim3(256), 0, c.stream, (float*)p.p, rcur, ds, dg, rg.nrec,
                         rg.loop);
    }
    hipLaunchKernelGGL(k_bms_coef, dim3(1), dim3(256), 0, c.stream, ds, dg, (float*)dcoef.p, (int*)dmeta.p);
    hipLaunchKernelGGL(k_bms_flush4<NR>, dim3(rf.nb), dim3(256), 0, c.stream, ring0, ring_stride, fpm, fpsim,
                       (const float*)dcoef.p, (const int*)dmeta.p, n, rf.nrec, rf.loop);
    MH_CHECK(hipGetLastError());
    return 0;
  };
  auto t0 = std::chrono::steady_clock::now();
  HipGraph gr;
  if ((rc = capture_graph(gr, window))) return rc;
  BMSG* hgp = nullptr;
  MH_CHECK(hipHostMalloc((void**)&hgp, sizeof(BMSG), hipHostMallocDefault));
  *hgp = hg;
  long iteration = 1;  // the initialisation counts as one (ks_multicg_offset.c:246)
  while (!hgp->halt_all && iteration < prm.max_iters) {
    rc = (int)hipGraphLaunch(gr.e, c.stream);
    if (!rc) rc = (int)hipMemcpyAsync(hgp, dg, sizeof(BMSG), hipMemcpyDeviceToHost, c.stream);
    if (!rc) rc = (int)hipStreamSynchronize(c.stream);
    if (rc) break;
    iteration = 1 + hgp->iters_max;
  }
  (void)hipHostFree(hgp);
  if (rc) return -1000 - rc;
  st.ms_iters = (int)iteration;
  MH_CHECK(hipMemcpyAsync(hs.data(), ds, NR * sizeof(BMSR), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  int conv = 1;
  for (int r = 0; r < NR; r++) conv &= hs[r].converged;
  st.converged = conv;
  for (int j = 0; j < n; j++)
    if ((rc = bconv<float, double>(psim[j].p, x64[j], NR))) return rc;
  MH_CHECK(hipStreamSynchronize(c.stream));
  st.ms_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// FP64 true residual r64 = b - A x (A = -D_eo D_oe + offset), |r64|^2 per rhs into out
template <int NR>
int btrue_residual(const void* b64, const void* x64, void* tmp64, void* t64, void* r64, const double* offsets, int pb,
                   double* sums_dev, double* out) {
  Context& c = ctx();
  int rc;
  double* part;
  int ps;
  if ((rc = bpartials(&part, &ps))) return rc;
  BDOp d1;
  d1.prec = P_F64;
  d1.nr = NR;
  d1.kind = BK_STORE;
  d1.pb = pb ^ 1;
  d1.src = x64;
  d1.dst = tmp64;
  if ((rc = bdslash(d1, nullptr))) return rc;
  BDOp d2 = d1;
  d2.kind = BK_MS;
  d2.pb = pb;
  d2.src = tmp64;
  d2.dst = t64;
  d2.p = x64;
  for (int r = 0; r < NR; r++) d2.coef[r] = -offsets[r];
  if ((rc = bdslash(d2, nullptr))) return rc;
  RecGrid rg = rec_grid(NR, ps);
  hipLaunchKernelGGL(k_bresid<NR>, dim3(rg.nb), dim3(256), 0, c.stream, (const double*)b64, (const double*)t64,
                     (double*)r64, rg.nrec, rg.loop, part, ps);
  MH_CHECK(hipGetLastError());
  if ((rc = breduce(NR, rg.nb, sums_dev))) return rc;
  MH_CHECK(hipMemcpyAsync(out, sums_dev, NR * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

template <int NR>
int bcg_mixed_nr(const void* b64, void* x64, const double* offsets, const BCGParams& prm, BCGStats& st) {
  Context& c = ctx();
  const int pb = pbit(prm.parity);
  int rc;
  double* part;
  int pstride;
  if ((rc = bpartials(&part, &pstride))) return rc;
  const RecGrid rg = rec_grid(NR, pstride);
  const int sp = prm.sloppy == P_H16 ? P_H16 : P_F32;  // operator / search-direction precision
  if (sp == P_H16 && (rc = ensure_links(P_H16))) return rc;
  int vs, v64;
  if ((rc = btune(P_F64, NR, &v64, nullptr)) || (rc = btune(sp, NR, &vs, nullptr))) return rc;
  DevBuf tmp64, t64, r64, ps, tmps, ttts, rs, xs, sums, dscal, dglob;
  if ((rc = balloc(&tmp64.p, P_F64, NR)) || (rc = balloc(&t64.p, P_F64, NR)) || (rc = balloc(&r64.p, P_F64, NR)) ||
      (rc = balloc(&ps.p, sp, NR)) || (rc = balloc(&tmps.p, sp, NR)) || (rc = balloc(&ttts.p, P_F32, NR)) ||
      (rc = balloc(&rs.p, P_F32, NR)) || (rc = balloc(&xs.p, P_F32, NR)))
    return rc;
  if ((rc = dev_alloc(&sums.p, 128 * sizeof(double))) || (rc = dev_alloc(&dscal.p, NR * sizeof(BCGR))) ||
      (rc = dev_alloc(&dglob.p, sizeof(BCGG))))
    return rc;
  double* dsum = (double*)sums.p;  // [0, 3 NR): CG dots; [64, 64 + NR): |r|^2
  double* dsum_rr = dsum + 64;
  BCGR* ds = (BCGR*)dscal.p;
  BCGG* dg = (BCGG*)dglob.p;
  double bn[BNR_MAX], rt[BNR_MAX];
  if ((rc = bnorms(b64, P_F64, NR, dsum, bn))) return rc;
  std::vector<BCGR> hs(NR);
  memset(hs.data(), 0, NR * sizeof(BCGR));
  BCGG hg0;
  memset(&hg0, 0, sizeof hg0);
  const double delta = prm.delta > 0 ? prm.delta : kBatchReliableDelta;
  // K sloppy iterations (d_congrad5_fn_milc.c:281-367, FEWSUMS)
  auto iterations = [&]() -> int {
    int r2;
    for (int k = 0; k < BKW; k++) {
      BDOp d1;
      d1.prec = sp;
      d1.nr = NR;
      d1.kind = BK_STORE;
      d1.pb = pb ^ 1;
      d1.src = ps.p;
      d1.dst = tmps.p;
      d1.halt_all = &dg->halt_all;
      d1.variant = vs;
      if ((r2 = bdslash(d1, nullptr))) return r2;
      BDOp d2 = d1;
      d2.kind = BK_CG;
      d2.pb = pb;
      d2.src = tmps.p;
      d2.dst = ttts.p;
      d2.p = ps.p;
      d2.r = rs.p;
      for (int r = 0; r < NR; r++) d2.coef[r] = -offsets[r];
      int nb2 = 0;
      if ((r2 = bdslash(d2, &nb2))) return r2;
      if ((r2 = breduce(3 * NR, nb2, dsum))) return r2;
      hipLaunchKernelGGL(k_bcg_scalars, dim3(1), dim3(64), 0, c.stream, dsum, dsum_rr, NR, ds, dg);
      if (sp == P_H16)
        hipLaunchKernelGGL(k_bcg_update_h16<NR>, dim3(rg.nb), dim3(256), 0, c.stream, (float*)xs.p, (int4*)ps.p,
                           (float*)rs.p, (const float*)ttts.p, ds, dg, rg.nrec, rg.loop, part, pstride);
      else
        hipLaunchKernelGGL(k_bcg_update<NR>, dim3(rg.nb), dim3(256), 0, c.stream, (float*)xs.p, (float*)ps.p,
                           (float*)rs.p, (const float*)ttts.p, ds, dg, rg.nrec, rg.loop, part, pstride);
      if ((r2 = breduce(NR, rg.nb, dsum_rr))) return r2;
    }
    MH_CHECK(hipGetLastError());
    return 0;
  };
  auto t0 = std::chrono::steady_clock::now();
  HipGraph gr;
  BCGG* hg = nullptr;
  MH_CHECK(hipHostMalloc((void**)&hg, sizeof(BCGG), hipHostMallocDefault));
  int converged[BNR_MAX] = {};
  long iteration = 0;
  int nru = 0;
  bool first = true;
  const size_t nel = (size_t)c.Vh * NR * 6;
  for (;;) {
    // reliable update: x += x_s, FP64 true residual
    if (!first) {
      hipLaunchKernelGGL(k_baccum, dim3((unsigned)((nel + 255) / 256)), dim3(256), 0, c.stream, (double*)x64,
                         (float*)xs.p, nel);
      if ((rc = (int)hipGetLastError())) break;
    }
    if ((rc = btrue_residual<NR>(b64, x64, tmp64.p, t64.p, r64.p, offsets, pb, dsum, rt))) break;
    iteration++;
    int all = 1;
    for (int r = 0; r < NR; r++) {
      const double res = prm.resid_r ? prm.resid_r[r] : prm.resid;
      converged[r] = bn[r] == 0.0 || rt[r] < res * res * bn[r];
      st.true_rel[r] = bn[r] > 0 ? std::sqrt(rt[r] / bn[r]) : 0.0;
      all &= converged[r];
    }
    if (all || iteration >= prm.max_iters) break;
    // sloppy phase until every unconverged rhs has dropped by delta^2
    for (int r = 0; r < NR; r++) {
      BCGR& S = hs[r];
      S.halt = converged[r];
      S.go = 0;
      S.iters = 0;
      S.rsq = rt[r];
      const double res = prm.resid_r ? prm.resid_r[r] : prm.resid;
      S.rsq_stop = std::max(res * res * bn[r], delta * delta * rt[r]);
    }
    if ((rc = (int)hipMemcpyAsync(ds, hs.data(), NR * sizeof(BCGR), hipMemcpyHostToDevice, c.stream))) break;
    hg0.iter_cap = (int)std::min<long>(prm.max_iters - iteration, 1L << 30);
    if ((rc = (int)hipMemcpyAsync(dg, &hg0, sizeof(BCGG), hipMemcpyHostToDevice, c.stream))) break;
    if (sp == P_H16)
      hipLaunchKernelGGL(k_bru_h16<NR>, dim3(rg.nb), dim3(256), 0, c.stream, (const double*)r64.p, (float*)rs.p,
                         (int4*)ps.p, ds, first ? 1 : 0, rg.nrec, rg.loop, part, pstride);
    else
      hipLaunchKernelGGL(k_bru<NR>, dim3(rg.nb), dim3(256), 0, c.stream, (const double*)r64.p, (float*)rs.p,
                         (float*)ps.p, ds, first ? 1 : 0, rg.nrec, rg.loop, part, pstride);
    if ((rc = (int)hipGetLastError())) break;
    if ((rc = breduce(NR, rg.nb, dsum_rr))) break;
    if (!first) nru++;
    first = false;
    if (!gr.e && (rc = capture_graph(gr, iterations))) break;
    for (;;) {
      if ((rc = (int)hipGraphLaunch(gr.e, c.stream))) break;
      if ((rc = (int)hipMemcpyAsync(hg, dg, sizeof(BCGG), hipMemcpyDeviceToHost, c.stream))) break;
      if ((rc = (int)hipStreamSynchronize(c.stream))) break;
      if (hg->halt_all || iteration + hg->iters_max >= prm.max_iters) break;
    }
    if (rc) break;
    iteration += hg->iters_max;
  }
  (void)hipHostFree(hg);
  st.iters = (int)iteration;
  st.reliable_updates = nru;
  for (int r = 0; r < NR; r++) st.converged[r] = converged[r];
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // namespace

int bmultishift_sets(const void* b64, int nr, int n, const double* offs, const double* stage_resid,
                     const double* refine_resid, const BMSParams& prm, std::vector<void*>& x64, BMSStats& st) {
  int rc;
  st = BMSStats();
  if (nr == 4) rc = bmultishift_nr<4>(b64, offs, n, stage_resid, prm, x64, st);
  else if (nr == 8) rc = bmultishift_nr<8>(b64, offs, n, stage_resid, prm, x64, st);
  else if (nr == 16) rc = bmultishift_nr<16>(b64, offs, n, stage_resid, prm, x64, st);
  else return -2;
  if (rc || !refine_resid) return rc;
  // per-shift refinement to the FP64 true-residual targets
  auto t0 = std::chrono::steady_clock::now();
  int conv = 1;
  for (int j = 0; j < n; j++) {
    double off[BNR_MAX];
    for (int r = 0; r < nr; r++) off[r] = offs[j * nr + r];
    BCGParams cp;
    cp.parity = prm.parity;
    cp.max_iters = prm.max_iters;
    cp.resid = refine_resid[0];
    cp.resid_r = refine_resid;
    cp.delta = prm.delta;
    cp.sloppy = prm.refine_sloppy;
    BCGStats cs;
    rc = bcg_mixed(b64, x64[j], nr, off, cp, cs);
    if (rc) return rc;
    st.refine_iters[j] = cs.iters;
    double mx = 0;
    for (int r = 0; r < nr; r++) {
      mx = std::max(mx, cs.true_rel[r]);
      st.true_rel_r[j][r] = cs.true_rel[r];
      conv &= cs.converged[r];
    }
    st.true_rel[j] = mx;
  }
  st.converged = conv;
  st.refine_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int bmultishift(const void* b64, int nr, const std::vector<double>& offsets, const BMSParams& prm,
                std::vector<void*>& x64, BMSStats& st) {
  const int n = (int)offsets.size();
  if (n <= 0 || n > BMS_MAX || nr <= 0 || nr > BNR_MAX) return -2;
  std::vector<double> offs((size_t)n * nr), stage(nr, prm.resid), refine(nr, prm.refine_resid);
  for (int j = 0; j < n; j++)
    for (int r = 0; r < nr; r++) offs[(size_t)j * nr + r] = offsets[j];
  return bmultishift_sets(b64, nr, n, offs.data(), stage.data(), prm.refine_resid > 0 ? refine.data() : nullptr,
                          prm, x64, st);
}

int bcg_mixed(const void* b64, void* x64, int nr, const double* offsets, const BCGParams& prm, BCGStats& st) {
  st = BCGStats();
  if (nr == 4) return bcg_mixed_nr<4>(b64, x64, offsets, prm, st);
  if (nr == 8) return bcg_mixed_nr<8>(b64, x64, offsets, prm, st);
  if (nr == 16) return bcg_mixed_nr<16>(b64, x64, offsets, prm, st);
  return -2;
}

// Refinement of one source's multi-shift solutions (the per-shift CG of
// ks_multicg_hybrid_field, generic_ks/ks_multicg.c:170-236): up to four
// consecutive shifts at a time as one batched mixed CG (the source in every
// slot, each slot with its own offset and target; unused slots are zero and
// converged from the start), a shift left alone by the one-source mixed CG.
// The batched stencil of four right-hand sides costs about twice the
// one-source stencil (48^3x96, 16-bit: 995 against 500-550 us; FP32: 1878
// against about 1000 us), and neighbouring shifts need similar iteration
// counts, so a group costs about 2.5 times its slowest shift (48^3x96, 12
// shifts to 1e-10: 119, 113 and 48 iterations for the three groups, 0.91 s).
// One group of the eight lightest shifts instead of two of four was measured
// at 0.87 s (an 8-rhs iteration costs 1.75 times a 4-rhs one), not enough for
// the larger work fields.
int refine_shifts(const VecField& b, std::vector<VecField>& x, const double* offs, const double* resid, int parity,
                  int max_iters, int sloppy, long* iters, double* fres, int* converged) {
  Context& c = ctx();
  const int n = (int)x.size();
  int rc;
  if ((rc = ensure_links(sloppy))) return rc;
  for (int t0 = 0; t0 < n; t0 += 4) {
    const int g = std::min(4, n - t0);
    if (g == 1) {
      SolveParams sp;
      sp.parity = parity;
      sp.max_iters = max_iters;
      sp.max_restarts = 1;
      sp.resid = resid[t0];
      sp.relresid = 0;
      sp.sloppy_prec = sloppy;
      sp.delta = kReliableDelta;
      sp.batch = 16;
      SolveStats s2;
      if ((rc = cg_solve(b, x[t0], 0.5 * std::sqrt(offs[t0]), sp, s2))) return rc;
      *iters += s2.iters;
      if (fres) fres[t0] = std::sqrt(s2.final_rsq);
      *converged &= s2.converged;
      continue;
    }
    DevBuf wb, wx;
    if ((rc = balloc(&wb.p, P_F64, 4)) || (rc = balloc(&wx.p, P_F64, 4))) return rc;
    double off[4], res[4];
    for (int q = 0; q < 4; q++) {
      off[q] = offs[t0 + std::min(q, g - 1)];
      res[q] = resid[t0 + std::min(q, g - 1)];
    }
    for (int q = 0; q < g; q++)
      if ((rc = bpack(b, wb.p, P_F64, 4, q)) || (rc = bpack(x[t0 + q], wx.p, P_F64, 4, q))) return rc;
    BCGParams cp;
    cp.parity = parity;
    cp.max_iters = max_iters;
    cp.resid = res[0];
    cp.resid_r = res;
    cp.delta = kBatchReliableDelta;
    cp.sloppy = sloppy;
    BCGStats cs;
    if ((rc = bcg_mixed(wb.p, wx.p, 4, off, cp, cs))) return rc;
    for (int q = 0; q < g; q++) {
      if ((rc = bunpack(wx.p, P_F64, 4, q, x[t0 + q]))) return rc;
      if (fres) fres[t0 + q] = cs.true_rel[q];
      *converged &= cs.converged[q];
    }
    *iters += cs.iters;
    MH_CHECK(hipStreamSynchronize(c.stream));
  }
  return 0;
}

}  // namespace mh
