// mh_bdslash.hip -- the fat + Naik stencil applied to NR right-hand sides at once.
//
// Same operator as generic_ks/dslash_fn_dblstore.c:311-563 (one output parity,
// dst(x) = sum_mu F_mu(x)s(x+mu) + L_mu(x)s(x+3mu) - F_mu(x-mu)^dag s(x-mu)
// - L_mu(x-3mu)^dag s(x-3mu)), applied to NR vectors that share the links.
// The link traffic (1152 B/site in FP32) is paid once for all NR vectors, so
// the per-rhs cost falls from 1200 B to 72 + 48 B per output site and the
// product becomes a small dense contraction: 3x3 link times 3xNR block.
//
// Two implementations compete in an autotuner:
//   VALU  tps threads per site, nr/tps right-hand sides per thread, links in
//         the solver's own F32/F64 formats.
//   MFMA  v_mfma_f32_4x4x1f32 (16 independent 4x4 outer-product blocks per
//         wavefront).  Block = (site, real-or-imaginary part, 4 rhs); its 4
//         rows are the colour rows (row 3 idle: 75% of the unit in use), its 4
//         columns 4 right-hand sides.  The links are re-laid out once per link
//         set as a double store with the sign folded in (W_d for the 16 hops:
//         F, L, -F^dag, -L^dag), split by row, so every lane fetches exactly its
//         row's three complex entries and every hop is the same code:
//           re: acc1 += Re W_ik Re v_k, acc2 += Im W_ik Im v_k, out = acc1 - acc2
//           im: acc1 += Re W_ik Im v_k, acc2 += Im W_ik Re v_k, out = acc1 + acc2
//         with the real and imaginary halves of each vector record fetched at
//         lane-dependent offsets (no per-lane selects in the loop).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mh_batch.h"
#include "mh_bkern.h"
#include "mh_dslash.h"

namespace mh {

// ---------------------------------------------------------------------------
// VALU variant
// ---------------------------------------------------------------------------
template <int P, int NR, int TPS, int KIND>
__global__ void __launch_bounds__(256) k_bds_valu(Geom g, LinkField fat, LinkField lng,
                                                  const typename BR<P>::T* __restrict__ src,
                                                  typename BR<P>::T* __restrict__ dst,
                                                  const typename BR<P>::T* __restrict__ pv,
                                                  const typename BR<P>::T* __restrict__ rv,
                                                  BCoef<typename BR<P>::T> coef, int pb, int loop, BOrder ord,
                                                  double* __restrict__ partials, int pstride,
                                                  const int* __restrict__ halt_all) {
  typedef typename BR<P>::T R;
  if (halt_all && *halt_all) return;
  constexpr int RPT = NR / TPS;
  constexpr int NV = KIND == BK_CG ? 3 : 1;
  const int q = threadIdx.x % TPS;
  const int gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int per_pass = (gridDim.x * blockDim.x) / TPS;
  double d[NV][RPT];
#pragma unroll
  for (int v = 0; v < NV; v++)
#pragma unroll
    for (int i = 0; i < RPT; i++) d[v][i] = 0.0;
  const int po = pb ^ 1;
  for (int it = 0; it < loop; it++) {
    int li = it * per_pass + gid / TPS;
    bool valid = li < g.Vh;
    if (!valid) li = g.Vh - 1;
    Coord c;
    const int cb = border_site(g, ord, li, pb, c);
    Vec3<R> acc[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) zero(acc[i]);
    // this thread's right-hand sides: RPT consecutive records r0 .. r0 + RPT - 1
    const int r0 = RPT * q;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
      Mat3<R> m;
      Vec3<R> v[RPT];
      int n = neighbour(g, c, cb, mu, 1);
      load_link<P>(fat, pb, mu, cb, m);
      bld_blk<R, NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) mv_acc(m, v[i], acc[i]);
      n = neighbour(g, c, cb, mu, 3);
      load_link<P>(lng, pb, mu, cb, m);
      bld_blk<R, NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) mv_acc(m, v[i], acc[i]);
      n = neighbour(g, c, cb, mu, -1);
      load_link<P>(fat, po, mu, n, m);
      bld_blk<R, NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) amv_sub(m, v[i], acc[i]);
      n = neighbour(g, c, cb, mu, -3);
      load_link<P>(lng, po, mu, n, m);
      bld_blk<R, NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) amv_sub(m, v[i], acc[i]);
    }
    if (KIND == BK_STORE) {
      if (valid) bst_blk<R, NR, RPT>(dst, cb, r0, acc);
      continue;
    }
    Vec3<R> pp[RPT], t[RPT];
    bld_blk<R, NR, RPT>(pv, cb, r0, pp);
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      R a = coef.c[r0 + i];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        t[i].c[k].re = acc[i].c[k].re + a * pp[i].c[k].re;
        t[i].c[k].im = acc[i].c[k].im + a * pp[i].c[k].im;
      }
    }
    if (!valid) continue;
    bst_blk<R, NR, RPT>(dst, cb, r0, t);
#pragma unroll
    for (int i = 0; i < RPT; i++)
#pragma unroll
      for (int k = 0; k < 3; k++)
        d[0][i] += (double)pp[i].c[k].re * (double)t[i].c[k].re + (double)pp[i].c[k].im * (double)t[i].c[k].im;
    if (KIND == BK_CG) {
      Vec3<R> rr[RPT];
      bld_blk<R, NR, RPT>(rv, cb, r0, rr);
#pragma unroll
      for (int i = 0; i < RPT; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
          d[1 % NV][i] += (double)rr[i].c[k].re * (double)t[i].c[k].re + (double)rr[i].c[k].im * (double)t[i].c[k].im;
          d[2 % NV][i] += (double)t[i].c[k].re * (double)t[i].c[k].re + (double)t[i].c[k].im * (double)t[i].c[k].im;
        }
    }
  }
  if (KIND != BK_STORE) bsum_valu<NV, NR, TPS>(d, partials, pstride);
}

// 16-bit form for the sloppy operator of the batched mixed CG: 16-bit links
// (fat nine elements; LR = 1: Naik links in the two-row form), 16-bit source
// records, FP32 arithmetic.  BK_STORE writes 16-bit records; BK_CG reads the
// 16-bit p and the FP32 r and writes the FP32 dst.
template <int NR, int TPS, int KIND, int LR>
__global__ void __launch_bounds__(256) k_bds_h16(Geom g, LinkField fat, LinkField lng, const int4* __restrict__ src,
                                                 void* __restrict__ dst, const int4* __restrict__ pv,
                                                 const float* __restrict__ rv, BCoef<float> coef, int pb, int loop,
                                                 BOrder ord, double* __restrict__ partials, int pstride,
                                                 const int* __restrict__ halt_all) {
  static_assert(KIND == BK_STORE || KIND == BK_CG, "16-bit batched stencil: store or CG form");
  if (halt_all && *halt_all) return;
  constexpr int RPT = NR / TPS;
  constexpr int NV = 3;
  const int q = threadIdx.x % TPS;
  const int gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int per_pass = (gridDim.x * blockDim.x) / TPS;
  double d[NV][RPT];
#pragma unroll
  for (int v = 0; v < NV; v++)
#pragma unroll
    for (int i = 0; i < RPT; i++) d[v][i] = 0.0;
  const int po = pb ^ 1;
  for (int it = 0; it < loop; it++) {
    int li = it * per_pass + gid / TPS;
    bool valid = li < g.Vh;
    if (!valid) li = g.Vh - 1;
    Coord c;
    const int cb = border_site(g, ord, li, pb, c);
    Vec3<float> acc[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) zero(acc[i]);
    const int r0 = RPT * q;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
      Mat3<float> m;
      Vec3<float> v[RPT];
      int n = neighbour(g, c, cb, mu, 1);
      load_stencil_link<P_H16, 0>(fat, pb, mu, cb, m);
      bld_blk_h16<NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) mv_acc(m, v[i], acc[i]);
      n = neighbour(g, c, cb, mu, 3);
      load_stencil_link<P_H16, LR>(lng, pb, mu, cb, m);
      bld_blk_h16<NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) mv_acc(m, v[i], acc[i]);
      n = neighbour(g, c, cb, mu, -1);
      load_stencil_link<P_H16, 0>(fat, po, mu, n, m);
      bld_blk_h16<NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) amv_sub(m, v[i], acc[i]);
      n = neighbour(g, c, cb, mu, -3);
      load_stencil_link<P_H16, LR>(lng, po, mu, n, m);
      bld_blk_h16<NR, RPT>(src, n, r0, v);
#pragma unroll
      for (int i = 0; i < RPT; i++) amv_sub(m, v[i], acc[i]);
    }
    if (KIND == BK_STORE) {
      if (valid) bst_blk_h16<NR, RPT>((int4*)dst, cb, r0, acc);
      continue;
    }
    Vec3<float> pp[RPT], t[RPT], rr[RPT];
    bld_blk_h16<NR, RPT>(pv, cb, r0, pp);
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const float a = coef.c[r0 + i];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        t[i].c[k].re = acc[i].c[k].re + a * pp[i].c[k].re;
        t[i].c[k].im = acc[i].c[k].im + a * pp[i].c[k].im;
      }
    }
    if (!valid) continue;
    bst_blk<float, NR, RPT>((float*)dst, cb, r0, t);
    bld_blk<float, NR, RPT>(rv, cb, r0, rr);
#pragma unroll
    for (int i = 0; i < RPT; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        d[0][i] += (double)pp[i].c[k].re * (double)t[i].c[k].re + (double)pp[i].c[k].im * (double)t[i].c[k].im;
        d[1][i] += (double)rr[i].c[k].re * (double)t[i].c[k].re + (double)rr[i].c[k].im * (double)t[i].c[k].im;
        d[2][i] += (double)t[i].c[k].re * (double)t[i].c[k].re + (double)t[i].c[k].im * (double)t[i].c[k].im;
      }
  }
  if (KIND != BK_STORE) bsum_valu<NV, NR, TPS>(d, partials, pstride);
}

// ---------------------------------------------------------------------------
// MFMA variant
// ---------------------------------------------------------------------------
// W: [pb][hop d = 0..15][row][cb][k][re, im] floats, d = mu + 4 * {+1, +3, -1, -3}
template <int NR, int KIND>
__global__ void __launch_bounds__(256) k_bds_mfma(Geom g, const float* __restrict__ W,
                                                  const float* __restrict__ src, float* __restrict__ dst,
                                                  const float* __restrict__ pv, const float* __restrict__ rv,
                                                  BCoef<float> coef, int pb, int loop, BOrder ord,
                                                  double* __restrict__ partials, int pstride,
                                                  const int* __restrict__ halt_all) {
  if (halt_all && *halt_all) return;
  constexpr int LPS = 2 * NR;   // lanes per site
  constexpr int SPW = 64 / LPS; // sites per wavefront and group
  constexpr int NV = KIND == BK_CG ? 3 : 1;
  const int lane = threadIdx.x & 63;
  const int wave = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * (blockDim.x >> 6);
  const int ls = lane & (LPS - 1);
  const int part = ls / NR;       // 0: real parts, 1: imaginary parts
  const int rhs = ls & (NR - 1);
  const int row = min(ls & 3, 2); // A-operand row of this lane (row 3 idle)
  const size_t plane = (size_t)g.Vh * 6;
  const float cf = coef.c[rhs];
  double d[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) d[v] = 0.0;
  for (int it = 0; it < loop; it++) {
    int li = (it * nwaves + wave) * SPW + lane / LPS;
    bool valid = li < g.Vh;
    if (!valid) li = g.Vh - 1;
    Coord c;
    const int cb = border_site(g, ord, li, pb, c);
    mf4 a1 = {0.f, 0.f, 0.f, 0.f}, a2 = {0.f, 0.f, 0.f, 0.f};
    const float* wrow = W + ((size_t)(pb * 16) * 3 + row) * plane + (size_t)cb * 6;
#pragma unroll 4
    for (int dd = 0; dd < 16; dd++) {
      const int mu = dd & 3, hop = dd >> 2;
      const int k = hop == 0 ? 1 : (hop == 1 ? 3 : (hop == 2 ? -1 : -3));
      const int n = neighbour(g, c, cb, mu, k);
      const float* w = wrow + (size_t)dd * 3 * plane;
      const float w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4], w5 = w[5];
      const float* v = src + ((size_t)n * NR + rhs) * 6;
      const float* v1 = v + 3 * part;
      const float* v2 = v + 3 - 3 * part;
      const float b10 = v1[0], b11 = v1[1], b12 = v1[2];
      const float b20 = v2[0], b21 = v2[1], b22 = v2[2];
      a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(w0, b10, a1, 0, 0, 0);
      a2 = __builtin_amdgcn_mfma_f32_4x4x1f32(w1, b20, a2, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(w2, b11, a1, 0, 0, 0);
      a2 = __builtin_amdgcn_mfma_f32_4x4x1f32(w3, b21, a2, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_4x4x1f32(w4, b12, a1, 0, 0, 0);
      a2 = __builtin_amdgcn_mfma_f32_4x4x1f32(w5, b22, a2, 0, 0, 0);
    }
    float o[3];
    o[0] = part ? a1[0] + a2[0] : a1[0] - a2[0];
    o[1] = part ? a1[1] + a2[1] : a1[1] - a2[1];
    o[2] = part ? a1[2] + a2[2] : a1[2] - a2[2];
    const size_t rec = ((size_t)cb * NR + rhs) * 6 + 3 * part;
    if (KIND == BK_STORE) {
      if (valid) {
        dst[rec] = o[0];
        dst[rec + 1] = o[1];
        dst[rec + 2] = o[2];
      }
      continue;
    }
    float p0 = pv[rec], p1 = pv[rec + 1], p2 = pv[rec + 2];
    float t0 = o[0] + cf * p0, t1 = o[1] + cf * p1, t2 = o[2] + cf * p2;
    if (!valid) continue;
    dst[rec] = t0;
    dst[rec + 1] = t1;
    dst[rec + 2] = t2;
    d[0] += (double)p0 * t0 + (double)p1 * t1 + (double)p2 * t2;
    if (KIND == BK_CG) {
      float r0 = rv[rec], r1 = rv[rec + 1], r2 = rv[rec + 2];
      d[1 % NV] += (double)r0 * t0 + (double)r1 * t1 + (double)r2 * t2;
      d[2 % NV] += (double)t0 * t0 + (double)t1 * t1 + (double)t2 * t2;
    }
  }
  if (KIND != BK_STORE) bsum_lanes<NV, NR>(d, partials, pstride);
}

__global__ void k_build_w(Geom g, LinkField fat, LinkField lng, float* __restrict__ W) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pd = blockIdx.y;  // pb * 16 + d
  if (cb >= g.Vh) return;
  int pb = pd >> 4, dd = pd & 15, mu = dd & 3, hop = dd >> 2;
  Mat3<double> m;
  if (hop < 2) {
    load_link<P_F64>(hop == 0 ? fat : lng, pb, mu, cb, m);
  } else {
    Coord c = site_coords(g, cb, pb);
    int n = neighbour(g, c, cb, mu, hop == 2 ? -1 : -3);
    load_link<P_F64>(hop == 2 ? fat : lng, pb ^ 1, mu, n, m);
  }
  for (int row = 0; row < 3; row++) {
    float* w = W + ((size_t)(pd * 3 + row) * g.Vh + cb) * 6;
    for (int k = 0; k < 3; k++) {
      // forward hops: W = U; backward hops: W = -U^dag
      double x = hop < 2 ? m.e[row][k].re : -m.e[k][row].re;
      double y = hop < 2 ? m.e[row][k].im : m.e[k][row].im;
      w[2 * k] = (float)x;
      w[2 * k + 1] = (float)y;
    }
  }
}

// channel sums of the partials: sums[ch] = sum_b partials[ch * pstride + b]
__global__ void __launch_bounds__(1024) k_breduce(const double* __restrict__ partials, int pstride, int nblocks,
                                                  double* __restrict__ sums) {
  const double* p = partials + (size_t)blockIdx.x * pstride;
  double s = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += p[b];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double sm[16];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sm[w];
    sums[blockIdx.x] = t;
  }
}

template <typename R, int NR>
__global__ void k_bpack(VecField in, R* __restrict__ out, int r, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<double> v;
  load_vec<P_F64>(in, cb, v);
  Vec3<R> w;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    w.c[k].re = (R)v.c[k].re;
    w.c[k].im = (R)v.c[k].im;
  }
  bst<R, NR>(out, cb, r, w);
}
template <typename R, int NR>
__global__ void k_bunpack(const R* __restrict__ in, int r, VecField out, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<R> w;
  bld<R, NR>(in, cb, r, w);
  Vec3<double> v;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v.c[k].re = (double)w.c[k].re;
    v.c[k].im = (double)w.c[k].im;
  }
  store_vec<P_F64>(out, cb, v);
}

template <int NR>
__global__ void k_bpack_h16(VecField in, int4* __restrict__ out, int r, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<double> v;
  load_vec<P_F64>(in, cb, v);
  Vec3<float> w;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    w.c[k].re = (float)v.c[k].re;
    w.c[k].im = (float)v.c[k].im;
  }
  out[(size_t)cb * NR + r] = h16_pack(w);
}
template <int NR>
__global__ void k_bunpack_h16(const int4* __restrict__ in, int r, VecField out, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<float> w;
  h16_unpack(in[(size_t)cb * NR + r], w);
  Vec3<double> v;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v.c[k].re = (double)w.c[k].re;
    v.c[k].im = (double)w.c[k].im;
  }
  store_vec<P_F64>(out, cb, v);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

const int kBlockCap = 16384;  // partial-sum records per channel

struct BWork {
  double* partials = nullptr;
  float* W = nullptr;
  unsigned W_version = ~0u;
  int tuned[3][3];      // [prec][nr index] -> variant, -1 untuned
  double tuned_us[3][3];
  BWork() {
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) tuned[a][b] = -1, tuned_us[a][b] = 0;
  }
};
BWork g_bw;

int nr_index(int nr) { return nr == 4 ? 0 : (nr == 8 ? 1 : (nr == 16 ? 2 : -1)); }

// site-order block shapes offered to the tuner (index 0: checkerboard order)
const int kOrders[4][4] = {{0, 0, 0, 0}, {8, 8, 8, 4}, {8, 4, 4, 4}, {4, 4, 4, 4}};
BOrder order_of(int k) {
  if (k <= 0 || k > 3) return BOrder{};
  return make_border(ctx().geom, kOrders[k][0], kOrders[k][1], kOrders[k][2], kOrders[k][3]);
}
int prec_index(int prec) { return prec == P_F64 ? 0 : (prec == P_F32 ? 1 : (prec == P_H16 ? 2 : -1)); }

int ensure_partials() {
  if (g_bw.partials) return 0;
  return dev_alloc((void**)&g_bw.partials, (size_t)3 * BNR_MAX * kBlockCap * sizeof(double));
}

int ensure_w() {
  Context& c = ctx();
  if (!c.links[P_F64].valid) return -3;
  if (g_bw.W && g_bw.W_version == c.links_version) return 0;
  int rc;
  if (!g_bw.W && (rc = dev_alloc((void**)&g_bw.W, (size_t)2 * 16 * 3 * c.Vh * 6 * sizeof(float)))) return rc;
  hipLaunchKernelGGL(k_build_w, dim3((c.Vh + 255) / 256, 32), dim3(256), 0, c.stream, c.geom, c.links[P_F64].fat,
                     c.links[P_F64].lng, g_bw.W);
  MH_CHECK(hipGetLastError());
  g_bw.W_version = c.links_version;
  return 0;
}

// smallest site loop that keeps the block count within the partial-sum cap
int loop_for(long blocks_at_loop1, int loop_min) {
  int loop = std::max(1, loop_min);
  while ((blocks_at_loop1 + loop - 1) / loop > kBlockCap) loop *= 2;
  return loop;
}

template <int P, int NR, int TPS>
int launch_valu(const BDOp& op, int block, int loop_min, int order, int* nb_out) {
  typedef typename BR<P>::T R;
  Context& c = ctx();
  long threads1 = (long)c.Vh * TPS;
  long blocks1 = (threads1 + block - 1) / block;
  int loop = loop_for(blocks1, loop_min);
  int nb = (int)((blocks1 + loop - 1) / loop);
  const BOrder ord = order_of(order);
  BCoef<R> cf;
  for (int r = 0; r < BNR_MAX; r++) cf.c[r] = (R)op.coef[r];
  const DevLinks& L = c.links[P];
  if (!L.valid) return -3;
  const R* src = (const R*)op.src;
  R* dst = (R*)op.dst;
  const R* pv = (const R*)op.p;
  const R* rv = (const R*)op.r;
  if (op.kind == BK_STORE)
    hipLaunchKernelGGL((k_bds_valu<P, NR, TPS, BK_STORE>), dim3(nb), dim3(block), 0, c.stream, c.geom, L.fat, L.lng,
                       src, dst, pv, rv, cf, op.pb, loop, ord, g_bw.partials, kBlockCap, op.halt_all);
  else if (op.kind == BK_MS)
    hipLaunchKernelGGL((k_bds_valu<P, NR, TPS, BK_MS>), dim3(nb), dim3(block), 0, c.stream, c.geom, L.fat, L.lng, src,
                       dst, pv, rv, cf, op.pb, loop, ord, g_bw.partials, kBlockCap, op.halt_all);
  else
    hipLaunchKernelGGL((k_bds_valu<P, NR, TPS, BK_CG>), dim3(nb), dim3(block), 0, c.stream, c.geom, L.fat, L.lng, src,
                       dst, pv, rv, cf, op.pb, loop, ord, g_bw.partials, kBlockCap, op.halt_all);
  MH_CHECK(hipGetLastError());
  if (nb_out) *nb_out = nb;
  return 0;
}

template <int NR, int TPS>
int launch_h16(const BDOp& op, int block, int loop_min, int order, int* nb_out) {
  Context& c = ctx();
  if (op.kind != BK_STORE && op.kind != BK_CG) return -2;
  long threads1 = (long)c.Vh * TPS;
  long blocks1 = (threads1 + block - 1) / block;
  int loop = loop_for(blocks1, loop_min);
  int nb = (int)((blocks1 + loop - 1) / loop);
  const BOrder ord = order_of(order);
  BCoef<float> cf;
  for (int r = 0; r < BNR_MAX; r++) cf.c[r] = (float)op.coef[r];
  const DevLinks& L = c.links[P_H16];
  if (!L.valid) return -3;
  const int4* src = (const int4*)op.src;
  const int4* pv = (const int4*)op.p;
  const float* rv = (const float*)op.r;
  const bool lr = L.lng.recon != 0;
#define MH_H16_LAUNCH(KIND, LR)                                                                                   \
  hipLaunchKernelGGL((k_bds_h16<NR, TPS, KIND, LR>), dim3(nb), dim3(block), 0, c.stream, c.geom, L.fat, L.lng, src, \
                     op.dst, pv, rv, cf, op.pb, loop, ord, g_bw.partials, kBlockCap, op.halt_all)
  if (op.kind == BK_STORE) {
    if (lr) MH_H16_LAUNCH(BK_STORE, 1);
    else MH_H16_LAUNCH(BK_STORE, 0);
  } else {
    if (lr) MH_H16_LAUNCH(BK_CG, 1);
    else MH_H16_LAUNCH(BK_CG, 0);
  }
#undef MH_H16_LAUNCH
  MH_CHECK(hipGetLastError());
  if (nb_out) *nb_out = nb;
  return 0;
}

template <int NR>
int launch_h16_tps(const BDOp& op, const BVariant& v, int* nb) {
  switch (v.tps) {
    case 1:
      if constexpr (NR <= 8) return launch_h16<NR, 1>(op, v.block, v.loop, v.order, nb);
      break;
    case 2:
      if constexpr (NR / 2 <= 8) return launch_h16<NR, 2>(op, v.block, v.loop, v.order, nb);
      break;
    case 4:
      return launch_h16<NR, 4>(op, v.block, v.loop, v.order, nb);
    case 8:
      if constexpr (NR >= 8) return launch_h16<NR, 8>(op, v.block, v.loop, v.order, nb);
      break;
  }
  return -2;
}

template <int NR>
int launch_mfma(const BDOp& op, int block, int loop_min, int order, int* nb_out) {
  Context& c = ctx();
  int rc = ensure_w();
  if (rc) return rc;
  constexpr int SPW = 64 / (2 * NR);
  long waves1 = ((long)c.Vh + SPW - 1) / SPW;
  int wpb = block / 64;
  long blocks1 = (waves1 + wpb - 1) / wpb;
  int loop = loop_for(blocks1, loop_min);
  // every site group of every wave must exist: round the wave count up
  long waves = (waves1 + loop - 1) / loop;
  int nb = (int)((waves + wpb - 1) / wpb);
  const BOrder ord = order_of(order);
  BCoef<float> cf;
  for (int r = 0; r < BNR_MAX; r++) cf.c[r] = (float)op.coef[r];
  const float* src = (const float*)op.src;
  float* dst = (float*)op.dst;
  const float* pv = (const float*)op.p;
  const float* rv = (const float*)op.r;
  if (op.kind == BK_STORE)
    hipLaunchKernelGGL((k_bds_mfma<NR, BK_STORE>), dim3(nb), dim3(block), 0, c.stream, c.geom, g_bw.W, src, dst, pv,
                       rv, cf, op.pb, loop, ord, g_bw.partials, kBlockCap, op.halt_all);
  else if (op.kind == BK_MS)
    hipLaunchKernelGGL((k_bds_mfma<NR, BK_MS>), dim3(nb), dim3(block), 0, c.stream, c.geom, g_bw.W, src, dst, pv, rv,
                       cf, op.pb, loop, ord, g_bw.partials, kBlockCap, op.halt_all);
  else
    hipLaunchKernelGGL((k_bds_mfma<NR, BK_CG>), dim3(nb), dim3(block), 0, c.stream, c.geom, g_bw.W, src, dst, pv, rv,
                       cf, op.pb, loop, ord, g_bw.partials, kBlockCap, op.halt_all);
  MH_CHECK(hipGetLastError());
  if (nb_out) *nb_out = nb;
  return 0;
}

template <int P, int NR>
int launch_valu_tps(const BDOp& op, const BVariant& v, int* nb) {
  constexpr int RMAX = P == P_F64 ? 4 : 8;  // accumulators per thread (more spill)
  switch (v.tps) {
    case 1:
      if constexpr (NR <= RMAX) return launch_valu<P, NR, 1>(op, v.block, v.loop, v.order, nb);
      break;
    case 2:
      if constexpr (NR / 2 <= RMAX) return launch_valu<P, NR, 2>(op, v.block, v.loop, v.order, nb);
      break;
    case 4:
      return launch_valu<P, NR, 4>(op, v.block, v.loop, v.order, nb);
    case 8:
      if constexpr (NR >= 8) return launch_valu<P, NR, 8>(op, v.block, v.loop, v.order, nb);
      break;
  }
  return -2;
}

int launch_variant(const BDOp& op, const BVariant& v, int* nb) {
  if (op.prec == P_H16) {
    if (v.mfma) return -2;
    if (op.nr == 4) return launch_h16_tps<4>(op, v, nb);
    if (op.nr == 8) return launch_h16_tps<8>(op, v, nb);
    return launch_h16_tps<16>(op, v, nb);
  }
  if (v.mfma) {
    if (op.prec != P_F32) return -2;
    if (op.nr == 4) return launch_mfma<4>(op, v.block, v.loop, v.order, nb);
    if (op.nr == 8) return launch_mfma<8>(op, v.block, v.loop, v.order, nb);
    return launch_mfma<16>(op, v.block, v.loop, v.order, nb);
  }
  if (op.prec == P_F64) {
    if (op.nr == 4) return launch_valu_tps<P_F64, 4>(op, v, nb);
    if (op.nr == 8) return launch_valu_tps<P_F64, 8>(op, v, nb);
    return launch_valu_tps<P_F64, 16>(op, v, nb);
  }
  if (op.nr == 4) return launch_valu_tps<P_F32, 4>(op, v, nb);
  if (op.nr == 8) return launch_valu_tps<P_F32, 8>(op, v, nb);
  return launch_valu_tps<P_F32, 16>(op, v, nb);
}

}  // namespace

size_t bvec_bytes(int prec, int nr) {
  if (prec == P_H16) return (size_t)ctx().Vh * nr * sizeof(int4);
  return (size_t)ctx().Vh * nr * 6 * (prec == P_F64 ? sizeof(double) : sizeof(float));
}

int balloc(void** p, int prec, int nr) {
  size_t n = bvec_bytes(prec, nr);
  int rc = dev_alloc(p, n);
  if (rc) return rc;
  MH_CHECK(hipMemsetAsync(*p, 0, n, ctx().stream));
  return 0;
}

std::vector<BVariant> bvariants(int prec, int nr) {
  std::vector<BVariant> v;
  int rpt_max = prec == P_F64 ? 4 : 8;
  std::vector<int> orders = {0};
  for (int k = 1; k < 4; k++)
    if (order_of(k).on) orders.push_back(k);
  for (int order : orders)
    for (int tps : {1, 2, 4, 8}) {
      if (tps > nr || nr / tps > rpt_max) continue;
      for (int block : {128, 256})
        for (int loop : {1, 2}) {
          if (order && loop > 1) continue;
          v.push_back(BVariant{0, tps, block, loop, order});
        }
    }
  if (prec == P_F32)
    for (int order : orders)
      for (int block : {128, 256})
        for (int loop : {1, 2, 4}) {
          if (order && loop > 2) continue;
          v.push_back(BVariant{1, 0, block, loop, order});
        }
  return v;
}

void bdescribe(int prec, int nr, int variant, char* buf, int n) {
  std::vector<BVariant> vs = bvariants(prec, nr);
  if (variant < 0 || variant >= (int)vs.size()) {
    snprintf(buf, n, "none");
    return;
  }
  const BVariant& v = vs[variant];
  char ord[48];
  if (v.order)
    snprintf(ord, sizeof ord, "site blocks %dx%dx%dx%d", kOrders[v.order][0], kOrders[v.order][1],
             kOrders[v.order][2], kOrders[v.order][3]);
  else
    snprintf(ord, sizeof ord, "checkerboard order");
  if (v.mfma)
    snprintf(buf, n, "mfma_f32_4x4x1 nr=%d sites/wave=%d block=%d loop=%d %s", nr, 32 / nr, v.block, v.loop, ord);
  else
    snprintf(buf, n, "valu %s nr=%d threads/site=%d rhs/thread=%d block=%d loop=%d %s",
             prec == P_F64 ? "f64" : (prec == P_H16 ? "16-bit" : "f32"), nr, v.tps, nr / v.tps, v.block, v.loop, ord);
}

int bpartials(double** p, int* pstride) {
  int rc = ensure_partials();
  if (rc) return rc;
  *p = g_bw.partials;
  *pstride = kBlockCap;
  return 0;
}

int btime_variant(int prec, int nr, int variant, int reps, double* usec) {
  Context& c = ctx();
  std::vector<BVariant> vs = bvariants(prec, nr);
  if (variant < 0 || variant >= (int)vs.size()) return -2;
  int rc;
  if ((rc = ensure_partials())) return rc;
  void *a = nullptr, *b = nullptr;
  if ((rc = balloc(&a, prec, nr)) || (rc = balloc(&b, prec, nr))) {
    dev_free(a);
    return rc;
  }
  BDOp op;
  op.prec = prec;
  op.nr = nr;
  op.kind = BK_STORE;
  op.pb = 0;
  op.src = a;
  op.dst = b;
  hipEvent_t e0, e1;
  MH_CHECK(hipEventCreate(&e0));
  MH_CHECK(hipEventCreate(&e1));
  for (int i = 0; i < 2 && !rc; i++) rc = launch_variant(op, vs[variant], nullptr);
  MH_CHECK(hipEventRecord(e0, c.stream));
  for (int i = 0; i < reps && !rc; i++) {
    op.pb = i & 1;
    rc = launch_variant(op, vs[variant], nullptr);
  }
  MH_CHECK(hipEventRecord(e1, c.stream));
  MH_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  MH_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  if (rc) return rc;
  *usec = 1000.0 * ms / reps;
  return 0;
}

int btune(int prec, int nr, int* chosen, double* usec_best) {
  int pi = prec_index(prec), ni = nr_index(nr);
  if (pi < 0 || ni < 0) return -2;
  // the timings depend on the lattice shape, not on the link values: a choice
  // holds until the context ends
  if (g_bw.tuned[pi][ni] >= 0) {
    *chosen = g_bw.tuned[pi][ni];
    if (usec_best) *usec_best = g_bw.tuned_us[pi][ni];
    return 0;
  }
  std::vector<BVariant> vs = bvariants(prec, nr);
  int best = -1;
  double best_us = 1e30;
  for (int v = 0; v < (int)vs.size(); v++) {
    double us;
    int rc = btime_variant(prec, nr, v, 5, &us);
    if (rc) return rc;
    if (us < best_us) {
      best_us = us;
      best = v;
    }
  }
  g_bw.tuned[pi][ni] = best;
  g_bw.tuned_us[pi][ni] = best_us;
  *chosen = best;
  if (usec_best) *usec_best = best_us;
  return 0;
}

int bdslash(const BDOp& op, int* nblocks_out) {
  int pi = prec_index(op.prec), ni = nr_index(op.nr);
  if (pi < 0 || ni < 0) return -2;
  int rc;
  if ((rc = ensure_partials())) return rc;
  int v = op.variant;
  if (v < 0) {
    v = g_bw.tuned[pi][ni];
    if (v < 0 && (rc = btune(op.prec, op.nr, &v, nullptr))) return rc;
  }
  std::vector<BVariant> vs = bvariants(op.prec, op.nr);
  if (v >= (int)vs.size()) return -2;
  return launch_variant(op, vs[v], nblocks_out);
}

int bpack(const VecField& in, void* batched, int prec, int nr, int r) {
  Context& c = ctx();
  dim3 grid((c.Vh + 255) / 256);
  if (prec == P_H16) {
    if (nr == 4) hipLaunchKernelGGL((k_bpack_h16<4>), grid, dim3(256), 0, c.stream, in, (int4*)batched, r, c.Vh);
    else if (nr == 8) hipLaunchKernelGGL((k_bpack_h16<8>), grid, dim3(256), 0, c.stream, in, (int4*)batched, r, c.Vh);
    else hipLaunchKernelGGL((k_bpack_h16<16>), grid, dim3(256), 0, c.stream, in, (int4*)batched, r, c.Vh);
  } else if (prec == P_F64) {
    if (nr == 4) hipLaunchKernelGGL((k_bpack<double, 4>), grid, dim3(256), 0, c.stream, in, (double*)batched, r, c.Vh);
    else if (nr == 8) hipLaunchKernelGGL((k_bpack<double, 8>), grid, dim3(256), 0, c.stream, in, (double*)batched, r, c.Vh);
    else hipLaunchKernelGGL((k_bpack<double, 16>), grid, dim3(256), 0, c.stream, in, (double*)batched, r, c.Vh);
  } else {
    if (nr == 4) hipLaunchKernelGGL((k_bpack<float, 4>), grid, dim3(256), 0, c.stream, in, (float*)batched, r, c.Vh);
    else if (nr == 8) hipLaunchKernelGGL((k_bpack<float, 8>), grid, dim3(256), 0, c.stream, in, (float*)batched, r, c.Vh);
    else hipLaunchKernelGGL((k_bpack<float, 16>), grid, dim3(256), 0, c.stream, in, (float*)batched, r, c.Vh);
  }
  MH_CHECK(hipGetLastError());
  return 0;
}

int bunpack(const void* batched, int prec, int nr, int r, VecField& out) {
  Context& c = ctx();
  dim3 grid((c.Vh + 255) / 256);
  if (prec == P_H16) {
    if (nr == 4) hipLaunchKernelGGL((k_bunpack_h16<4>), grid, dim3(256), 0, c.stream, (const int4*)batched, r, out, c.Vh);
    else if (nr == 8) hipLaunchKernelGGL((k_bunpack_h16<8>), grid, dim3(256), 0, c.stream, (const int4*)batched, r, out, c.Vh);
    else hipLaunchKernelGGL((k_bunpack_h16<16>), grid, dim3(256), 0, c.stream, (const int4*)batched, r, out, c.Vh);
  } else if (prec == P_F64) {
    if (nr == 4) hipLaunchKernelGGL((k_bunpack<double, 4>), grid, dim3(256), 0, c.stream, (const double*)batched, r, out, c.Vh);
    else if (nr == 8) hipLaunchKernelGGL((k_bunpack<double, 8>), grid, dim3(256), 0, c.stream, (const double*)batched, r, out, c.Vh);
    else hipLaunchKernelGGL((k_bunpack<double, 16>), grid, dim3(256), 0, c.stream, (const double*)batched, r, out, c.Vh);
  } else {
    if (nr == 4) hipLaunchKernelGGL((k_bunpack<float, 4>), grid, dim3(256), 0, c.stream, (const float*)batched, r, out, c.Vh);
    else if (nr == 8) hipLaunchKernelGGL((k_bunpack<float, 8>), grid, dim3(256), 0, c.stream, (const float*)batched, r, out, c.Vh);
    else hipLaunchKernelGGL((k_bunpack<float, 16>), grid, dim3(256), 0, c.stream, (const float*)batched, r, out, c.Vh);
  }
  MH_CHECK(hipGetLastError());
  return 0;
}

// per-channel sums of the partials left by the last kernel (device array sums[nch])
int breduce(int nch, int nblocks, double* sums) {
  Context& c = ctx();
  hipLaunchKernelGGL(k_breduce, dim3(nch), dim3(1024), 0, c.stream, g_bw.partials, kBlockCap, nblocks, sums);
  MH_CHECK(hipGetLastError());
  return 0;
}

void batch_release() {
  dev_free(g_bw.partials);
  dev_free(g_bw.W);
  g_bw = BWork();
}

}  // namespace mh
