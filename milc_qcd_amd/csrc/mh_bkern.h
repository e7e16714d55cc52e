// mh_bkern.h -- device helpers shared by the batched kernels: batched record
// loads/stores and per-right-hand-side block reductions.
#pragma once
#include "mh_batch.h"

namespace mh {

template <int P> struct BR { typedef float T; };
template <> struct BR<P_F64> { typedef double T; };
template <typename R> struct BCoef { R c[BNR_MAX]; };

typedef float mf4 __attribute__((ext_vector_type(4)));

// Site order of the batched stencil.  on = 0: checkerboard index order.
// on = 1: 4-D blocks of (2 bx2) x by x bz x bt sites in lexicographic block
// order, lexicographic (x fastest) inside a block.  After the xcd_block remap
// consecutive workgroups run on one XCD, so that XCD sweeps a compact region
// and the depth-3 halo of NR-rhs vectors around it stays in its L2 instead of
// being re-fetched from the Infinity Cache for each of the 16 hops.
struct BOrder {
  int on;
  int bx2, by, bz, bt, nbx, nby, nbz, bsites;
  FastDiv d_bs, d_bx2, d_by, d_bz, d_nbx, d_nby, d_nbz;
};

// checkerboard index and coordinates of logical site i of parity pb
__host__ __device__ __forceinline__ int border_site(const Geom& g, const BOrder& o, int i, int pb, Coord& c) {
  if (!o.on) {
    c = site_coords(g, i, pb);
    return i;
  }
  uint32_t b = o.d_bs.div((uint32_t)i);
  uint32_t s = (uint32_t)i - b * o.bsites;
  uint32_t r = o.d_bx2.div(s);
  int xl = (int)(s - r * o.bx2);
  uint32_t r2 = o.d_by.div(r);
  int yl = (int)(r - r2 * o.by);
  uint32_t r3 = o.d_bz.div(r2);
  int zl = (int)(r2 - r3 * o.bz);
  int tl = (int)r3;
  uint32_t q = o.d_nbx.div(b);
  int Bx = (int)(b - q * o.nbx);
  uint32_t q2 = o.d_nby.div(q);
  int By = (int)(q - q2 * o.nby);
  uint32_t q3 = o.d_nbz.div(q2);
  int Bz = (int)(q2 - q3 * o.nbz);
  int Bt = (int)q3;
  int xh = Bx * o.bx2 + xl;
  c.y = By * o.by + yl;
  c.z = Bz * o.bz + zl;
  c.t = Bt * o.bt + tl;
  c.x = 2 * xh + ((c.y + c.z + c.t + pb) & 1);
  int cb = xh + g.Xh * (c.y + g.X[1] * (c.z + g.X[2] * c.t));
  c.rowbase = cb - xh;
  return cb;
}

inline BOrder make_border(const Geom& g, int bx, int by, int bz, int bt) {
  BOrder o{};
  o.on = 0;
  if (bx <= 0 || bx % 2 || g.X[0] % bx || g.X[1] % by || g.X[2] % bz || g.X[3] % bt) return o;
  o.on = 1;
  o.bx2 = bx / 2;
  o.by = by;
  o.bz = bz;
  o.bt = bt;
  o.nbx = g.X[0] / bx;
  o.nby = g.X[1] / by;
  o.nbz = g.X[2] / bz;
  o.bsites = o.bx2 * by * bz * bt;
  o.d_bs = FastDiv((uint32_t)o.bsites);
  o.d_bx2 = FastDiv((uint32_t)o.bx2);
  o.d_by = FastDiv((uint32_t)by);
  o.d_bz = FastDiv((uint32_t)bz);
  o.d_nbx = FastDiv((uint32_t)o.nbx);
  o.d_nby = FastDiv((uint32_t)o.nby);
  o.d_nbz = FastDiv((uint32_t)o.nbz);
  return o;
}

template <typename R, int NR>
__device__ __forceinline__ void bld(const R* __restrict__ f, int cb, int r, Vec3<R>& v) {
  const R* a = f + ((size_t)cb * NR + r) * 6;
  v.c[0].re = a[0]; v.c[1].re = a[1]; v.c[2].re = a[2];
  v.c[0].im = a[3]; v.c[1].im = a[4]; v.c[2].im = a[5];
}
template <typename R, int NR>
__device__ __forceinline__ void bst(R* __restrict__ f, int cb, int r, const Vec3<R>& v) {
  R* a = f + ((size_t)cb * NR + r) * 6;
  a[0] = v.c[0].re; a[1] = v.c[1].re; a[2] = v.c[2].re;
  a[3] = v.c[0].im; a[4] = v.c[1].im; a[5] = v.c[2].im;
}

// RPT consecutive records (rhs r0 .. r0+RPT-1) of site cb with 16-byte loads
// (8-byte for a lone FP32 record): the records of a site are contiguous, so the
// threads of a site read whole cache lines with few instructions.
template <typename R> struct Chunk16 { typedef float4 T; static constexpr int n = 4; };
template <> struct Chunk16<double> { typedef double2 T; static constexpr int n = 2; };

template <typename R, int NR, int RPT>
__device__ __forceinline__ void bld_blk(const R* __restrict__ f, int cb, int r0, Vec3<R> (&v)[RPT]) {
  constexpr int NE = 6 * RPT;
  R e[NE];
  const R* a = f + ((size_t)cb * NR + r0) * 6;
  if constexpr ((NE * sizeof(R)) % 16 == 0) {
    typedef typename Chunk16<R>::T C;
    constexpr int EPC = Chunk16<R>::n;
    const C* p = (const C*)a;
#pragma unroll
    for (int k = 0; k < NE / EPC; k++) {
      C w = p[k];
      const R* wr = (const R*)&w;
#pragma unroll
      for (int j = 0; j < EPC; j++) e[k * EPC + j] = wr[j];
    }
  } else {
    const float2* p = (const float2*)a;
#pragma unroll
    for (int k = 0; k < NE / 2; k++) {
      float2 w = p[k];
      e[2 * k] = w.x;
      e[2 * k + 1] = w.y;
    }
  }
#pragma unroll
  for (int i = 0; i < RPT; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[i].c[k].re = e[6 * i + k];
      v[i].c[k].im = e[6 * i + 3 + k];
    }
}
template <typename R, int NR, int RPT>
__device__ __forceinline__ void bst_blk(R* __restrict__ f, int cb, int r0, const Vec3<R> (&v)[RPT]) {
  constexpr int NE = 6 * RPT;
  R e[NE];
#pragma unroll
  for (int i = 0; i < RPT; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      e[6 * i + k] = v[i].c[k].re;
      e[6 * i + 3 + k] = v[i].c[k].im;
    }
  R* a = f + ((size_t)cb * NR + r0) * 6;
  if constexpr ((NE * sizeof(R)) % 16 == 0) {
    typedef typename Chunk16<R>::T C;
    constexpr int EPC = Chunk16<R>::n;
    C* p = (C*)a;
#pragma unroll
    for (int k = 0; k < NE / EPC; k++) {
      C w;
      R* wr = (R*)&w;
#pragma unroll
      for (int j = 0; j < EPC; j++) wr[j] = e[k * EPC + j];
      p[k] = w;
    }
  } else {
    float2* p = (float2*)a;
#pragma unroll
    for (int k = 0; k < NE / 2; k++) p[k] = make_float2(e[2 * k], e[2 * k + 1]);
  }
}

// 16-bit batched records (the sloppy vectors of the batched mixed CG): record
// (site cb, rhs r) = one int4 at cb * NR + r holding (re, im) of colours 0, 1,
// 2 as 16-bit fixed point of the record's norm (its largest |component|), the
// norm itself as a float in the fourth word: the single-rhs 16-bit vector
// record, so a site's NR records are NR * 16 contiguous bytes.
__device__ __forceinline__ void h16_unpack(const int4& w, Vec3<float>& v) {
  const float n = __int_as_float(w.w) * (1.0f / 32767.0f);
  v.c[0].re = q16lo(w.x) * n;
  v.c[0].im = q16hi(w.x) * n;
  v.c[1].re = q16lo(w.y) * n;
  v.c[1].im = q16hi(w.y) * n;
  v.c[2].re = q16lo(w.z) * n;
  v.c[2].im = q16hi(w.z) * n;
}
__device__ __forceinline__ int4 h16_pack(const Vec3<float>& v) {
  float m = 0.f;
#pragma unroll
  for (int k = 0; k < 3; k++) m = fmaxf(m, fmaxf(fabsf(v.c[k].re), fabsf(v.c[k].im)));
  const float inv = m > 0.f ? 1.0f / m : 0.f;
  int4 o;
  o.x = q16pack(v.c[0].re * inv, v.c[0].im * inv);
  o.y = q16pack(v.c[1].re * inv, v.c[1].im * inv);
  o.z = q16pack(v.c[2].re * inv, v.c[2].im * inv);
  o.w = __float_as_int(m);
  return o;
}
template <int NR, int RPT>
__device__ __forceinline__ void bld_blk_h16(const int4* __restrict__ f, int cb, int r0, Vec3<float> (&v)[RPT]) {
  const int4* a = f + (size_t)cb * NR + r0;
#pragma unroll
  for (int i = 0; i < RPT; i++) h16_unpack(a[i], v[i]);
}
template <int NR, int RPT>
__device__ __forceinline__ void bst_blk_h16(int4* __restrict__ f, int cb, int r0, const Vec3<float> (&v)[RPT]) {
  int4* a = f + (size_t)cb * NR + r0;
#pragma unroll
  for (int i = 0; i < RPT; i++) a[i] = h16_pack(v[i]);
}

// Per-rhs block sums when rhs = lane % NR: written to LDS, then one value per
// (dot, rhs) channel to partials[ch * pstride + blockIdx.x].
template <int NV, int NR>
__device__ __forceinline__ void bsum_lanes(double (&d)[NV], double* __restrict__ partials, int pstride) {
#pragma unroll
  for (int v = 0; v < NV; v++)
    for (int off = NR; off < 64; off <<= 1) d[v] += __shfl_xor(d[v], off, 64);
  __shared__ double sm[NV][NR][4];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < NR)
#pragma unroll
    for (int v = 0; v < NV; v++) sm[v][lane][w] = d[v];
  __syncthreads();
  int t = threadIdx.x;
  if (t < NV * NR) {
    int v = t / NR, r = t % NR;
    double s = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sm[v][r][k];
    partials[(size_t)t * pstride + blockIdx.x] = s;
  }
}

// Same for a thread owning rhs RPT q + i (q = lane % TPS, RPT = NR / TPS).
template <int NV, int NR, int TPS>
__device__ __forceinline__ void bsum_valu(double (&d)[NV][NR / TPS], double* __restrict__ partials, int pstride) {
  constexpr int RPT = NR / TPS;
#pragma unroll
  for (int v = 0; v < NV; v++)
#pragma unroll
    for (int i = 0; i < RPT; i++)
      for (int off = TPS; off < 64; off <<= 1) d[v][i] += __shfl_xor(d[v][i], off, 64);
  __shared__ double sm[NV][NR][4];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < TPS)
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
      for (int i = 0; i < RPT; i++) sm[v][RPT * lane + i][w] = d[v][i];
  __syncthreads();
  int t = threadIdx.x;
  if (t < NV * NR) {
    int v = t / NR, r = t % NR;
    double s = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sm[v][r][k];
    partials[(size_t)t * pstride + blockIdx.x] = s;
  }
}

}  // namespace mh
