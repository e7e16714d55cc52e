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

// Same for a thread owning rhs q + TPS i (q = lane % TPS).
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
      for (int i = 0; i < RPT; i++) sm[v][lane + TPS * i][w] = d[v][i];
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
