// mh_bkern.h -- device helpers shared by the batched kernels: batched record
// loads/stores and per-right-hand-side block reductions.
#pragma once
#include "mh_batch.h"

namespace mh {

template <int P> struct BR { typedef float T; };
template <> struct BR<P_F64> { typedef double T; };
template <typename R> struct BCoef { R c[BNR_MAX]; };

typedef float mf4 __attribute__((ext_vector_type(4)));

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
