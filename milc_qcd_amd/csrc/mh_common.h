// mh_common.h -- shared device types for the MI355X staggered (HISQ/asqtad) library.
//
// Device data layout (one lattice, dims nx*ny*nz*nt, nx even):
//   * checkerboard index cb = lex/2 within a parity, lex = x + nx(y + ny(z + nz t));
//     this is MILC's node_index order restricted to one parity
//     (generic/layout_hyper_prime.c:509-520), so the AoS<->SoA reorder is a
//     per-parity transpose;
//   * colour vectors: structure of arrays over sites, one array per colour
//     component ("plane"), so a wavefront's 64 lanes read 64 consecutive sites;
//   * link fields: [parity][dir][element plane][cb], every element plane
//     contiguous over cb, forward links only (the backward hop reads the
//     neighbour's forward link and applies the adjoint in registers, so no
//     MILC-style double store is kept).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stdio.h>

#define MH_CHECK(call)                                                              \
  do {                                                                              \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess) {                                                         \
      fprintf(stderr, "milc_hip: HIP error %s at %s:%d: %s\n", hipGetErrorName(_e), \
              __FILE__, __LINE__, hipGetErrorString(_e));                           \
      return -(int)_e - 1000;                                                       \
    }                                                                               \
  } while (0)

namespace mh {

// MILC parity codes (include/macros.h:17-19)
enum { EVEN = 2, ODD = 1, EVENANDODD = 3 };

// Parity bit of the device layout: 0 = even sites, 1 = odd sites.
__host__ __device__ inline int pbit(int milc_parity) { return milc_parity == ODD ? 1 : 0; }

// ---------------------------------------------------------------------------
// Unsigned division by a runtime constant d (Granlund-Montgomery):
//   shift = ceil(log2 d), magic = floor(2^32 (2^shift - d) / d) + 1,
//   q = (mulhi(n, magic) + n) >> shift, exact for n < 2^31.
// ---------------------------------------------------------------------------
struct FastDiv {
  uint32_t d, magic, shift;
  FastDiv() : d(1), magic(1), shift(0) {}
  explicit FastDiv(uint32_t dv) : d(dv) {
    shift = 0;
    while ((1ull << shift) < d) shift++;
    magic = (uint32_t)(((1ull << 32) * ((1ull << shift) - d)) / d + 1);
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t hi = __umulhi(n, magic);
#else
    uint32_t hi = (uint32_t)(((uint64_t)n * magic) >> 32);
#endif
    return (uint32_t)(((uint64_t)hi + n) >> shift);
  }
};

// Lattice geometry passed by value to every stencil kernel.
struct Geom {
  int X[4];     // nx, ny, nz, nt
  int Xh;       // nx/2
  int Vh;       // sites per parity
  int sy, sz, st;  // cb strides of y, z, t
  FastDiv dXh, dY, dZ;
};

inline Geom make_geom(const int dims[4]) {
  Geom g;
  for (int i = 0; i < 4; i++) g.X[i] = dims[i];
  g.Xh = dims[0] / 2;
  g.Vh = dims[0] * dims[1] * dims[2] * dims[3] / 2;
  g.sy = g.Xh;
  g.sz = g.Xh * dims[1];
  g.st = g.Xh * dims[1] * dims[2];
  g.dXh = FastDiv((uint32_t)g.Xh);
  g.dY = FastDiv((uint32_t)dims[1]);
  g.dZ = FastDiv((uint32_t)dims[2]);
  return g;
}

// Coordinates of checkerboard site cb of parity bit pb.
struct Coord {
  int x, y, z, t, rowbase;
};
__host__ __device__ __forceinline__ Coord site_coords(const Geom& g, int cb, int pb) {
  Coord c;
  uint32_t r = g.dXh.div((uint32_t)cb);
  int xh = cb - (int)r * g.Xh;
  uint32_t r2 = g.dY.div(r);
  c.y = (int)r - (int)r2 * g.X[1];
  uint32_t r3 = g.dZ.div(r2);
  c.z = (int)r2 - (int)r3 * g.X[2];
  c.t = (int)r3;
  c.x = 2 * xh + ((c.y + c.z + c.t + pb) & 1);
  c.rowbase = cb - xh;
  return c;
}

// cb index of the neighbour at distance k (=+-1, +-3) along mu.
__host__ __device__ __forceinline__ int neighbour(const Geom& g, const Coord& c, int cb, int mu, int k) {
  if (mu == 0) {
    int x = c.x + k;
    x += (x < 0) ? g.X[0] : 0;
    x -= (x >= g.X[0]) ? g.X[0] : 0;
    return c.rowbase + (x >> 1);
  }
  int coord = mu == 1 ? c.y : (mu == 2 ? c.z : c.t);
  int L = g.X[mu];
  int stride = mu == 1 ? g.sy : (mu == 2 ? g.sz : g.st);
  int n = coord + k;
  int d = k;
  if (n < 0) d += L;
  else if (n >= L) d -= L;
  return cb + d * stride;
}

// ---------------------------------------------------------------------------
// complex arithmetic in registers
// ---------------------------------------------------------------------------
template <typename R> struct Cplx { R re, im; };

template <typename R> struct Vec3 {
  Cplx<R> c[3];
};
template <typename R> struct Mat3 {
  Cplx<R> e[3][3];
};

template <typename R>
__device__ __forceinline__ void zero(Vec3<R>& v) {
#pragma unroll
  for (int i = 0; i < 3; i++) v.c[i].re = v.c[i].im = R(0);
}

// acc += M v
template <typename R>
__device__ __forceinline__ void mv_acc(const Mat3<R>& m, const Vec3<R>& v, Vec3<R>& acc) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      acc.c[i].re += m.e[i][k].re * v.c[k].re - m.e[i][k].im * v.c[k].im;
      acc.c[i].im += m.e[i][k].re * v.c[k].im + m.e[i][k].im * v.c[k].re;
    }
  }
}
// acc -= M^dag v
template <typename R>
__device__ __forceinline__ void amv_sub(const Mat3<R>& m, const Vec3<R>& v, Vec3<R>& acc) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      acc.c[i].re -= m.e[k][i].re * v.c[k].re + m.e[k][i].im * v.c[k].im;
      acc.c[i].im -= m.e[k][i].re * v.c[k].im - m.e[k][i].im * v.c[k].re;
    }
  }
}

// ---------------------------------------------------------------------------
// Precision formats of device fields.
//   F64: double2 per complex.
//   F32: float2 per complex.
//   H16: 16-bit fixed point (int16, value = q / 32767 * scale).  Links share
//        one scale per field (max |element|); vectors carry one float norm per
//        site (max |component| of that site), stored in a separate norm plane.
// ---------------------------------------------------------------------------
enum Prec { P_F64 = 0, P_F32 = 1, P_H16 = 2 };

template <int P> struct PrecTraits;
template <> struct PrecTraits<P_F64> {
  typedef double Real;      // arithmetic type
  typedef double2 Store;    // storage of one complex
  static constexpr int bytes_per_complex = 16;
};
template <> struct PrecTraits<P_F32> {
  typedef float Real;
  typedef float2 Store;
  static constexpr int bytes_per_complex = 8;
};
template <> struct PrecTraits<P_H16> {
  typedef float Real;
  typedef short2 Store;
  static constexpr int bytes_per_complex = 4;
};

// A colour-vector field of one parity in SoA layout: plane k (k = colour) at
// data + k*stride; H16 fields also have a norm plane.
struct VecField {
  void* data;
  float* norm;   // H16 only
  int stride;    // elements per plane (>= Vh)
};

// A link field (fat or long) for both parities: element plane e (0..8) of
// direction mu and parity bit pb at data + ((pb*4 + mu)*9 + e)*stride.
struct LinkField {
  void* data;
  float scale;  // H16: value = q * scale / 32767
  int stride;
};

__device__ __forceinline__ float h16_to_f(short q) { return (float)q; }
__device__ __forceinline__ short f_to_h16(float v) {
  // v in [-1, 1] after normalisation; round to nearest
  float s = v * 32767.0f;
  s = fminf(fmaxf(s, -32767.0f), 32767.0f);
  return (short)__float2int_rn(s);
}

template <int P>
__device__ __forceinline__ void load_vec(const VecField& f, int cb, Vec3<typename PrecTraits<P>::Real>& v);

template <>
__device__ __forceinline__ void load_vec<P_F64>(const VecField& f, int cb, Vec3<double>& v) {
  const double2* p = (const double2*)f.data;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double2 a = p[k * f.stride + cb];
    v.c[k].re = a.x; v.c[k].im = a.y;
  }
}
template <>
__device__ __forceinline__ void load_vec<P_F32>(const VecField& f, int cb, Vec3<float>& v) {
  const float2* p = (const float2*)f.data;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float2 a = p[k * f.stride + cb];
    v.c[k].re = a.x; v.c[k].im = a.y;
  }
}
template <>
__device__ __forceinline__ void load_vec<P_H16>(const VecField& f, int cb, Vec3<float>& v) {
  const short2* p = (const short2*)f.data;
  float n = f.norm[cb] * (1.0f / 32767.0f);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    short2 a = p[k * f.stride + cb];
    v.c[k].re = h16_to_f(a.x) * n; v.c[k].im = h16_to_f(a.y) * n;
  }
}

template <int P>
__device__ __forceinline__ void store_vec(const VecField& f, int cb, const Vec3<typename PrecTraits<P>::Real>& v);

template <>
__device__ __forceinline__ void store_vec<P_F64>(const VecField& f, int cb, const Vec3<double>& v) {
  double2* p = (double2*)f.data;
#pragma unroll
  for (int k = 0; k < 3; k++) p[k * f.stride + cb] = make_double2(v.c[k].re, v.c[k].im);
}
template <>
__device__ __forceinline__ void store_vec<P_F32>(const VecField& f, int cb, const Vec3<float>& v) {
  float2* p = (float2*)f.data;
#pragma unroll
  for (int k = 0; k < 3; k++) p[k * f.stride + cb] = make_float2(v.c[k].re, v.c[k].im);
}
template <>
__device__ __forceinline__ void store_vec<P_H16>(const VecField& f, int cb, const Vec3<float>& v) {
  short2* p = (short2*)f.data;
  float m = 0.f;
#pragma unroll
  for (int k = 0; k < 3; k++) m = fmaxf(m, fmaxf(fabsf(v.c[k].re), fabsf(v.c[k].im)));
  float inv = m > 0.f ? 1.0f / m : 0.f;
  f.norm[cb] = m;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    short2 q;
    q.x = f_to_h16(v.c[k].re * inv);
    q.y = f_to_h16(v.c[k].im * inv);
    p[k * f.stride + cb] = q;
  }
}

template <int P>
__device__ __forceinline__ void load_link(const LinkField& f, int pb, int mu, int cb,
                                          Mat3<typename PrecTraits<P>::Real>& m);

template <>
__device__ __forceinline__ void load_link<P_F64>(const LinkField& f, int pb, int mu, int cb, Mat3<double>& m) {
  const double2* p = (const double2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    double2 a = p[(size_t)e * f.stride];
    m.e[e / 3][e % 3].re = a.x; m.e[e / 3][e % 3].im = a.y;
  }
}
template <>
__device__ __forceinline__ void load_link<P_F32>(const LinkField& f, int pb, int mu, int cb, Mat3<float>& m) {
  const float2* p = (const float2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    float2 a = p[(size_t)e * f.stride];
    m.e[e / 3][e % 3].re = a.x; m.e[e / 3][e % 3].im = a.y;
  }
}
template <>
__device__ __forceinline__ void load_link<P_H16>(const LinkField& f, int pb, int mu, int cb, Mat3<float>& m) {
  const short2* p = (const short2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride + cb;
  float s = f.scale * (1.0f / 32767.0f);
#pragma unroll
  for (int e = 0; e < 9; e++) {
    short2 a = p[(size_t)e * f.stride];
    m.e[e / 3][e % 3].re = h16_to_f(a.x) * s; m.e[e / 3][e % 3].im = h16_to_f(a.y) * s;
  }
}

// XCD-aware block remap: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; remap so each XCD walks one contiguous range of sites and
// its private L2 sees the neighbour halos of its own range.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
  const int nx = 8;
  int per = nblocks / nx;
  if (per == 0 || bid >= per * nx) return bid;
  return (bid % nx) * per + bid / nx;
}

}  // namespace mh
