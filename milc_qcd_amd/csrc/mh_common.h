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

// Block order for kernels that read neighbours in every direction.  Workgroups
// are dealt to the 8 XCDs in turn (block b runs on XCD b % 8), so a 1-D grid of
// xcd_grid(nbx, ny) blocks gives XCD k the site blocks [k per, (k+1) per) and,
// for each, all ny values of the second index in a row: the neighbour rows a
// site block reads are then fetched through the same L2.
__host__ __device__ __forceinline__ int xcd_per(int nbx) { return (nbx + 7) >> 3; }
inline dim3 xcd_grid(int nbx, int ny) { return dim3(8 * xcd_per(nbx) * ny); }
__device__ __forceinline__ bool xcd_block(int nbx, int ny, int& bx, int& by) {
  const int j = blockIdx.x >> 3;
  by = j % ny;
  bx = (blockIdx.x & 7) * xcd_per(nbx) + j / ny;
  return bx < nbx;
}

// MILC's KS phase of link (c, mu) with antiperiodic time boundary
// (generic_ks/rephase.c:64-118): eta_x = (-1)^t, eta_y = (-1)^(t+x),
// eta_z = (-1)^(t+x+y), eta_t = 1, times -1 on the last time slice.
__device__ __forceinline__ double ks_phase(const Geom& g, const Coord& c, int mu) {
  double ph[4];
  ph[3] = 1.0;
  ph[0] = (c.t % 2 == 1) ? -1.0 : 1.0;
  ph[1] = (c.x % 2 == 1) ? -ph[0] : ph[0];
  ph[2] = (c.y % 2 == 1) ? -ph[1] : ph[1];
  if (c.t == g.X[3] - 1) ph[3] = -ph[3];
  return ph[mu];
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
// Precision formats of device fields.  Every per-site record is read with
// 16-byte loads where its size allows, and link loads are non-temporal: each
// link is used once per Dslash, so it should not displace the colour vectors
// (read 16 times each) from L2 and the Infinity Cache.
//
//   F64  vector: 3 planes of double2 (one per colour).
//        link  : 9 planes of double2 per (parity, dir).
//   F32  vector: one float4 plane (colours 0,1) + one float2 plane (colour 2).
//        link  : 4 float4 planes (elements 0..7) + 1 float2 plane (element 8)
//                per (parity, dir); the float2 planes follow all float4 planes.
//   H16  16-bit fixed point, value = q/32767 * scale.
//        vector: one 16-byte record per site: 3 complex as (short,short) words
//                and the site's float norm (max |component|) in the last word.
//        link  : 2 int4 planes (elements 0..7) + 1 int plane (element 8) per
//                (parity, dir); one scale per field (max |element|).
// ---------------------------------------------------------------------------
enum Prec { P_F64 = 0, P_F32 = 1, P_H16 = 2 };

template <int P> struct PrecTraits;
template <> struct PrecTraits<P_F64> {
  typedef double Real;
  static constexpr int vec_bytes = 48;
  static constexpr int link_bytes = 144;
};
template <> struct PrecTraits<P_F32> {
  typedef float Real;
  static constexpr int vec_bytes = 24;
  static constexpr int link_bytes = 72;
};
template <> struct PrecTraits<P_H16> {
  typedef float Real;
  static constexpr int vec_bytes = 16;
  static constexpr int link_bytes = 36;
};
inline int vec_site_bytes(int prec) { return prec == P_F64 ? 48 : (prec == P_F32 ? 24 : 16); }
inline int link_site_bytes(int prec) { return prec == P_F64 ? 144 : (prec == P_F32 ? 72 : 36); }

// A colour-vector field of one parity (stride = sites per plane, >= Vh).
struct VecField {
  void* data = nullptr;
  float* norm = nullptr;   // unused (the 16-bit norm lives in the site record)
  int stride = 0;
};

// A link field (fat or long) for both parities and all four directions.
// H16 with recon = 1 (long links that are a real multiple of a U(3) matrix,
// L = c U): rows 0 and 1 are stored, row 2 = z conj(row0 x row1) with the
// per-link z = det(U) / c kept as its phase (16 bits; the modulus 1/c follows
// from the rows), i.e. 26 bytes per link instead of 36.
struct LinkField {
  void* data = nullptr;
  float scale = 1.0f;  // H16: value = q * scale / 32767
  int stride = 0;
  float zscale = 1.0f; // H16 recon: largest |Re z|, |Im z| (reported only)
  int recon = 0;
};

typedef double nat_d2 __attribute__((ext_vector_type(2)));
typedef float nat_f2 __attribute__((ext_vector_type(2)));
typedef float nat_f4 __attribute__((ext_vector_type(4)));
typedef int nat_i4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ nat_d2 ldnt(const double2* p) { return __builtin_nontemporal_load((const nat_d2*)p); }
__device__ __forceinline__ nat_f4 ldnt(const float4* p) { return __builtin_nontemporal_load((const nat_f4*)p); }
__device__ __forceinline__ nat_f2 ldnt(const float2* p) { return __builtin_nontemporal_load((const nat_f2*)p); }
__device__ __forceinline__ nat_i4 ldnt(const int4* p) { return __builtin_nontemporal_load((const nat_i4*)p); }
__device__ __forceinline__ int ldnt(const int* p) { return __builtin_nontemporal_load(p); }
typedef int nat_i2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ nat_i2 ldnt(const int2* p) { return __builtin_nontemporal_load((const nat_i2*)p); }

__device__ __forceinline__ float q16lo(int w) { return (float)(short)(w & 0xffff); }
__device__ __forceinline__ float q16hi(int w) { return (float)(w >> 16); }
__device__ __forceinline__ int q16pack(float a, float b) {
  int qa = __float2int_rn(fminf(fmaxf(a * 32767.0f, -32767.0f), 32767.0f));
  int qb = __float2int_rn(fminf(fmaxf(b * 32767.0f, -32767.0f), 32767.0f));
  return (qa & 0xffff) | (qb << 16);
}

// ---- vectors ----------------------------------------------------------------
template <int P>
__device__ __forceinline__ void load_vec(const VecField& f, int cb, Vec3<typename PrecTraits<P>::Real>& v);
template <>
__device__ __forceinline__ void load_vec<P_F64>(const VecField& f, int cb, Vec3<double>& v) {
  const double2* p = (const double2*)f.data;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double2 a = p[(size_t)k * f.stride + cb];
    v.c[k].re = a.x; v.c[k].im = a.y;
  }
}
template <>
__device__ __forceinline__ void load_vec<P_F32>(const VecField& f, int cb, Vec3<float>& v) {
  float4 a = ((const float4*)f.data)[cb];
  float2 b = ((const float2*)((const float4*)f.data + f.stride))[cb];
  v.c[0].re = a.x; v.c[0].im = a.y; v.c[1].re = a.z; v.c[1].im = a.w; v.c[2].re = b.x; v.c[2].im = b.y;
}
template <>
__device__ __forceinline__ void load_vec<P_H16>(const VecField& f, int cb, Vec3<float>& v) {
  int4 a = ((const int4*)f.data)[cb];
  float n = __int_as_float(a.w) * (1.0f / 32767.0f);
  v.c[0].re = q16lo(a.x) * n; v.c[0].im = q16hi(a.x) * n;
  v.c[1].re = q16lo(a.y) * n; v.c[1].im = q16hi(a.y) * n;
  v.c[2].re = q16lo(a.z) * n; v.c[2].im = q16hi(a.z) * n;
}

template <int P>
__device__ __forceinline__ void store_vec(const VecField& f, int cb, const Vec3<typename PrecTraits<P>::Real>& v);
template <>
__device__ __forceinline__ void store_vec<P_F64>(const VecField& f, int cb, const Vec3<double>& v) {
  double2* p = (double2*)f.data;
#pragma unroll
  for (int k = 0; k < 3; k++) p[(size_t)k * f.stride + cb] = make_double2(v.c[k].re, v.c[k].im);
}
template <>
__device__ __forceinline__ void store_vec<P_F32>(const VecField& f, int cb, const Vec3<float>& v) {
  ((float4*)f.data)[cb] = make_float4(v.c[0].re, v.c[0].im, v.c[1].re, v.c[1].im);
  ((float2*)((float4*)f.data + f.stride))[cb] = make_float2(v.c[2].re, v.c[2].im);
}
template <>
__device__ __forceinline__ void store_vec<P_H16>(const VecField& f, int cb, const Vec3<float>& v) {
  float m = 0.f;
#pragma unroll
  for (int k = 0; k < 3; k++) m = fmaxf(m, fmaxf(fabsf(v.c[k].re), fabsf(v.c[k].im)));
  float inv = m > 0.f ? 1.0f / m : 0.f;
  int4 o;
  o.x = q16pack(v.c[0].re * inv, v.c[0].im * inv);
  o.y = q16pack(v.c[1].re * inv, v.c[1].im * inv);
  o.z = q16pack(v.c[2].re * inv, v.c[2].im * inv);
  o.w = __float_as_int(m);
  ((int4*)f.data)[cb] = o;
}

// Non-temporal form of store_vec for fields that the next kernel reads (the
// stencil and CG outputs): the record goes out without being kept dirty in
// the writing XCD's L2, which measured 4% faster for a stencil whose source
// the previous launch wrote (32^3x64, 16-bit: 96.4 us against 100.7 us).
template <int P>
__device__ __forceinline__ void store_vec_nt(const VecField& f, int cb, const Vec3<typename PrecTraits<P>::Real>& v);
template <>
__device__ __forceinline__ void store_vec_nt<P_F64>(const VecField& f, int cb, const Vec3<double>& v) {
  nat_d2* p = (nat_d2*)f.data;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    nat_d2 a;
    a.x = v.c[k].re;
    a.y = v.c[k].im;
    __builtin_nontemporal_store(a, p + (size_t)k * f.stride + cb);
  }
}
template <>
__device__ __forceinline__ void store_vec_nt<P_F32>(const VecField& f, int cb, const Vec3<float>& v) {
  nat_f4 a;
  a.x = v.c[0].re; a.y = v.c[0].im; a.z = v.c[1].re; a.w = v.c[1].im;
  nat_f2 b;
  b.x = v.c[2].re; b.y = v.c[2].im;
  __builtin_nontemporal_store(a, (nat_f4*)f.data + cb);
  __builtin_nontemporal_store(b, (nat_f2*)((nat_f4*)f.data + f.stride) + cb);
}
template <>
__device__ __forceinline__ void store_vec_nt<P_H16>(const VecField& f, int cb, const Vec3<float>& v) {
  float m = 0.f;
#pragma unroll
  for (int k = 0; k < 3; k++) m = fmaxf(m, fmaxf(fabsf(v.c[k].re), fabsf(v.c[k].im)));
  float inv = m > 0.f ? 1.0f / m : 0.f;
  nat_i4 o;
  o.x = q16pack(v.c[0].re * inv, v.c[0].im * inv);
  o.y = q16pack(v.c[1].re * inv, v.c[1].im * inv);
  o.z = q16pack(v.c[2].re * inv, v.c[2].im * inv);
  o.w = __float_as_int(m);
  __builtin_nontemporal_store(o, (nat_i4*)f.data + cb);
}

// ---- links --------------------------------------------------------------------
template <int P>
__device__ __forceinline__ void load_link(const LinkField& f, int pb, int mu, int cb,
                                          Mat3<typename PrecTraits<P>::Real>& m);
template <>
__device__ __forceinline__ void load_link<P_F64>(const LinkField& f, int pb, int mu, int cb, Mat3<double>& m) {
  const double2* p = (const double2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    nat_d2 a = ldnt(p + (size_t)e * f.stride);
    m.e[e / 3][e % 3].re = a.x; m.e[e / 3][e % 3].im = a.y;
  }
}
template <>
__device__ __forceinline__ void load_link<P_F32>(const LinkField& f, int pb, int mu, int cb, Mat3<float>& m) {
  const float4* p4 = (const float4*)f.data + (size_t)((pb * 4 + mu) * 4) * f.stride + cb;
  const float2* p2 = (const float2*)((const float4*)f.data + (size_t)32 * f.stride) + (size_t)(pb * 4 + mu) * f.stride + cb;
  float r[18];
#pragma unroll
  for (int ch = 0; ch < 4; ch++) {
    nat_f4 a = ldnt(p4 + (size_t)ch * f.stride);
    r[4 * ch] = a.x; r[4 * ch + 1] = a.y; r[4 * ch + 2] = a.z; r[4 * ch + 3] = a.w;
  }
  nat_f2 b = ldnt(p2);
  r[16] = b.x; r[17] = b.y;
#pragma unroll
  for (int e = 0; e < 9; e++) { m.e[e / 3][e % 3].re = r[2 * e]; m.e[e / 3][e % 3].im = r[2 * e + 1]; }
}
// 16-bit link, all nine elements
__device__ __forceinline__ void load_link_h16f(const LinkField& f, int pb, int mu, int cb, Mat3<float>& m) {
  const int4* p4 = (const int4*)f.data + (size_t)((pb * 4 + mu) * 2) * f.stride + cb;
  const int* p1 = (const int*)((const int4*)f.data + (size_t)16 * f.stride) + (size_t)(pb * 4 + mu) * f.stride + cb;
  float s = f.scale * (1.0f / 32767.0f);
  int w[9];
#pragma unroll
  for (int ch = 0; ch < 2; ch++) {
    nat_i4 a = ldnt(p4 + (size_t)ch * f.stride);
    w[4 * ch] = a.x; w[4 * ch + 1] = a.y; w[4 * ch + 2] = a.z; w[4 * ch + 3] = a.w;
  }
  w[8] = ldnt(p1);
#pragma unroll
  for (int e = 0; e < 9; e++) { m.e[e / 3][e % 3].re = q16lo(w[e]) * s; m.e[e / 3][e % 3].im = q16hi(w[e]) * s; }
}

// 16-bit long link in the two-row form: planes of int4 (row-0 elements and
// element (1,0)), int2 (elements (1,1), (1,2)) and short (the phase of z),
// each plane 8 * stride records; row 2 = z conj(row0 x row1) with
// z = exp(i phi) / c, phi = q pi / 32767 and 1/c = (|row0|^2 + |row1|^2)/2)^-1/2
// (the rows of L = c U both have norm c): 26 bytes per link.
__device__ __forceinline__ void load_link_h16r(const LinkField& f, int pb, int mu, int cb, Mat3<float>& m) {
  const size_t S = (size_t)f.stride, k = (size_t)(pb * 4 + mu) * S + cb;
  const int4* p4 = (const int4*)f.data;
  const int2* p2 = (const int2*)(p4 + 8 * S);
  const short* ph = (const short*)(p2 + 8 * S);
  nat_i4 a = ldnt(p4 + k);
  nat_i2 b = ldnt(p2 + k);
  const short zq = __builtin_nontemporal_load(ph + k);
  const float s = f.scale * (1.0f / 32767.0f);
  const int w[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
  float n2 = 0.f;
#pragma unroll
  for (int e = 0; e < 6; e++) {
    m.e[e / 3][e % 3].re = q16lo(w[e]) * s;
    m.e[e / 3][e % 3].im = q16hi(w[e]) * s;
    n2 += m.e[e / 3][e % 3].re * m.e[e / 3][e % 3].re + m.e[e / 3][e % 3].im * m.e[e / 3][e % 3].im;
  }
  const float inv_c = __frsqrt_rn(0.5f * n2);
  float sn, cs;
  __sincosf((float)zq * (3.14159265358979f / 32767.0f), &sn, &cs);
  const float zr = cs * inv_c, zi = sn * inv_c;
#pragma unroll
  for (int k3 = 0; k3 < 3; k3++) {
    const int i = (k3 + 1) % 3, j = (k3 + 2) % 3;
    // cross_k = r0_i r1_j - r0_j r1_i
    const Cplx<float>&a0 = m.e[0][i], &b1 = m.e[1][j], &a1 = m.e[0][j], &b0 = m.e[1][i];
    float cr = a0.re * b1.re - a0.im * b1.im - (a1.re * b0.re - a1.im * b0.im);
    float ci = a0.re * b1.im + a0.im * b1.re - (a1.re * b0.im + a1.im * b0.re);
    // z conj(cross)
    m.e[2][k3].re = zr * cr + zi * ci;
    m.e[2][k3].im = zi * cr - zr * ci;
  }
}

template <>
__device__ __forceinline__ void load_link<P_H16>(const LinkField& f, int pb, int mu, int cb, Mat3<float>& m) {
  if (f.recon) load_link_h16r(f, pb, mu, cb, m);
  else load_link_h16f(f, pb, mu, cb, m);
}

// Store one link; for H16 `inv_scale` = 1/scale of the destination field.
template <int P, typename R>
__device__ __forceinline__ void store_link(const LinkField& f, int pb, int mu, int cb, const Mat3<R>& m,
                                           float inv_scale);
template <>
__device__ __forceinline__ void store_link<P_F64, double>(const LinkField& f, int pb, int mu, int cb,
                                                          const Mat3<double>& m, float) {
  double2* p = (double2*)f.data + (size_t)((pb * 4 + mu) * 9) * f.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) p[(size_t)e * f.stride] = make_double2(m.e[e / 3][e % 3].re, m.e[e / 3][e % 3].im);
}
template <>
__device__ __forceinline__ void store_link<P_F32, double>(const LinkField& f, int pb, int mu, int cb,
                                                          const Mat3<double>& m, float) {
  float4* p4 = (float4*)f.data + (size_t)((pb * 4 + mu) * 4) * f.stride + cb;
  float2* p2 = (float2*)((float4*)f.data + (size_t)32 * f.stride) + (size_t)(pb * 4 + mu) * f.stride + cb;
  float r[18];
#pragma unroll
  for (int e = 0; e < 9; e++) { r[2 * e] = (float)m.e[e / 3][e % 3].re; r[2 * e + 1] = (float)m.e[e / 3][e % 3].im; }
#pragma unroll
  for (int ch = 0; ch < 4; ch++) p4[(size_t)ch * f.stride] = make_float4(r[4 * ch], r[4 * ch + 1], r[4 * ch + 2], r[4 * ch + 3]);
  *p2 = make_float2(r[16], r[17]);
}
template <>
__device__ __forceinline__ void store_link<P_H16, double>(const LinkField& f, int pb, int mu, int cb,
                                                          const Mat3<double>& m, float inv_scale) {
  int4* p4 = (int4*)f.data + (size_t)((pb * 4 + mu) * 2) * f.stride + cb;
  int* p1 = (int*)((int4*)f.data + (size_t)16 * f.stride) + (size_t)(pb * 4 + mu) * f.stride + cb;
  int w[9];
#pragma unroll
  for (int e = 0; e < 9; e++)
    w[e] = q16pack((float)(m.e[e / 3][e % 3].re * inv_scale), (float)(m.e[e / 3][e % 3].im * inv_scale));
  p4[0] = make_int4(w[0], w[1], w[2], w[3]);
  p4[f.stride] = make_int4(w[4], w[5], w[6], w[7]);
  *p1 = w[8];
}

// Sum of NV values over `nblocks` block records parts[b*4 + k] (k < NV) by
// one block: thread t takes records t, t+blockDim, ... with four independent
// 32-byte loads in flight per pass.  Fixed association order, so the result
// is bit-reproducible.  Only thread 0 holds the result.
template <int NV>
__device__ __forceinline__ void reduce_records(const double* __restrict__ parts, int nblocks, double (&out)[NV]) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0.0;
  const double4* p4 = (const double4*)parts;
  const int bd = blockDim.x;
  for (int base = threadIdx.x; base < nblocks; base += 4 * bd) {
    double4 r[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int i = base + u * bd;
      r[u] = i < nblocks ? p4[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      acc[0] += r[u].x;
      if (NV > 1) acc[NV > 1 ? 1 : 0] += r[u].y;
      if (NV > 2) acc[NV > 2 ? 2 : 0] += r[u].z;
      if (NV > 3) acc[NV > 3 ? 3 : 0] += r[u].w;
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, 64);
  __shared__ double sm[NV][16];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sm[k][wid] = acc[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double t = 0.0;
      for (int w = 0; w < (bd >> 6); w++) t += sm[k][w];
      out[k] = t;
    }
  }
}

// reduce_records over two record arrays at once (NA values of pa, NB of pb,
// both with `nblocks` records): the loads of both are in flight together, and
// every value is summed in the same order as reduce_records sums it, so the
// results are bit-identical to two separate calls.  Only thread 0 holds them.
template <int NA, int NB>
__device__ __forceinline__ void reduce_records_pair(const double* __restrict__ pa, const double* __restrict__ pb,
                                                    int nblocks, double (&outa)[NA], double (&outb)[NB]) {
  constexpr int NV = NA + NB;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0.0;
  const double4* a4 = (const double4*)pa;
  const double4* b4 = (const double4*)pb;
  const int bd = blockDim.x;
  for (int base = threadIdx.x; base < nblocks; base += 4 * bd) {
    double4 ra[4], rb[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int i = base + u * bd;
      ra[u] = i < nblocks ? a4[i] : make_double4(0.0, 0.0, 0.0, 0.0);
      rb[u] = i < nblocks ? b4[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double va[4] = {ra[u].x, ra[u].y, ra[u].z, ra[u].w};
      const double vb[4] = {rb[u].x, rb[u].y, rb[u].z, rb[u].w};
#pragma unroll
      for (int k = 0; k < NA; k++) acc[k] += va[k];
#pragma unroll
      for (int k = 0; k < NB; k++) acc[NA + k] += vb[k];
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, 64);
  __shared__ double sm[NV][16];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sm[k][wid] = acc[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double t = 0.0;
      for (int w = 0; w < (bd >> 6); w++) t += sm[k][w];
      if (k < NA) outa[k] = t;
      else outb[k - NA] = t;
    }
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
