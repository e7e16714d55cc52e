// mh_core.hip -- context, field memory, MILC<->device layout, Dslash and CG kernels.
//
// Reference behaviour restated here (MILC 7.8.1):
//   Dslash  generic_ks/dslash_fn_dblstore.c:311-563 (dslash_fn_field_special)
//   CG      generic_ks/d_congrad5_fn_milc.c:60-407 (ks_congrad_parity_cpu, FEWSUMS)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iterator>
#include <map>
#include <unordered_map>
#include <vector>

#include "mh_dslash.h"
#include "mh_internal.h"

namespace mh {

static Context g_ctx;
Context& ctx() { return g_ctx; }

// Storage of r_s, x_s and M^dag M p_s in the 16-bit mixed CG: FP32.  Keeping
// them in 16-bit as well saves 40 B per site and iteration but was measured
// slower at 32^3x64, m = 0.01 (95.8 against 94.6 ms per solve on one box):
// the 16-bit residual recurrence needs 369-401 iterations instead of 365-368.
static constexpr int kH16Arith = P_F32;

size_t vec_bytes(int prec, int stride) { return (size_t)vec_site_bytes(prec) * stride; }
size_t link_bytes(int prec, int stride) { return (size_t)8 * link_site_bytes(prec) * stride; }

HipGraph::~HipGraph() {
  if (e) (void)hipGraphExecDestroy(e);
  if (g) (void)hipGraphDestroy(g);
}

int capture_graph(HipGraph& gr, const std::function<int()>& enqueue) {
  Context& c = ctx();
  MH_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue();
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c.stream, &g);
  gr.g = g;
  if (rc) return rc;
  MH_CHECK(e);
  MH_CHECK(hipGraphInstantiate(&gr.e, gr.g, nullptr, nullptr, 0));
  return 0;
}

// Device memory: freed blocks are kept by size and handed out again, so the
// solvers' per-call fields cost neither hipMalloc nor hipFree's implicit device
// synchronisation.  All work is ordered on the context stream, so a block
// reused by a later call is touched only after the kernels of the earlier one.
// The pool keeps at most a quarter of the device memory and gives everything
// back when an allocation fails or the context ends.
namespace {
struct Pool {
  std::unordered_map<void*, size_t> live;
  std::multimap<size_t, void*> spare;
  size_t spare_bytes = 0, cap = 0;
};
Pool g_pool;

void pool_release_spare(size_t keep) {
  if (g_pool.spare.empty()) return;
  if (ctx().stream) (void)hipStreamSynchronize(ctx().stream);
  // largest blocks first
  while (!g_pool.spare.empty() && g_pool.spare_bytes > keep) {
    auto it = std::prev(g_pool.spare.end());
    (void)hipFree(it->second);
    g_pool.spare_bytes -= it->first;
    g_pool.spare.erase(it);
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free solve (milc_hip_dev_invert_noalloc): one
// file-scope device array of kScratchBytes.  Every device the library runs on
// holds its own copy until the process ends; its address on the current
// device is looked up once per device at init (hipGetSymbolAddress) and only
// that address is handed to kernels.  While an arena over it is active, the
// solver's work vectors are carved from it and any other device allocation is
// refused.  Calls that overlap on one device would share it: the library's
// calls must not overlap (one context per process).
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];

int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return -2;
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  MH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return -5;
  g_scratch_ptr[device] = p;
  return 0;
}

namespace {
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
};
Arena g_arena;
}  // namespace

ScratchArena::ScratchArena() {
  Context& c = ctx();
  void* p = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch_ptr[c.device] : nullptr;
  ok = p != nullptr;
  g_arena.base = (char*)p;
  g_arena.size = ok ? kScratchBytes : 0;
  g_arena.used = 0;
}
ScratchArena::~ScratchArena() { g_arena = Arena(); }

static bool in_arena(const void* p) {
  return g_arena.base && (const char*)p >= g_arena.base && (const char*)p < g_arena.base + g_arena.size;
}

__global__ void k_zero16(uint4* __restrict__ p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Zero `bytes` (a multiple of 16) with a fill kernel on the context stream.
static int zero_fill(void* p, size_t bytes) {
  size_t n = bytes / 16;
  if (!n) return 0;
  unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_zero16, dim3(nb), dim3(256), 0, ctx().stream, (uint4*)p, n);
  MH_CHECK(hipGetLastError());
  return 0;
}

int dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return 0;
  bytes = (bytes + 255) & ~(size_t)255;
  if (g_arena.base) {
    fprintf(stderr, "milc_hip: device allocation of %zu bytes inside the allocation-free path\n", bytes);
    return -7;
  }
  auto it = g_pool.spare.find(bytes);
  if (it != g_pool.spare.end()) {
    *p = it->second;
    g_pool.spare.erase(it);
    g_pool.spare_bytes -= bytes;
    g_pool.live[*p] = bytes;
    return 0;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    pool_release_spare(0);
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    MH_CHECK(e);
  }
  g_pool.live[*p] = bytes;
  return 0;
}

void dev_free(void* p) {
  if (!p || in_arena(p)) return;
  auto it = g_pool.live.find(p);
  if (it == g_pool.live.end()) {
    (void)hipFree(p);
    return;
  }
  if (!g_pool.cap) {
    size_t fr = 0, tot = 0;
    g_pool.cap = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 4 : ((size_t)16 << 30);
  }
  g_pool.spare.emplace(it->second, p);
  g_pool.spare_bytes += it->second;
  g_pool.live.erase(it);
  if (g_pool.spare_bytes > g_pool.cap) pool_release_spare(g_pool.cap / 2);
}

void dev_pool_release() { pool_release_spare(0); }

void* pinned_scratch(size_t bytes) {
  Workspace& w = ctx().ws;
  if (w.pinned_bytes >= bytes) return w.pinned;
  if (g_arena.base) return nullptr;  // no allocation in the allocation-free path
  if (w.pinned) {
    if (ctx().stream) (void)hipStreamSynchronize(ctx().stream);
    (void)hipHostFree(w.pinned);
  }
  w.pinned = nullptr;
  w.pinned_bytes = 0;
  size_t n = std::max<size_t>(bytes, 64 << 10);
  if (hipHostMalloc(&w.pinned, n, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    w.pinned = nullptr;
    return nullptr;
  }
  w.pinned_bytes = n;
  return w.pinned;
}

int alloc_vec(VecField& v, int prec, int stride) {
  v.stride = stride;
  v.norm = nullptr;
  if (g_arena.base) {
    // carve from the scratch arena; zeroed with a fill kernel (no memset node)
    size_t bytes = (vec_bytes(prec, stride) + 255) & ~(size_t)255;
    if (g_arena.used + bytes > g_arena.size) {
      v.data = nullptr;
      fprintf(stderr, "milc_hip: scratch of %zu bytes exhausted (%zu more needed)\n", g_arena.size,
              g_arena.used + bytes - g_arena.size);
      return -6;
    }
    v.data = g_arena.base + g_arena.used;
    g_arena.used += bytes;
    return zero_fill(v.data, bytes);
  }
  int rc = dev_alloc(&v.data, vec_bytes(prec, stride));
  if (rc) return rc;
  MH_CHECK(hipMemsetAsync(v.data, 0, vec_bytes(prec, stride), ctx().stream));
  return 0;
}
void free_vec(VecField& v) {
  dev_free(v.data);
  v.data = nullptr;
}
int alloc_links(LinkField& l, int prec, int stride) {
  l.stride = stride;
  l.scale = 1.0f;
  l.zscale = 1.0f;
  l.recon = 0;
  int rc = dev_alloc(&l.data, link_bytes(prec, stride));
  if (rc) return rc;
  // pooled blocks keep their old contents: zero the field, padding included
  MH_CHECK(hipMemsetAsync(l.data, 0, link_bytes(prec, stride), ctx().stream));
  return 0;
}
void free_links(LinkField& l) {
  dev_free(l.data);
  l.data = nullptr;
}

static inline int nblocks_for(int n, int block) { return (n + block - 1) / block; }

// ===========================================================================
// Layout kernels: MILC AoS (node_index order) <-> per-parity SoA planes
// ===========================================================================
template <int P>
__global__ void k_vec_from_milc(const double* __restrict__ aos, VecField out, int Vh, int pb) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  const double* s = aos + 6 * ((size_t)pb * Vh + cb);
  Vec3<typename PrecTraits<P>::Real> v;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v.c[k].re = (typename PrecTraits<P>::Real)s[2 * k];
    v.c[k].im = (typename PrecTraits<P>::Real)s[2 * k + 1];
  }
  store_vec<P>(out, cb, v);
}
template <int P>
__global__ void k_vec_to_milc(VecField in, double* __restrict__ aos, int Vh, int pb) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<typename PrecTraits<P>::Real> v;
  load_vec<P>(in, cb, v);
  double* d = aos + 6 * ((size_t)pb * Vh + cb);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    d[2 * k] = (double)v.c[k].re;
    d[2 * k + 1] = (double)v.c[k].im;
  }
}

// links: aos[(4*i + mu)*18], i = MILC site index = pb*Vh + cb
template <int P>
__global__ void k_links_from_milc(const double* __restrict__ aos, LinkField out, int Vh, float inv_scale) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;  // pb*4 + mu
  if (cb >= Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  const double* s = aos + 18 * (4 * ((size_t)pb * Vh + cb) + mu);
  Mat3<double> m;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    m.e[e / 3][e % 3].re = s[2 * e];
    m.e[e / 3][e % 3].im = s[2 * e + 1];
  }
  store_link<P, double>(out, pb, mu, cb, m, inv_scale);
}
template <int P>
__global__ void k_links_to_milc(LinkField in, double* __restrict__ aos, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Mat3<typename PrecTraits<P>::Real> m;
  load_link<P>(in, pb, mu, cb, m);
  double* d = aos + 18 * (4 * ((size_t)pb * Vh + cb) + mu);
  for (int e = 0; e < 9; e++) {
    d[2 * e] = (double)m.e[e / 3][e % 3].re;
    d[2 * e + 1] = (double)m.e[e / 3][e % 3].im;
  }
}

// max |element| of an FP64 device link field, for the 16-bit scale
// largest |Re|, |Im| over the valid sites of an FP64 field of n / stride planes
// (the padding of each plane, past Vh, holds whatever the memory held before)
__global__ void k_link_absmax(const double2* __restrict__ p, size_t n, int stride, int Vh, double* out_partial) {
  double m = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if ((int)(i % stride) >= Vh) continue;
    double2 a = p[i];
    m = fmax(m, fmax(fabs(a.x), fabs(a.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double sm[16];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, sm[w]);
    out_partial[blockIdx.x] = m;
  }
}

template <int PI, int PO>
__global__ void k_convert_links(LinkField in, LinkField out, int Vh, float inv_scale) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= Vh) return;
  int pb = pm >> 2, mu = pm & 3;
  Mat3<typename PrecTraits<PI>::Real> m;
  load_link<PI>(in, pb, mu, cb, m);
  Mat3<double> d;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    d.e[e / 3][e % 3].re = (double)m.e[e / 3][e % 3].re;
    d.e[e / 3][e % 3].im = (double)m.e[e / 3][e % 3].im;
  }
  store_link<PO, double>(out, pb, mu, cb, d, inv_scale);
}

// Two-row form of scaled-unitary long links L = c U (c > 0 real, U in U(3)):
// c^2 = |L|_F^2 / 3, z = det(U) / c = det(L) / c^4, row 2 = z conj(row0 x row1).
__device__ __forceinline__ void long_link_z(const Mat3<double>& L, double& c2, double& zr, double& zi,
                                            double& dev) {
  double n2 = 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) n2 += L.e[i][j].re * L.e[i][j].re + L.e[i][j].im * L.e[i][j].im;
  c2 = n2 / 3.0;
  // |L^dag L - c^2 1|_F / c^2
  double d = 0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double re = 0, im = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        re += L.e[k][i].re * L.e[k][j].re + L.e[k][i].im * L.e[k][j].im;
        im += L.e[k][i].re * L.e[k][j].im - L.e[k][i].im * L.e[k][j].re;
      }
      if (i == j) re -= c2;
      d += re * re + im * im;
    }
  dev = c2 > 0 ? sqrt(d) / c2 : 1e300;
  // det(L) along row 0 with the cofactors of rows 1, 2
  double dr = 0, di = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int i = (k + 1) % 3, j = (k + 2) % 3;
    const Cplx<double>&a = L.e[1][i], &b = L.e[2][j], &e = L.e[1][j], &f = L.e[2][i];
    double cr = a.re * b.re - a.im * b.im - (e.re * f.re - e.im * f.im);
    double ci = a.re * b.im + a.im * b.re - (e.re * f.im + e.im * f.re);
    dr += L.e[0][k].re * cr - L.e[0][k].im * ci;
    di += L.e[0][k].re * ci + L.e[0][k].im * cr;
  }
  const double inv = c2 > 0 ? 1.0 / (c2 * c2) : 0.0;
  zr = dr * inv;
  zi = di * inv;
}

// per-block max of the deviation from a scaled unitary matrix and of |z| components
__global__ void __launch_bounds__(256) k_long_check(LinkField in, int Vh, double* __restrict__ partials) {
  double dmax = 0, zmax = 0;
  const size_t n = (size_t)8 * Vh;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int pm = (int)(t / Vh), cb = (int)(t % Vh);
    Mat3<double> L;
    load_link<P_F64>(in, pm >> 2, pm & 3, cb, L);
    double c2, zr, zi, dev;
    long_link_z(L, c2, zr, zi, dev);
    dmax = fmax(dmax, dev);
    zmax = fmax(zmax, fmax(fabs(zr), fabs(zi)));
  }
  for (int off = 32; off > 0; off >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
    zmax = fmax(zmax, __shfl_xor(zmax, off, 64));
  }
  __shared__ double sm[2][4];
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = dmax;
    sm[1][threadIdx.x >> 6] = zmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) {
      dmax = fmax(dmax, sm[0][w]);
      zmax = fmax(zmax, sm[1][w]);
    }
    partials[2 * blockIdx.x] = dmax;
    partials[2 * blockIdx.x + 1] = zmax;
  }
}

// 16-bit two-row store (the layout read by load_link_h16r)
__global__ void k_store_h16r(LinkField in, LinkField out, int Vh, float inv_scale, float inv_zscale) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  if (cb >= Vh) return;
  Mat3<double> L;
  load_link<P_F64>(in, pm >> 2, pm & 3, cb, L);
  double c2, zr, zi, dev;
  long_link_z(L, c2, zr, zi, dev);
  int w[6];
#pragma unroll
  for (int e = 0; e < 6; e++)
    w[e] = q16pack((float)(L.e[e / 3][e % 3].re * inv_scale), (float)(L.e[e / 3][e % 3].im * inv_scale));
  const size_t S = (size_t)out.stride, k = (size_t)pm * S + cb;
  int4* p4 = (int4*)out.data;
  int2* p2 = (int2*)(p4 + 8 * S);
  short* ph = (short*)(p2 + 8 * S);
  p4[k] = make_int4(w[0], w[1], w[2], w[3]);
  p2[k] = make_int2(w[4], w[5]);
  // z = det(U) / c: its modulus 1/c follows from the stored rows, its phase
  // arg det(U) is kept as 16-bit fixed point in units of pi / 32767
  (void)inv_zscale;
  const double phi = atan2(zi, zr);
  ph[k] = (short)__double2int_rn(fmin(fmax(phi * (32767.0 / M_PI), -32767.0), 32767.0));
}

template <int PI, int PO>
__global__ void k_convert_vec(VecField in, VecField out, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<typename PrecTraits<PI>::Real> a;
  load_vec<PI>(in, cb, a);
  Vec3<typename PrecTraits<PO>::Real> b;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    b.c[k].re = (typename PrecTraits<PO>::Real)a.c[k].re;
    b.c[k].im = (typename PrecTraits<PO>::Real)a.c[k].im;
  }
  store_vec<PO>(out, cb, b);
}

#define DISPATCH_PREC(prec, FN, ...)                 \
  do {                                               \
    if ((prec) == P_F64) FN<P_F64>(__VA_ARGS__);     \
    else if ((prec) == P_F32) FN<P_F32>(__VA_ARGS__); \
    else FN<P_H16>(__VA_ARGS__);                     \
  } while (0)

template <int P>
static void launch_vec_from_milc(const double* aos, VecField& out, int pb) {
  Context& c = ctx();
  hipLaunchKernelGGL(k_vec_from_milc<P>, dim3(nblocks_for(c.Vh, 256)), dim3(256), 0, c.stream, aos, out, c.Vh, pb);
}
template <int P>
static void launch_vec_to_milc(const VecField& in, double* aos, int pb) {
  Context& c = ctx();
  hipLaunchKernelGGL(k_vec_to_milc<P>, dim3(nblocks_for(c.Vh, 256)), dim3(256), 0, c.stream, in, aos, c.Vh, pb);
}

int vec_from_milc_dev(const double* dev_aos, VecField& out, int prec, int pb) {
  DISPATCH_PREC(prec, launch_vec_from_milc, dev_aos, out, pb);
  MH_CHECK(hipGetLastError());
  return 0;
}
int vec_to_milc_dev(const VecField& in, int prec, int pb, double* dev_aos) {
  DISPATCH_PREC(prec, launch_vec_to_milc, in, dev_aos, pb);
  MH_CHECK(hipGetLastError());
  return 0;
}

static int ensure_staging(size_t elems) {
  Context& c = ctx();
  if (c.staging_elems >= elems) return 0;
  dev_free(c.staging);
  c.staging = nullptr;
  c.staging_elems = 0;
  int rc = dev_alloc((void**)&c.staging, elems * sizeof(double));
  if (rc) return rc;
  c.staging_elems = elems;
  return 0;
}

int upload_vec_milc(const double* host_aos, VecField& out, int prec, int pb) {
  Context& c = ctx();
  int rc = ensure_staging((size_t)2 * c.Vh * 6);
  if (rc) return rc;
  // only the parity half is needed
  MH_CHECK(hipMemcpyAsync(c.staging + (size_t)pb * c.Vh * 6, host_aos + (size_t)pb * c.Vh * 6,
                          (size_t)c.Vh * 6 * sizeof(double), hipMemcpyHostToDevice, c.stream));
  return vec_from_milc_dev(c.staging, out, prec, pb);
}
int download_vec_milc(const VecField& in, int prec, int pb, double* host_aos) {
  Context& c = ctx();
  int rc = ensure_staging((size_t)2 * c.Vh * 6);
  if (rc) return rc;
  rc = vec_to_milc_dev(in, prec, pb, c.staging);
  if (rc) return rc;
  MH_CHECK(hipMemcpyAsync(host_aos + (size_t)pb * c.Vh * 6, c.staging + (size_t)pb * c.Vh * 6,
                          (size_t)c.Vh * 6 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

static int link_absmax_f64(const LinkField& f, double* out) {
  Context& c = ctx();
  const int nb = 512;
  size_t n = 2ull * 4 * 9 * f.stride;
  hipLaunchKernelGGL(k_link_absmax, dim3(nb), dim3(256), 0, c.stream, (const double2*)f.data, n, f.stride, c.Vh,
                     c.ws.partials);
  std::vector<double> h(nb);
  MH_CHECK(hipMemcpyAsync(h.data(), c.ws.partials, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double m = 0;
  for (double v : h) m = std::max(m, v);
  *out = m;
  return 0;
}

int links_from_milc_dev(const double* dev_aos, LinkField& out, int prec) {
  Context& c = ctx();
  dim3 grid(nblocks_for(c.Vh, 256), 8);
  if (prec == P_F64) {
    hipLaunchKernelGGL(k_links_from_milc<P_F64>, grid, dim3(256), 0, c.stream, dev_aos, out, c.Vh, 1.0f);
  } else if (prec == P_F32) {
    hipLaunchKernelGGL(k_links_from_milc<P_F32>, grid, dim3(256), 0, c.stream, dev_aos, out, c.Vh, 1.0f);
  } else {
    // scale from an FP64 copy
    LinkField tmp;
    int rc = alloc_links(tmp, P_F64, out.stride);
    if (rc) return rc;
    hipLaunchKernelGGL(k_links_from_milc<P_F64>, grid, dim3(256), 0, c.stream, dev_aos, tmp, c.Vh, 1.0f);
    double m;
    rc = link_absmax_f64(tmp, &m);
    free_links(tmp);
    if (rc) return rc;
    out.scale = (float)(m > 0 ? m : 1.0);
    hipLaunchKernelGGL(k_links_from_milc<P_H16>, grid, dim3(256), 0, c.stream, dev_aos, out, c.Vh,
                       (float)(1.0 / out.scale));
  }
  MH_CHECK(hipGetLastError());
  return 0;
}

int links_to_milc_dev(const LinkField& in, int prec, double* dev_aos) {
  Context& c = ctx();
  dim3 grid(nblocks_for(c.Vh, 256), 8);
  if (prec == P_F64) hipLaunchKernelGGL(k_links_to_milc<P_F64>, grid, dim3(256), 0, c.stream, in, dev_aos, c.Vh);
  else if (prec == P_F32) hipLaunchKernelGGL(k_links_to_milc<P_F32>, grid, dim3(256), 0, c.stream, in, dev_aos, c.Vh);
  else hipLaunchKernelGGL(k_links_to_milc<P_H16>, grid, dim3(256), 0, c.stream, in, dev_aos, c.Vh);
  MH_CHECK(hipGetLastError());
  return 0;
}

// FP64 -> lower precision link conversion (the input must be FP64)
int convert_links(const LinkField& in, int pin, LinkField& out, int pout) {
  Context& c = ctx();
  if (pin != P_F64) return -2;
  dim3 grid(nblocks_for(c.Vh, 256), 8);
  if (pout == P_F64) {
    hipLaunchKernelGGL((k_convert_links<P_F64, P_F64>), grid, dim3(256), 0, c.stream, in, out, c.Vh, 1.0f);
  } else if (pout == P_F32) {
    hipLaunchKernelGGL((k_convert_links<P_F64, P_F32>), grid, dim3(256), 0, c.stream, in, out, c.Vh, 1.0f);
  } else {
    double m;
    int rc = link_absmax_f64(in, &m);
    if (rc) return rc;
    out.scale = (float)(m > 0 ? m : 1.0);
    hipLaunchKernelGGL((k_convert_links<P_F64, P_H16>), grid, dim3(256), 0, c.stream, in, out, c.Vh,
                       (float)(1.0 / out.scale));
  }
  MH_CHECK(hipGetLastError());
  return 0;
}

// FP64 long links -> 16-bit two-row form when every link is a real multiple of
// a U(3) matrix to tol (relative); returns 1 (out untouched) when they are not.
int convert_long_h16r(const LinkField& in, LinkField& out, double tol) {
  Context& c = ctx();
  const int nb = 512;
  hipLaunchKernelGGL(k_long_check, dim3(nb), dim3(256), 0, c.stream, in, c.Vh, c.ws.partials);
  MH_CHECK(hipGetLastError());
  std::vector<double> h(2 * nb);
  MH_CHECK(hipMemcpyAsync(h.data(), c.ws.partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double dmax = 0, zmax = 0;
  for (int b = 0; b < nb; b++) {
    dmax = std::max(dmax, h[2 * b]);
    zmax = std::max(zmax, h[2 * b + 1]);
  }
  if (!(dmax <= tol) || !(zmax > 0)) return 1;
  double m;
  int rc = link_absmax_f64(in, &m);
  if (rc) return rc;
  out.scale = (float)(m > 0 ? m : 1.0);
  out.zscale = (float)zmax;
  out.recon = 1;
  dim3 grid(nblocks_for(c.Vh, 256), 8);
  hipLaunchKernelGGL(k_store_h16r, grid, dim3(256), 0, c.stream, in, out, c.Vh, (float)(1.0 / out.scale),
                     (float)(1.0 / out.zscale));
  MH_CHECK(hipGetLastError());
  return 0;
}

int convert_vec(const VecField& in, int pin, VecField& out, int pout) {
  Context& c = ctx();
  dim3 grid(nblocks_for(c.Vh, 256));
#define CV(a, b)                                                                                   \
  if (pin == a && pout == b) {                                                                     \
    hipLaunchKernelGGL((k_convert_vec<a, b>), grid, dim3(256), 0, c.stream, in, out, c.Vh);       \
    MH_CHECK(hipGetLastError());                                                                   \
    return 0;                                                                                      \
  }
  CV(P_F64, P_F64) CV(P_F64, P_F32) CV(P_F64, P_H16) CV(P_F32, P_F64) CV(P_F32, P_F32) CV(P_F32, P_H16)
  CV(P_H16, P_F64) CV(P_H16, P_F32) CV(P_H16, P_H16)
#undef CV
  return -2;
}

// ===========================================================================
// Dslash
//   dst(x) = sum_mu F_mu(x) s(x+mu) + L_mu(x) s(x+3mu)
//                 - F_mu(x-mu)^dag s(x-mu) - L_mu(x-3mu)^dag s(x-3mu)
// One thread per output site; the epilogue variants fuse the CG algebra.
// ===========================================================================
template <int P>
__global__ void __launch_bounds__(256) k_dslash(Geom g, LinkField fat, LinkField lng, VecField src, VecField dst,
                                                int pb, const int* __restrict__ stop) {
  if (stop && *stop) return;
  int bid = xcd_block(blockIdx.x, gridDim.x);
  int cb = bid * blockDim.x + threadIdx.x;
  if (cb >= g.Vh) return;
  Vec3<typename PrecTraits<P>::Real> acc;
  dslash_site<P>(g, fat, lng, src, cb, pb, acc);
  store_vec_nt<P>(dst, cb, acc);
}

// Block reduction of N doubles into partials[blockIdx.x * 4 + k].
template <int N>
__device__ __forceinline__ void block_sum_store(double (&v)[N], double* __restrict__ partials) {
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

// Second half of the normal operator plus the CG dot products
// (d_congrad5_fn_milc.c:281-308):
//   ttt = D tmp - msq_x4 * p   (= -M^dag M p)
//   pkp = <p,ttt>, c_tr = <ttt,r>, c_tt = <ttt,ttt>   (real parts)
// The scalar step runs in its own one-block kernel (k_cg_scalars): folding it
// into the last block of this kernel needs a device-scope release fence in
// every block, which writes back the XCD's L2 and was measured to triple the
// kernel time (16-bit, 32^3x64: 312 us against 107 us).
template <int PL, int PP, int PA>
__global__ void __launch_bounds__(256) k_dslash_mdagm(Geom g, LinkField fat, LinkField lng, VecField tmp, VecField p,
                                                      VecField r, VecField ttt, int pb, double msq_x4,
                                                      double* __restrict__ partials, const int* __restrict__ stop) {
  if (*stop) return;
  typedef typename PrecTraits<PA>::Real RA;
  int bid = xcd_block(blockIdx.x, gridDim.x);
  int cb = bid * blockDim.x + threadIdx.x;
  double s[3] = {0, 0, 0};
  if (cb < g.Vh) {
    Vec3<typename PrecTraits<PL>::Real> acc;
    dslash_site<PL>(g, fat, lng, tmp, cb, pb, acc);
    Vec3<typename PrecTraits<PP>::Real> pv;
    load_vec<PP>(p, cb, pv);
    Vec3<RA> rv, t;
    load_vec<PA>(r, cb, rv);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      t.c[k].re = (RA)((RA)acc.c[k].re - (RA)msq_x4 * (RA)pv.c[k].re);
      t.c[k].im = (RA)((RA)acc.c[k].im - (RA)msq_x4 * (RA)pv.c[k].im);
      s[0] += (double)pv.c[k].re * (double)t.c[k].re + (double)pv.c[k].im * (double)t.c[k].im;
      s[1] += (double)t.c[k].re * (double)rv.c[k].re + (double)t.c[k].im * (double)rv.c[k].im;
      s[2] += (double)t.c[k].re * (double)t.c[k].re + (double)t.c[k].im * (double)t.c[k].im;
    }
    store_vec_nt<PA>(ttt, cb, t);
  }
  block_sum_store<3>(s, partials);
}

// Restart block (d_congrad5_fn_milc.c:177-240): with acc = D tmp,
//   ttt = acc - msq_x4*x; resid = src + ttt (= src - M^dag M x); p = resid;
// |resid|^2 partials go to partials_rr.  Also used for the FP64 true residual
// of the reliable updates (then p/ttt are not written).
template <int P>
__global__ void __launch_bounds__(256) k_dslash_resid(Geom g, LinkField fat, LinkField lng, VecField tmp, VecField x,
                                                      VecField src, VecField resid, VecField p, int write_p,
                                                      int pb, double msq_x4, double* __restrict__ partials_rr) {
  typedef typename PrecTraits<P>::Real R;
  int bid = xcd_block(blockIdx.x, gridDim.x);
  int cb = bid * blockDim.x + threadIdx.x;
  double s[2] = {0, 0};
  if (cb < g.Vh) {
    Vec3<R> acc, xv, bv, r;
    dslash_site<P>(g, fat, lng, tmp, cb, pb, acc);
    load_vec<P>(x, cb, xv);
    load_vec<P>(src, cb, bv);
    double xx = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R tre = acc.c[k].re + (R)(-msq_x4) * xv.c[k].re;
      R tim = acc.c[k].im + (R)(-msq_x4) * xv.c[k].im;
      r.c[k].re = bv.c[k].re + tre;
      r.c[k].im = bv.c[k].im + tim;
      s[0] += (double)r.c[k].re * (double)r.c[k].re + (double)r.c[k].im * (double)r.c[k].im;
      xx += (double)xv.c[k].re * (double)xv.c[k].re + (double)xv.c[k].im * (double)xv.c[k].im;
    }
    // per-site term of the Fermilab relative residual (d_congrad5_fn_milc.c:45-47)
    s[1] = xx == 0.0 ? 1.0 : s[0] / xx;
    store_vec<P>(resid, cb, r);
    if (write_p) store_vec<P>(p, cb, r);
  }
  block_sum_store<2>(s, partials_rr);
}

// after a restart: rsq = actual_rsq = |resid|^2 and the relative residual
// sqrt(sum_i |r_i|^2/|x_i|^2 / Vh) (d_congrad5_fn_milc.c:52-55, one parity);
// reset the device loop state
__global__ void k_restart_scalars(const double* __restrict__ partials_rr, int nblocks, int Vh, CGScalars* sc) {
  double t[2];
  reduce_records<2>(partials_rr, nblocks, t);
  if (threadIdx.x == 0) {
    double rr = t[0];
    sc->relrsq = sqrt(t[1] / (double)Vh);
    sc->rsq = rr;
    sc->actual_rsq = rr;
    sc->rsq_first = rr;
    sc->rsq_max = rr;
    sc->halt = 0;
    sc->go = 0;
    sc->iters = 0;
  }
}

// FEWSUMS scalar step (d_congrad5_fn_milc.c:309-362), thread 0 of the
// calling block, from the reduced sums (pkp, c_tr, c_tt) and |r|^2 of the
// last update:
//   oldrsq = actual_rsq; a = -rsq/pkp; rsq = oldrsq + 2a c_tr + a^2 c_tt; b = rsq/oldrsq
// Halts the loop once rsq < rsq_stop (the target, or for the mixed solver the
// reliable-update trigger).
__device__ __forceinline__ void cg_scalar_step(const double (&d3)[3], const double (&d1)[1], CGScalars* sc) {
  double pkp = d3[0], c_tr = d3[1], c_tt = d3[2], oldrsq = d1[0];
  double a = -sc->rsq / pkp;
  double rsq = oldrsq + 2.0 * a * c_tr + a * a * c_tt;
  double b = rsq / oldrsq;
  sc->oldrsq = oldrsq;
  sc->actual_rsq = oldrsq;
  sc->pkp = pkp;
  sc->c_tr = c_tr;
  sc->c_tt = c_tt;
  sc->a = a;
  sc->b = b;
  sc->rsq = rsq;
  sc->go = 1;
  sc->iters += 1;
  if (rsq > sc->rsq_max) sc->rsq_max = rsq;
  if (!sc->rel_mode && rsq < sc->rsq_stop) sc->halt = 1;
}

// Both record arrays are read in one pass (their loads in flight together)
// before the halt flag is looked at, so the kernel waits for memory once.
__global__ void k_cg_scalars(const double* __restrict__ partials, const double* __restrict__ partials_rr,
                             int nblocks, CGScalars* sc) {
  double d3[3], d1[1];
  reduce_records_pair<3, 1>(partials, partials_rr, nblocks, d3, d1);
  if (sc->halt) {
    if (threadIdx.x == 0) sc->go = 0;
    return;
  }
  if (threadIdx.x == 0) cg_scalar_step(d3, d1, sc);
}

// With a relative-residual target the halt test waits for the update: the
// recurrence residual and my_relative_residue(resid, dest) of the updated
// fields must both be below their targets (d_congrad5_fn_milc.c:177-180,353-354).
__global__ void k_cg_relcheck(const double* __restrict__ partials_rr, int nblocks, int Vh, CGScalars* sc) {
  if (!sc->go) return;
  double t[2];
  reduce_records<2>(partials_rr, nblocks, t);
  if (threadIdx.x == 0) {
    double rel = sqrt(t[1] / (double)Vh);
    sc->relrsq = rel;
    bool r_ok = sc->rsq_stop <= 0.0 || sc->rsq < sc->rsq_stop;
    bool rel_ok = sc->relrsq_stop <= 0.0 || rel < sc->relrsq_stop;
    if (r_ok && rel_ok) sc->halt = 1;
  }
}

// dest += a p; resid += a ttt; |resid|^2 partials; p = resid + b p
// (d_congrad5_fn_milc.c:318-367)
// With REL the per-site terms |r_i|^2/|x_i|^2 of the relative residual of the
// updated fields go to the second record slot (d_congrad5_fn_milc.c:351-352).
template <int PP, int PA, int REL>
__global__ void __launch_bounds__(256) k_cg_update(VecField x, VecField p, VecField r, VecField ttt, int Vh,
                                                   const CGScalars* __restrict__ sc, double* __restrict__ partials_rr) {
  if (!sc->go) return;
  typedef typename PrecTraits<PA>::Real RA;
  typedef typename PrecTraits<PP>::Real RP;
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  double s[1 + REL] = {0};
  if (cb < Vh) {
    RA a = (RA)sc->a, b = (RA)sc->b;
    Vec3<RA> xv, rv, tv;
    Vec3<RP> pv;
    load_vec<PA>(x, cb, xv);
    load_vec<PA>(r, cb, rv);
    load_vec<PA>(ttt, cb, tv);
    load_vec<PP>(p, cb, pv);
    Vec3<RP> pn;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      xv.c[k].re = xv.c[k].re + a * (RA)pv.c[k].re;
      xv.c[k].im = xv.c[k].im + a * (RA)pv.c[k].im;
      rv.c[k].re = rv.c[k].re + a * tv.c[k].re;
      rv.c[k].im = rv.c[k].im + a * tv.c[k].im;
      s[0] += (double)rv.c[k].re * (double)rv.c[k].re + (double)rv.c[k].im * (double)rv.c[k].im;
      pn.c[k].re = (RP)(rv.c[k].re + b * (RA)pv.c[k].re);
      pn.c[k].im = (RP)(rv.c[k].im + b * (RA)pv.c[k].im);
    }
    if (REL) {
      double xx = 0;
#pragma unroll
      for (int k = 0; k < 3; k++)
        xx += (double)xv.c[k].re * (double)xv.c[k].re + (double)xv.c[k].im * (double)xv.c[k].im;
      s[REL] = xx == 0.0 ? 1.0 : s[0] / xx;
    }
    store_vec<PA>(x, cb, xv);
    store_vec<PA>(r, cb, rv);
    store_vec<PP>(p, cb, pn);
  }
  block_sum_store<1 + REL>(s, partials_rr);
}

// Reliable update, vector part: x_hi += x_s; x_s = 0 (x_s in precision PA)
template <int PA>
__global__ void k_accum_x(VecField xhi, VecField xs, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<double> a;
  Vec3<float> b, z;
  load_vec<P_F64>(xhi, cb, a);
  load_vec<PA>(xs, cb, b);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    a.c[k].re += (double)b.c[k].re;
    a.c[k].im += (double)b.c[k].im;
    z.c[k].re = z.c[k].im = 0.f;
  }
  store_vec<P_F64>(xhi, cb, a);
  store_vec<PA>(xs, cb, z);
}
// Reliable update, sloppy vectors: p_s += r_hi - r_s; r_s = r_hi, so that p_s
// equals r_hi + b p_old as if the true residual had been used in the last step
// (r_s in precision PA, p_s in PS).
template <int PS, int PA>
__global__ void k_ru_vectors(VecField rhi, VecField rs, VecField ps, int Vh) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  Vec3<double> rh;
  Vec3<float> r;
  Vec3<typename PrecTraits<PS>::Real> p;
  load_vec<P_F64>(rhi, cb, rh);
  load_vec<PA>(rs, cb, r);
  load_vec<PS>(ps, cb, p);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    p.c[k].re = (float)((double)p.c[k].re + (rh.c[k].re - (double)r.c[k].re));
    p.c[k].im = (float)((double)p.c[k].im + (rh.c[k].im - (double)r.c[k].im));
    r.c[k].re = (float)rh.c[k].re;
    r.c[k].im = (float)rh.c[k].im;
  }
  store_vec<PA>(rs, cb, r);
  store_vec<PS>(ps, cb, p);
}

template <int P>
__global__ void k_norm2(VecField v, int Vh, double* __restrict__ partials) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  double s[1] = {0};
  if (cb < Vh) {
    Vec3<typename PrecTraits<P>::Real> a;
    load_vec<P>(v, cb, a);
#pragma unroll
    for (int k = 0; k < 3; k++) s[0] += (double)a.c[k].re * a.c[k].re + (double)a.c[k].im * a.c[k].im;
  }
  block_sum_store<1>(s, partials);
}

// ===========================================================================
// host side
// ===========================================================================
static int dslash_grid() { return nblocks_for(ctx().Vh, 256); }

template <int P>
static void launch_dslash(const DevLinks& L, const VecField& src, VecField& dst, int pb, const int* stop) {
  Context& c = ctx();
  hipLaunchKernelGGL(k_dslash<P>, dim3(dslash_grid()), dim3(256), 0, c.stream, c.geom, L.fat, L.lng, src, dst,
                     pb, stop);
}

int dslash(const DevLinks& L, const VecField& src, VecField& dst, int pb_out, int prec) {
  if (!L.valid || L.prec != prec) return -3;
  DISPATCH_PREC(prec, launch_dslash, L, src, dst, pb_out, (const int*)nullptr);
  MH_CHECK(hipGetLastError());
  return 0;
}

int vec_norm2(const VecField& v, int prec, double* out) {
  Context& c = ctx();
  int nb = nblocks_for(c.Vh, 256);
  if (prec == P_F64) hipLaunchKernelGGL(k_norm2<P_F64>, dim3(nb), dim3(256), 0, c.stream, v, c.Vh, c.ws.partials);
  else if (prec == P_F32) hipLaunchKernelGGL(k_norm2<P_F32>, dim3(nb), dim3(256), 0, c.stream, v, c.Vh, c.ws.partials);
  else hipLaunchKernelGGL(k_norm2<P_H16>, dim3(nb), dim3(256), 0, c.stream, v, c.Vh, c.ws.partials);
  MH_CHECK(hipGetLastError());
  size_t bytes = (size_t)nb * 4 * sizeof(double);
  double* h = (double*)pinned_scratch(bytes);
  if (!h) return -5;
  MH_CHECK(hipMemcpyAsync(h, c.ws.partials, bytes, hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double s = 0;
  for (int i = 0; i < nb; i++) s += h[(size_t)i * 4];
  *out = s;
  return 0;
}

int vec_zero(VecField& v, int prec) {
  Context& c = ctx();
  MH_CHECK(hipMemsetAsync(v.data, 0, vec_bytes(prec, v.stride), c.stream));
  return 0;
}
int vec_copy(const VecField& in, VecField& out, int prec) {
  Context& c = ctx();
  MH_CHECK(hipMemcpyAsync(out.data, in.data, vec_bytes(prec, in.stride), hipMemcpyDeviceToDevice, c.stream));
  return 0;
}

// The scalar block moves through its pinned host mirror in both directions.
static int read_scalars(CGScalars* h) {
  Context& c = ctx();
  MH_CHECK(hipMemcpyAsync(c.ws.scal_host, c.ws.scal, sizeof(CGScalars), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  *h = *c.ws.scal_host;
  return 0;
}
static int write_scalars(const CGScalars* h) {
  Context& c = ctx();
  // the mirror may still feed an earlier copy
  MH_CHECK(hipStreamSynchronize(c.stream));
  *c.ws.scal_host = *h;
  MH_CHECK(hipMemcpyAsync(c.ws.scal, c.ws.scal_host, sizeof(CGScalars), hipMemcpyHostToDevice, c.stream));
  return 0;
}

// partials_rr lives in the second half of the partials buffer
static double* partials_rr() { return ctx().ws.partials + (size_t)ctx().ws.max_blocks * 4; }

// True residual resid = src - M^dag M x on parity pb with FP64 links; p = resid
// if write_p.  Leaves rsq in the device scalars (restart semantics).
static int true_residual(const VecField& src, const VecField& x, VecField& tmp_other, VecField& resid, VecField& p,
                         int write_p, int pb, double msq_x4) {
  Context& c = ctx();
  const DevLinks& L = c.links[P_F64];
  if (!L.valid) return -3;
  launch_dslash<P_F64>(L, x, tmp_other, pb ^ 1, nullptr);
  hipLaunchKernelGGL(k_dslash_resid<P_F64>, dim3(dslash_grid()), dim3(256), 0, c.stream, c.geom, L.fat, L.lng,
                     tmp_other, x, src, resid, p, write_p, pb, msq_x4, partials_rr());
  hipLaunchKernelGGL(k_restart_scalars, dim3(1), dim3(1024), 0, c.stream, partials_rr(), dslash_grid(), c.Vh,
                     c.ws.scal);
  MH_CHECK(hipGetLastError());
  return 0;
}

template <int PL, int PP, int PA>
static void enqueue_cg_iteration(const DevLinks& L, VecField& p, VecField& tmp, VecField& r, VecField& ttt,
                                 VecField& x, int pb, double msq_x4, bool rel = false) {
  Context& c = ctx();
  int nb = dslash_grid();
  const int* halt = &c.ws.scal->halt;
  hipLaunchKernelGGL(k_dslash<PL>, dim3(nb), dim3(256), 0, c.stream, c.geom, L.fat, L.lng, p, tmp, pb ^ 1, halt);
  hipLaunchKernelGGL((k_dslash_mdagm<PL, PP, PA>), dim3(nb), dim3(256), 0, c.stream, c.geom, L.fat, L.lng, tmp, p,
                     r, ttt, pb, msq_x4, c.ws.partials, halt);
  hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(1024), 0, c.stream, c.ws.partials, partials_rr(), nb, c.ws.scal);
  int nu = nblocks_for(c.Vh, 256);
  if (!rel) {
    hipLaunchKernelGGL((k_cg_update<PP, PA, 0>), dim3(nu), dim3(256), 0, c.stream, x, p, r, ttt, c.Vh, c.ws.scal,
                       partials_rr());
  } else {
    hipLaunchKernelGGL((k_cg_update<PP, PA, 1>), dim3(nu), dim3(256), 0, c.stream, x, p, r, ttt, c.Vh, c.ws.scal,
                       partials_rr());
    hipLaunchKernelGGL(k_cg_relcheck, dim3(1), dim3(1024), 0, c.stream, partials_rr(), nu, c.Vh, c.ws.scal);
  }
}

// Iterations to enqueue before the next host check.  The recurrence residual
// of a CG phase falls roughly geometrically, so from the rate of the batches
// so far (log rsq per iteration, < 0) the iterations still needed to reach
// rsq_stop are estimated.  Three quarters of that (plus one) is enqueued,
// between 4 and `cap`: an iteration queued after the halt still costs its
// four dispatches (about 13 us at 32^3x64), an extra host check about twice
// that, and the residual's fall is not smooth, so the estimate is taken low.
// Without a usable rate the default batch is kept.
static int next_batch(int dflt, double log_rate, double rsq_now, double rsq_stop, int cap = 64) {
  if (!(log_rate < 0) || !(rsq_now > 0) || !(rsq_stop > 0)) return std::min(dflt, cap);
  double need = std::log(rsq_stop / rsq_now) / log_rate;
  if (!(need > 0)) return 4;
  return (int)std::max(4.0, std::min((double)cap, std::ceil(0.75 * need) + 1.0));
}
static double batch_rate(double log_rate, double rsq_before, double rsq_now, int done_now) {
  if (done_now <= 0 || !(rsq_now > 0) || !(rsq_before > rsq_now)) return log_rate;
  return std::log(rsq_now / rsq_before) / done_now;
}

// Plain FP64 CG, MILC's restart/iteration semantics (d_congrad5_fn_milc.c:60-407),
// with the optional Fermilab relative residual: its target is met when
// my_relative_residue < relresid^2 (relrsqmin, :92), tested beside the
// residual at every restart block (:213-220) and after every update (:351-354).
static int cg_solve_f64(const VecField& src, VecField& x, double mass, const SolveParams& sp, SolveStats& st) {
  Context& c = ctx();
  const DevLinks& L = c.links[P_F64];
  if (!L.valid) return -3;
  int pb = pbit(sp.parity);
  double msq_x4 = 4.0 * mass * mass;
  double rsqmin = sp.resid * sp.resid;
  double relrsqmin = sp.relresid * sp.relresid;
  const bool rel = relrsqmin > 0;
  int niter = sp.max_iters, max_restarts = sp.max_restarts;
  long max_cg = (long)max_restarts * niter;
  st = SolveStats();
  double source_norm;
  int rc = vec_norm2(src, P_F64, &source_norm);
  if (rc) return rc;
  if (source_norm == 0.0) {
    st.converged = 1;
    return vec_zero(x, P_F64);
  }
  VecField ttt, p, r, tmp;
  VecScope owned{&ttt, &p, &r, &tmp};
  if ((rc = alloc_vec(ttt, P_F64, c.stride)) || (rc = alloc_vec(p, P_F64, c.stride)) ||
      (rc = alloc_vec(r, P_F64, c.stride)) || (rc = alloc_vec(tmp, P_F64, c.stride)))
    return rc;
  CGScalars h;
  memset(&h, 0, sizeof h);
  h.source_norm = source_norm;
  h.rsq_stop = rsqmin * source_norm;
  h.relrsq_stop = relrsqmin;
  h.rel_mode = rel ? 1 : 0;
  if ((rc = write_scalars(&h))) return rc;
  long iteration = 0;
  int nrestart = 0;
  double final_rsq = 0, final_relrsq = 0, size_r = 0;
  int batch = sp.batch > 0 ? sp.batch : 16;
  double log_rate = 0.0;  // log of the recurrence residual's fall per iteration
  auto met = [&](double r2, double rel2) {
    return (rsqmin <= 0 || rsqmin > r2) && (relrsqmin <= 0 || relrsqmin > rel2);
  };
  for (;;) {
    // restart block: true residual (and relative residual)
    if ((rc = true_residual(src, x, tmp, r, p, 1, pb, msq_x4))) break;
    if ((rc = read_scalars(&h))) break;
    final_rsq = h.rsq / source_norm;
    final_relrsq = h.relrsq;
    iteration++;
    if (iteration >= max_cg || nrestart >= max_restarts || met(final_rsq, final_relrsq)) break;
    nrestart++;
    // normal iterations until the next restart point or the recurrence target
    long left = niter - (iteration % niter);
    if (left == niter) left = 0;
    left = std::min(left, max_cg - iteration);
    int nb_next = next_batch(batch, log_rate, h.rsq, h.rsq_stop, 16);
    while (left > 0) {
      int nb = (int)std::min<long>(nb_next, left);
      for (int k = 0; k < nb; k++) enqueue_cg_iteration<P_F64, P_F64, P_F64>(L, p, tmp, r, ttt, x, pb, msq_x4, rel);
      MH_CHECK(hipGetLastError());
      int before = h.iters;
      double rsq_before = h.rsq;
      if ((rc = read_scalars(&h))) break;
      int done_now = h.iters - before;
      iteration += done_now;
      left -= done_now;
      size_r = h.rsq / source_norm;
      if (h.halt) break;
      log_rate = batch_rate(log_rate, rsq_before, h.rsq, done_now);
      nb_next = next_batch(batch, log_rate, h.rsq, h.rsq_stop);
    }
    if (rc) break;
  }
  st.final_rsq = final_rsq;
  st.final_relrsq = final_relrsq;
  st.size_r = size_r;
  st.iters = (int)iteration;
  st.restarts = nrestart;
  // converged = the FP64 true residuals met the targets (the GPU glue's test,
  // d_congrad5_fn_gpu.c:153, with the relative residual of :92 when asked for)
  st.converged = met(final_rsq, final_relrsq) ? 1 : 0;
  return rc;
}

// Mixed precision CG: sloppy (PS = F32 or H16 links and p, F32 r/x/Ap)
// iterations with the FEWSUMS recurrence, FP64 reliable updates when the
// sloppy residual has fallen by delta (or reached the target), FP64 true
// residual (and relative residual, when asked for) for the stopping test.
// PA: precision of the sloppy solution, residual and M^dag M p (r_s, x_s,
// ttt_s).
template <int PS, int PA>
static int cg_solve_mixed(const VecField& src, VecField& x, double mass, const SolveParams& sp, SolveStats& st) {
  Context& c = ctx();
  const DevLinks& LS = c.links[PS];
  if (!LS.valid || !c.links[P_F64].valid) return -3;
  int pb = pbit(sp.parity);
  double msq_x4 = 4.0 * mass * mass;
  double rsqmin = sp.resid * sp.resid;
  double relrsqmin = sp.relresid * sp.relresid;
  long max_cg = (long)sp.max_restarts * sp.max_iters;
  double delta = sp.delta > 0 ? sp.delta : kReliableDelta;
  st = SolveStats();
  double source_norm;
  int rc = vec_norm2(src, P_F64, &source_norm);
  if (rc) return rc;
  if (source_norm == 0.0) {
    st.converged = 1;
    return vec_zero(x, P_F64);
  }
  VecField rhi, tmp64, ps, tmps, ttts, rs, xs, dummy{};
  VecScope owned{&rhi, &tmp64, &ps, &tmps, &ttts, &rs, &xs};
  if ((rc = alloc_vec(rhi, P_F64, c.stride)) || (rc = alloc_vec(tmp64, P_F64, c.stride)) ||
      (rc = alloc_vec(ps, PS, c.stride)) || (rc = alloc_vec(tmps, PS, c.stride)) ||
      (rc = alloc_vec(ttts, PA, c.stride)) || (rc = alloc_vec(rs, PA, c.stride)) ||
      (rc = alloc_vec(xs, PA, c.stride)))
    return rc;
  CGScalars h;
  memset(&h, 0, sizeof h);
  h.source_norm = source_norm;
  if ((rc = write_scalars(&h))) return rc;
  int nb = nblocks_for(c.Vh, 256);
  long iteration = 0;
  int nru = 0;
  double final_rsq = 0, final_relrsq = 0;
  int batch = sp.batch > 0 ? sp.batch : 16;
  double log_rate = 0.0;  // log of the recurrence residual's fall per iteration
  bool first = true;
  auto met = [&](double r2, double rel2) {
    return (rsqmin <= 0 || rsqmin > r2) && (relrsqmin <= 0 || relrsqmin > rel2);
  };
  for (;;) {
    // reliable update (or the initial residual): FP64 true residual
    if (!first) hipLaunchKernelGGL(k_accum_x<PA>, dim3(nb), dim3(256), 0, c.stream, x, xs, c.Vh);
    if ((rc = true_residual(src, x, tmp64, rhi, dummy, 0, pb, msq_x4))) break;
    if (first) {
      if ((rc = convert_vec(rhi, P_F64, rs, PA)) || (rc = convert_vec(rhi, P_F64, ps, PS))) break;
    } else {
      hipLaunchKernelGGL((k_ru_vectors<PS, PA>), dim3(nb), dim3(256), 0, c.stream, rhi, rs, ps, c.Vh);
      nru++;
    }
    first = false;
    if ((rc = read_scalars(&h))) break;
    final_rsq = h.rsq / source_norm;
    final_relrsq = h.relrsq;
    iteration++;
    if (met(final_rsq, final_relrsq) || iteration >= max_cg) break;
    // sloppy iterations until |r|^2 fell by delta^2 or reached the target
    // (once the residual target is met and only the relative residual is
    // left, by delta^2 alone)
    bool r_done = rsqmin <= 0 || rsqmin > final_rsq;
    h.rsq_stop = r_done ? delta * delta * h.rsq : std::max(rsqmin * source_norm, delta * delta * h.rsq);
    h.rel_mode = 0;
    h.halt = 0;
    h.iters = 0;
    if ((rc = write_scalars(&h))) break;
    long left = max_cg - iteration;
    int n_next = next_batch(batch, log_rate, h.rsq, h.rsq_stop, 16);
    while (left > 0) {
      int n = (int)std::min<long>(n_next, left);
      for (int k = 0; k < n; k++) enqueue_cg_iteration<PS, PS, PA>(LS, ps, tmps, rs, ttts, xs, pb, msq_x4);
      MH_CHECK(hipGetLastError());
      int before = h.iters;
      double rsq_before = h.rsq;
      if ((rc = read_scalars(&h))) break;
      int done_now = h.iters - before;
      iteration += done_now;
      left -= done_now;
      if (h.halt) break;
      log_rate = batch_rate(log_rate, rsq_before, h.rsq, done_now);
      n_next = next_batch(batch, log_rate, h.rsq, h.rsq_stop);
    }
    if (rc) break;
  }
  st.final_rsq = final_rsq;
  st.final_relrsq = final_relrsq;
  st.size_r = final_rsq;
  st.iters = (int)iteration;
  st.restarts = nru;
  st.reliable_updates = nru;
  st.converged = met(final_rsq, final_relrsq) ? 1 : 0;
  return rc;
}

// Per-kernel times of the 16-bit mixed CG iteration inside the solve's own
// sequence: from the source's residual (x = 0) the iteration kernels run as in
// cg_solve_mixed, never halting, and HIP events between the kernels time each
// one over `reps` iterations (the events add their own small gaps); then the
// same number of iterations runs without events for the time per iteration.
// us[0..4] = stencil D p, stencil + M^dag M p dots, scalars, update, whole
// iteration.
int cg_iteration_profile(const VecField& src, double mass, int parity, int reps, double us[5]) {
  constexpr int PS = P_H16, PA = kH16Arith;
  Context& c = ctx();
  const DevLinks& LS = c.links[PS];
  if (!LS.valid || !c.links[P_F64].valid) return -3;
  if (reps < 1 || reps > 1000) return -2;
  const int pb = pbit(parity);
  const double msq_x4 = 4.0 * mass * mass;
  VecField x64, rhi, tmp64, ps, tmps, ttts, rs, xs, dummy{};
  VecScope owned{&x64, &rhi, &tmp64, &ps, &tmps, &ttts, &rs, &xs};
  int rc;
  if ((rc = alloc_vec(x64, P_F64, c.stride)) || (rc = alloc_vec(rhi, P_F64, c.stride)) ||
      (rc = alloc_vec(tmp64, P_F64, c.stride)) || (rc = alloc_vec(ps, PS, c.stride)) ||
      (rc = alloc_vec(tmps, PS, c.stride)) || (rc = alloc_vec(ttts, PA, c.stride)) ||
      (rc = alloc_vec(rs, PA, c.stride)) || (rc = alloc_vec(xs, PA, c.stride)))
    return rc;
  if ((rc = vec_zero(x64, P_F64))) return rc;
  CGScalars h;
  memset(&h, 0, sizeof h);
  if ((rc = vec_norm2(src, P_F64, &h.source_norm))) return rc;
  if ((rc = write_scalars(&h))) return rc;
  if ((rc = true_residual(src, x64, tmp64, rhi, dummy, 0, pb, msq_x4))) return rc;
  if ((rc = convert_vec(rhi, P_F64, rs, PA)) || (rc = convert_vec(rhi, P_F64, ps, PS))) return rc;
  if ((rc = read_scalars(&h))) return rc;
  h.rsq_stop = 0.0;  // never halts
  h.rel_mode = 0;
  h.halt = 0;
  h.iters = 0;
  if ((rc = write_scalars(&h))) return rc;
  for (int k = 0; k < 4; k++) enqueue_cg_iteration<PS, PS, PA>(LS, ps, tmps, rs, ttts, xs, pb, msq_x4);
  std::vector<hipEvent_t> ev(4 * reps + 1);
  for (auto& e : ev) MH_CHECK(hipEventCreate(&e));
  OnExit events([&] {
    for (auto& e : ev) (void)hipEventDestroy(e);
  });
  const int nb = dslash_grid(), nu = nblocks_for(c.Vh, 256);
  const int* halt = &c.ws.scal->halt;
  MH_CHECK(hipEventRecord(ev[0], c.stream));
  for (int i = 0; i < reps; i++) {
    hipLaunchKernelGGL(k_dslash<PS>, dim3(nb), dim3(256), 0, c.stream, c.geom, LS.fat, LS.lng, ps, tmps, pb ^ 1, halt);
    MH_CHECK(hipEventRecord(ev[4 * i + 1], c.stream));
    hipLaunchKernelGGL((k_dslash_mdagm<PS, PS, PA>), dim3(nb), dim3(256), 0, c.stream, c.geom, LS.fat, LS.lng, tmps,
                       ps, rs, ttts, pb, msq_x4, c.ws.partials, halt);
    MH_CHECK(hipEventRecord(ev[4 * i + 2], c.stream));
    hipLaunchKernelGGL(k_cg_scalars, dim3(1), dim3(1024), 0, c.stream, c.ws.partials, partials_rr(), nb, c.ws.scal);
    MH_CHECK(hipEventRecord(ev[4 * i + 3], c.stream));
    hipLaunchKernelGGL((k_cg_update<PS, PA, 0>), dim3(nu), dim3(256), 0, c.stream, xs, ps, rs, ttts, c.Vh, c.ws.scal,
                       partials_rr());
    MH_CHECK(hipEventRecord(ev[4 * i + 4], c.stream));
  }
  MH_CHECK(hipGetLastError());
  MH_CHECK(hipEventSynchronize(ev[4 * reps]));
  for (int k = 0; k < 5; k++) us[k] = 0.0;
  for (int i = 0; i < reps; i++)
    for (int k = 0; k < 4; k++) {
      float ms = 0;
      MH_CHECK(hipEventElapsedTime(&ms, ev[4 * i + k], ev[4 * i + k + 1]));
      us[k] += 1000.0 * ms / reps;
    }
  MH_CHECK(hipEventRecord(ev[0], c.stream));
  for (int i = 0; i < reps; i++) enqueue_cg_iteration<PS, PS, PA>(LS, ps, tmps, rs, ttts, xs, pb, msq_x4);
  MH_CHECK(hipEventRecord(ev[1], c.stream));
  MH_CHECK(hipEventSynchronize(ev[1]));
  float ms = 0;
  MH_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
  us[4] = 1000.0 * ms / reps;
  return read_scalars(&h);
}

// The work vectors cg_solve_f64 / cg_solve_mixed allocate, 256-byte aligned.
size_t cg_scratch_bytes(int ps, int stride) {
  auto a = [&](int prec) { return (vec_bytes(prec, stride) + 255) & ~(size_t)255; };
  if (ps == P_F64) return 4 * a(P_F64);
  return 2 * a(P_F64) + 2 * a(ps) + 3 * a(ps == P_H16 ? kH16Arith : P_F32);
}

int cg_solve(const VecField& src, VecField& x, double mass, const SolveParams& sp, SolveStats& st) {
  auto t0 = std::chrono::steady_clock::now();
  int rc;
  if (sp.sloppy_prec == P_F64) rc = cg_solve_f64(src, x, mass, sp, st);
  else if (sp.sloppy_prec == P_F32) rc = cg_solve_mixed<P_F32, P_F32>(src, x, mass, sp, st);
  else rc = cg_solve_mixed<P_H16, kH16Arith>(src, x, mass, sp, st);
  (void)hipStreamSynchronize(ctx().stream);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // namespace mh
