// mh_prn.hip -- MILC's per-site random numbers on the device, so that momenta
// and pseudofermion sources are MILC's own, draw for draw (SITERAND builds).
//
// Restates (MILC 7.8.1):
//   initialize_prn / myrand     generic/ranstuff.c:66-115 (default variant):
//                               7-word 24-bit feedback shift register XORed with
//                               bits 40..63 of a 64-bit linear congruential
//                               generator, both seeded from (iseed, index) by a
//                               32-bit congruential sequence
//   initialize_site_prn_from_seed  generic/ranstuff.c:117-131, index = the
//                               site's lexicographic index (make_lattice.c:49)
//   complex_gaussian_rand_no    libraries/gaussrand.c (default variant): polar
//                               Box-Muller, weight exp(-x^2), (v2 fac, v1 fac)
//   random_anti_hermitian       libraries/rand_ahmat.c (default variant)
//   ranmom                      generic/ranmom.c:10-41 (every site, XUP..TUP)
//   grsource_imp_rhmc draws     ks_imp_rhmc/grsource_rhmc.c:30-37 (three complex
//                               draws per site of the parity) and its
//                               "GRSOURCE: sum" (FAST magsq_su3vec,
//                               libraries/msq_su3vec.c:38-48)
//
// single = 1 rounds as a MILC_PRECISION=1 build: the rejection radius, fac and
// the products in float (no contraction), fac from a double sqrt and log;
// single = 0 as a MILC_PRECISION=2 build.  Every site owns its generator, so
// one thread per site draws in MILC's order without any coordination.
#include <cmath>
#include <vector>

#include "mh_internal.h"

namespace mh {

namespace {

struct SitePrn {
  uint32_t r[7];
  uint32_t mult;
  uint64_t ic;
};

__host__ __device__ inline void prn_init(SitePrn& p, int seed, int index) {
  uint32_t a = 69607u + 8u * (uint32_t)index, s = (uint32_t)seed;
  for (int k = 0; k < 7; k++) {
    s = a * s + 12345u;
    p.r[k] = (s >> 8) & 0xffffffu;
  }
  s = a * s + 12345u;
  p.mult = 100000005u + 8u * (uint32_t)index;
  p.ic = (uint64_t)(int64_t)(int32_t)s;
}

__host__ __device__ inline uint32_t prn_next(SitePrn& p) {
  uint32_t t = (((p.r[5] >> 7) | (p.r[6] << 17)) ^ ((p.r[4] >> 1) | (p.r[5] << 23))) & 0xffffffu;
#pragma unroll
  for (int k = 6; k > 0; k--) p.r[k] = p.r[k - 1];
  p.r[0] = t;
  p.ic = p.ic * (uint64_t)p.mult + 12345u;
  return t ^ (uint32_t)((p.ic >> 40) & 0xffffffu);
}

// one complex draw; v = 2u - 1 is exact in float for the 24-bit u
__device__ inline void cgauss_single(SitePrn& p, double& re, double& im) {
#pragma clang fp contract(off)
  float v1, v2, r;
  do {
    v1 = (float)(2.0 * ((double)prn_next(p) * (1.0 / 16777216.0)) - 1.0);
    v2 = (float)(2.0 * ((double)prn_next(p) * (1.0 / 16777216.0)) - 1.0);
    float a = v1 * v1;
    float b = v2 * v2;
    r = a + b;
  } while (r >= 1.0f);
  float fac = (float)sqrt(-log((double)r) / (double)r);
  float x = v2 * fac;
  float y = v1 * fac;
  re = x;
  im = y;
}

__device__ inline void cgauss_double(SitePrn& p, double& re, double& im) {
#pragma clang fp contract(off)
  double v1, v2, r;
  do {
    v1 = 2.0 * ((double)prn_next(p) * (1.0 / 16777216.0)) - 1.0;
    v2 = 2.0 * ((double)prn_next(p) * (1.0 / 16777216.0)) - 1.0;
    double a = v1 * v1;
    double b = v2 * v2;
    r = a + b;
  } while (r >= 1.0);
  double fac = sqrt(-log(r) / r);
  re = v2 * fac;
  im = v1 * fac;
}

__device__ inline void cgauss(SitePrn& p, int single, double& re, double& im) {
  if (single)
    cgauss_single(p, re, im);
  else
    cgauss_double(p, re, im);
}

__global__ void k_prn_seed(Geom g, int iseed, SitePrn* __restrict__ st) {
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pb = blockIdx.y;
  if (cb >= g.Vh) return;
  Coord c = site_coords(g, cb, pb);
  int index = c.x + g.X[0] * (c.y + g.X[1] * (c.z + g.X[2] * c.t));
  SitePrn p;
  prn_init(p, iseed, index);
  st[(size_t)pb * g.Vh + cb] = p;
}

// ranmom: packed anti_hermitmat [node_index][dir][10] = m01, m02, m12 (re, im),
// m00im, m11im, m22im, 0
__global__ void k_ranmom_prn(Geom g, SitePrn* __restrict__ st, int single, double* __restrict__ mom) {
#pragma clang fp contract(off)
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  int pb = blockIdx.y;
  if (cb >= g.Vh) return;
  size_t i = (size_t)pb * g.Vh + cb;
  SitePrn p = st[i];
  for (int dir = 0; dir < 4; dir++) {
    double* m = mom + (i * 4 + dir) * 10;
    double a, b;
    cgauss(p, single, a, b);
    if (single) {
      float s3 = (float)sqrt(1.0 / 3.0), fa = (float)a, fb = (float)b;
      float q = s3 * fb;
      float d0 = fa + q;
      float d1 = -fa + q;
      m[6] = d0;
      m[7] = d1;
      m[8] = (float)(-2.0 * (double)s3 * (double)fb);
    } else {
      double s3 = sqrt(1.0 / 3.0);
      double q = s3 * b;
      m[6] = a + q;
      m[7] = -a + q;
      m[8] = -2.0 * s3 * b;
    }
    for (int k = 0; k < 3; k++) cgauss(p, single, m[2 * k], m[2 * k + 1]);
    m[9] = 0.0;
  }
  st[i] = p;
}

// three complex draws per site of parity pb into v; per-block partials of the
// sum of |v|^2 per site (in single mode rounded per site as MILC's FAST
// magsq_su3vec)
__global__ void k_gaussian_prn(Geom g, SitePrn* __restrict__ st, int single, int pb, VecField v,
                               double* __restrict__ partials) {
#pragma clang fp contract(off)
  int cb = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0;
  if (cb < g.Vh) {
    size_t i = (size_t)pb * g.Vh + cb;
    SitePrn p = st[i];
    Vec3<double> a;
    for (int k = 0; k < 3; k++) cgauss(p, single, a.c[k].re, a.c[k].im);
    st[i] = p;
    store_vec<P_F64>(v, cb, a);
    if (single) {
      float f = 0.0f;
      for (int k = 0; k < 3; k++) {
        float x = (float)a.c[k].re, y = (float)a.c[k].im;
        float xx = x * x;
        f = f + xx;
        float yy = y * y;
        f = f + yy;
      }
      s = f;
    } else {
      for (int k = 0; k < 3; k++) s += a.c[k].re * a.c[k].re + a.c[k].im * a.c[k].im;
    }
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

struct PrnState {
  SitePrn* st = nullptr;
  size_t nsites = 0;
  int single = 1;
  SitePrn node{};
  bool valid = false;
};
PrnState g_prn;

}  // namespace

int prn_seed_dev(int iseed, int single) {
  Context& c = ctx();
  size_t n = (size_t)2 * c.Vh;
  if (g_prn.st && g_prn.nsites != n) prn_release();
  if (!g_prn.st) {
    int rc = dev_alloc((void**)&g_prn.st, n * sizeof(SitePrn));
    if (rc) return rc;
    g_prn.nsites = n;
  }
  hipLaunchKernelGGL(k_prn_seed, dim3((c.Vh + 255) / 256, 2), dim3(256), 0, c.stream, c.geom, iseed, g_prn.st);
  MH_CHECK(hipGetLastError());
  // the node generator of the accept/reject draw: index = volume + node number
  // (ks_imp_rhmc/setup.c:140)
  prn_init(g_prn.node, iseed, 2 * c.Vh);
  g_prn.single = single ? 1 : 0;
  g_prn.valid = true;
  MH_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

bool prn_valid() { return g_prn.valid && g_prn.nsites == (size_t)2 * ctx().Vh; }

int ranmom_prn_dev(double* mom_dev) {
  Context& c = ctx();
  if (!prn_valid()) return -3;
  hipLaunchKernelGGL(k_ranmom_prn, dim3((c.Vh + 255) / 256, 2), dim3(256), 0, c.stream, c.geom, g_prn.st,
                     g_prn.single, mom_dev);
  MH_CHECK(hipGetLastError());
  return 0;
}

int gaussian_prn_dev(int pb, VecField& v, double* sum) {
  Context& c = ctx();
  if (!prn_valid()) return -3;
  int nb = (c.Vh + 255) / 256;
  hipLaunchKernelGGL(k_gaussian_prn, dim3(nb), dim3(256), 0, c.stream, c.geom, g_prn.st, g_prn.single, pb, v,
                     c.ws.partials);
  MH_CHECK(hipGetLastError());
  std::vector<double> h((size_t)nb * 4);
  MH_CHECK(hipMemcpyAsync(h.data(), c.ws.partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MH_CHECK(hipStreamSynchronize(c.stream));
  double s = 0;
  for (int i = 0; i < nb; i++) s += h[(size_t)i * 4];
  if (sum) *sum = s;
  return 0;
}

// one uniform draw of the node generator (update(): the Metropolis test
// exp(-dS) < myrand(&node_prn), ks_imp_rhmc/update_rhmc.c)
double prn_node_uniform() { return prn_next(g_prn.node) * (1.0 / 16777216.0); }

void prn_release() {
  dev_free(g_prn.st);
  g_prn = PrnState{};
}

}  // namespace mh
