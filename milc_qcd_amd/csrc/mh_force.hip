// mh_force.hip -- level-0 fermion-force outer products on the device.
//
// Restates generic_ks/fermion_force_hisq_multi.c:1478-1636
// (fn_fermion_force_multi_hisq_smearing0) and the outer-product stage of
// fermion_force_asqtad_gpu.c:22-69:
//   O1_mu(x) = sum_t c1_t Z_t(x+mu)  Z_t(x)^dag
//   O3_mu(x) = sum_t c3_t Z_t(x+3mu) Z_t(x)^dag
// over all sites, with Z_t given on both parities.
#include <vector>

#include "mh_internal.h"

namespace mh {

#define OP_MAX 32
struct TermSet {
  VecField even[OP_MAX];
  VecField odd[OP_MAX];
  double c1[OP_MAX];
  double c3[OP_MAX];
};

// Four lanes per site, one per direction mu (lane = 4 * site + mu), 64 sites
// of one parity per block; the blocks of a site range are dealt to one XCD
// with both parities side by side (xcd_block), so the here-vector Z_t(x) is
// one request for the four lanes and the x/y neighbours come from that XCD's
// L2.  Each lane keeps its O1_mu and O3_mu (two FP64 3x3 accumulators) over
// all terms and writes them once, non-temporally (streamed out instead of
// displacing from the L2 the vectors that the other lanes are still gathering).
__global__ void __launch_bounds__(256) k_oprod(Geom g, TermSet ts, int nterms, LinkField o1, LinkField o3) {
  const int nbx = (g.Vh + 63) / 64;
  int bx, pb;
  if (!xcd_block(nbx, 2, bx, pb)) return;
  const int mu = threadIdx.x & 3;
  const int cb = bx * 64 + (threadIdx.x >> 2);
  if (cb >= g.Vh) return;
  Coord c = site_coords(g, cb, pb);
  int n1 = neighbour(g, c, cb, mu, 1);
  int n3 = neighbour(g, c, cb, mu, 3);
  Mat3<double> a1, a3;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) a1.e[i][j].re = a1.e[i][j].im = a3.e[i][j].re = a3.e[i][j].im = 0.0;
  for (int t = 0; t < nterms; t++) {
    const VecField& here = pb ? ts.odd[t] : ts.even[t];
    const VecField& there = pb ? ts.even[t] : ts.odd[t];
    Vec3<double> z, y1, y3;
    load_vec<P_F64>(here, cb, z);
    load_vec<P_F64>(there, n1, y1);
    load_vec<P_F64>(there, n3, y3);
    double w1 = ts.c1[t], w3 = ts.c3[t];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        // y_i conj(z_j)  (libraries/su3_proj.c)
        double re1 = y1.c[i].re * z.c[j].re + y1.c[i].im * z.c[j].im;
        double im1 = y1.c[i].im * z.c[j].re - y1.c[i].re * z.c[j].im;
        double re3 = y3.c[i].re * z.c[j].re + y3.c[i].im * z.c[j].im;
        double im3 = y3.c[i].im * z.c[j].re - y3.c[i].re * z.c[j].im;
        a1.e[i][j].re += w1 * re1;
        a1.e[i][j].im += w1 * im1;
        a3.e[i][j].re += w3 * re3;
        a3.e[i][j].im += w3 * im3;
      }
  }
  const int pm = pb * 4 + mu;
  double2* p1 = (double2*)o1.data + (size_t)(pm * 9) * o1.stride + cb;
  double2* p3 = (double2*)o3.data + (size_t)(pm * 9) * o3.stride + cb;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    nat_d2 v1, v3;
    v1.x = a1.e[e / 3][e % 3].re;
    v1.y = a1.e[e / 3][e % 3].im;
    v3.x = a3.e[e / 3][e % 3].re;
    v3.y = a3.e[e / 3][e % 3].im;
    __builtin_nontemporal_store(v1, (nat_d2*)(p1 + (size_t)e * o1.stride));
    __builtin_nontemporal_store(v3, (nat_d2*)(p3 + (size_t)e * o3.stride));
  }
}

int force_oprod_coeff(const std::vector<VecField>& xe, const std::vector<VecField>& xo, const std::vector<double>& c1,
                      const std::vector<double>& c3, LinkField& o1, LinkField& o3) {
  Context& c = ctx();
  int n = (int)xe.size();
  if (n <= 0 || n > OP_MAX || (int)xo.size() != n) return -2;
  TermSet ts{};
  for (int t = 0; t < n; t++) {
    ts.even[t] = xe[t];
    ts.odd[t] = xo[t];
    ts.c1[t] = c1[t];
    ts.c3[t] = c3[t];
  }
  hipLaunchKernelGGL(k_oprod, xcd_grid((c.Vh + 63) / 64, 2), dim3(256), 0, c.stream, c.geom, ts, n, o1, o3);
  MH_CHECK(hipGetLastError());
  return 0;
}

int force_oprod(const std::vector<VecField>& xe, const std::vector<VecField>& xo, const std::vector<double>& residues,
                LinkField& o1, LinkField& o3) {
  return force_oprod_coeff(xe, xo, residues, residues, o1, o3);
}

}  // namespace mh
