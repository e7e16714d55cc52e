// mh_dslash.h -- the fat + Naik staggered stencil for one output site.
//
//   dst(x) = sum_mu F_mu(x) s(x+mu) + L_mu(x) s(x+3mu)
//                 - F_mu(x-mu)^dag s(x-mu) - L_mu(x-3mu)^dag s(x-3mu)
// (generic_ks/dslash_fn_dblstore.c:311-563; the back links of the double
// store, fn_links_milc.c:115-195, are the adjoints of the neighbours' forward
// links, taken here in registers).
#pragma once
#include "mh_common.h"

namespace mh {

// Directions processed per scheduling window: the compiler otherwise hoists
// all sixteen link loads of the 16-bit and FP32 formats and runs out of
// registers; two directions keep 8 matrices in flight per thread.
template <int P> struct DslashUnroll { static constexpr int value = 4; };
template <> struct DslashUnroll<P_H16> { static constexpr int value = 2; };

// Link loads of the stencil: the 16-bit format is fixed at compile time (full
// nine elements for fat links; LR = 1: long links in the two-row form).
template <int P, int LR>
__device__ __forceinline__ void load_stencil_link(const LinkField& f, int pb, int mu, int cb,
                                                  Mat3<typename PrecTraits<P>::Real>& m) {
  if constexpr (P == P_H16) {
    if constexpr (LR) load_link_h16r(f, pb, mu, cb, m);
    else load_link_h16f(f, pb, mu, cb, m);
  } else {
    load_link<P>(f, pb, mu, cb, m);
  }
}

template <int P, int LR>
__device__ __forceinline__ void dslash_site_impl(const Geom& g, const LinkField& fat, const LinkField& lng,
                                                 const VecField& src, int cb, int pb,
                                                 Vec3<typename PrecTraits<P>::Real>& acc) {
  typedef typename PrecTraits<P>::Real R;
  Coord c = site_coords(g, cb, pb);
  int po = pb ^ 1;
  zero(acc);
#pragma unroll DslashUnroll<P>::value
  for (int mu = 0; mu < 4; mu++) {
    Mat3<R> m;
    Vec3<R> v;
    int n;
    // forward one hop
    n = neighbour(g, c, cb, mu, 1);
    load_stencil_link<P, 0>(fat, pb, mu, cb, m);
    load_vec<P>(src, n, v);
    mv_acc(m, v, acc);
    // forward three hops
    n = neighbour(g, c, cb, mu, 3);
    load_stencil_link<P, LR>(lng, pb, mu, cb, m);
    load_vec<P>(src, n, v);
    mv_acc(m, v, acc);
    // backward one hop: F_mu(x-mu) lives on the other parity
    n = neighbour(g, c, cb, mu, -1);
    load_stencil_link<P, 0>(fat, po, mu, n, m);
    load_vec<P>(src, n, v);
    amv_sub(m, v, acc);
    // backward three hops
    n = neighbour(g, c, cb, mu, -3);
    load_stencil_link<P, LR>(lng, po, mu, n, m);
    load_vec<P>(src, n, v);
    amv_sub(m, v, acc);
  }
}

// The 16-bit long-link format is uniform over the launch: one branch per thread.
template <int P>
__device__ __forceinline__ void dslash_site(const Geom& g, const LinkField& fat, const LinkField& lng,
                                            const VecField& src, int cb, int pb,
                                            Vec3<typename PrecTraits<P>::Real>& acc) {
  if constexpr (P == P_H16) {
    if (lng.recon) {
      dslash_site_impl<P, 1>(g, fat, lng, src, cb, pb, acc);
      return;
    }
  }
  dslash_site_impl<P, 0>(g, fat, lng, src, cb, pb, acc);
}

}  // namespace mh
