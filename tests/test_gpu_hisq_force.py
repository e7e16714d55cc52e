"""HISQ thin-link fermion force on the GPU through the C-ABI, against the FP64
oracle (which is pinned to MILC's compiled eo_fermion_force_multi and to
central differences of the action):

* milc_hip_eo_fermion_force_multi (MILC's CPU argument form: residues, eps,
  V and W rebuilt on the device from the thin links);
* milc_hip_hisq_force (qudaHisqForce form: per-term one-hop / three-hop
  weights 2 r_t and naik 2 r_t, W / V / U supplied);
* the momentum is added to, not overwritten.

Workload: 8^4 and 6x4x8x10 lattices, thin links from gauge seeds 1020/1021
(scale 0.4), three terms on all sites from seeds 300+, FP64, tolerance 1e-12.
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

C1 = [1 / 8, 0.0, -1 / 16, 1 / 64, -1 / 384, 0.0]
C2 = [1 / 8 + 12 / 16 + 1 / 8, -1 / 24, -1 / 16, 1 / 64, -1 / 384, -2 / 16]


def pack_ah(F):
    out = np.zeros(F.shape[:2] + (10,))
    out[..., 0:2] = F[..., 0, 1, :]
    out[..., 2:4] = F[..., 0, 2, :]
    out[..., 4:6] = F[..., 1, 2, :]
    out[..., 6] = F[..., 0, 0, 1]
    out[..., 7] = F[..., 1, 1, 1]
    out[..., 8] = F[..., 2, 2, 1]
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("dims,seed", [((8, 8, 8, 8), 1020), ((6, 4, 8, 10), 1021)])
def test_eo_fermion_force_multi_hisq_vs_oracle(dims, seed):
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(seed))
    xs = [L.gaussian_vector(300 + t, ko.EVENANDODD) for t in range(3)]
    res = [0.25, -0.8, 1.3]
    eps = 0.35
    ref = pack_ah(L.hisq_force(U, xs, res, eps=eps))
    with mq.Lattice(dims) as G:
        mom = np.zeros((L.V, 4, 10))
        G.eo_fermion_force_multi(eps, res, xs, U, mom)
        e = rel(mom, ref)
        print(f"HISQ force GPU vs oracle {dims}: {e:.2e}")
        assert e < 1e-12
        # accumulates into the momentum it is given
        G.eo_fermion_force_multi(eps, res, xs, U, mom)
        assert rel(mom, 2 * ref) < 1e-12


def test_hisq_force_quda_form_vs_oracle():
    dims = (8, 8, 8, 8)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1022))
    _, _, _, W = L.hisq_links(U, want_W=True)
    Vf = L.fat_links(C1, U)
    xs = [L.gaussian_vector(310 + t, ko.EVENANDODD) for t in range(2)]
    res = [0.6, 0.9]
    eps = 0.2
    ref = pack_ah(L.hisq_force(U, xs, res, eps=eps))
    coeff = [(2 * r, C2[1] * 2 * r) for r in res]
    with mq.Lattice(dims) as G:
        mom = np.zeros((L.V, 4, 10))
        G.hisq_force(eps, coeff, xs, C2, C1, W, Vf, U, mom)
    assert rel(mom, ref) < 1e-12
