"""HISQ thin-link fermion force on the GPU through the C-ABI, against the FP64
oracle (which is pinned to MILC's compiled eo_fermion_force_multi and to
central differences of the action):

* milc_hip_eo_fermion_force_multi (MILC's CPU argument form: residues, eps,
  V and W rebuilt on the device from the thin links);
* milc_hip_hisq_force (qudaHisqForce form: per-term one-hop / three-hop
  weights 2 r_t and naik 2 r_t, W / V / U supplied);
* the momentum is added to, not overwritten;
* a second Naik set (a charm quark of a 2+1+1 run, eps_naik = -0.2): its
  links from milc_hip_load_hisq_links_naik and the force with the last two of
  four terms in that set (oracle pinned to MILC's compiled code with
  n_naiks = 2 and to central differences on those links).

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


def test_dev_hisq_force_from_resident_solutions():
    """milc_hip_dev_hisq_force: X_t on even sites in device slots, links built
    on the device from the seed; the oracle gets the full Z_t = (X_t, D_oe X_t)
    and the same thin links"""
    dims = (6, 4, 8, 10)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1025))
    fat, lng, _ = L.hisq_links(U)
    res = [0.4, 1.2, -0.3]
    eps = 0.45
    E = L.parity_slice(ko.EVEN)
    zs = []
    for t in range(len(res)):
        x = L.gaussian_vector(340 + t, ko.EVEN)
        zs.append(x + L.dslash(fat, lng, x, ko.ODD))
        assert np.array_equal(zs[-1][E], x[E])
    ref = pack_ah(L.hisq_force(U, zs, res, eps=eps))
    with mq.Lattice(dims) as G:
        G.dev_gauge_hisq(1025)
        for t in range(len(res)):
            G.dev_source(340 + t, mq.EVEN, 5 + t)
        mom = np.zeros((L.V, 4, 10))
        st = G.dev_hisq_force(5, res, eps, mom)
        print(st)
        assert rel(mom, ref) < 1e-12
        # |increment|^2 as reported; the packed form holds each off-diagonal once
        assert st["mom_norm2"] > 0 and abs(st["mom_norm2"] - float(np.sum(mom ** 2))) < 1e-10 * st["mom_norm2"]


@pytest.mark.parametrize("form", [0, 1, 2])
def test_every_device_form_vs_oracle(form):
    """the path-table gather (0), the default choice (1: staple form where it is
    exact for any links) and the staple form everywhere (2) each match the
    oracle's path table, HISQ and asqtad FN, on unitary links"""
    dims = (6, 4, 8, 10)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1023))
    xs = [L.gaussian_vector(320 + t, ko.EVENANDODD) for t in range(2)]
    res = [0.7, -0.4]
    eps = 0.3
    ref_h = pack_ah(L.hisq_force(U, xs, res, eps=eps))
    c = ko.asqtad_coeffs(0.85)
    ref_f = pack_ah(L.fn_force(c, U, xs, res, eps=eps))
    with mq.Lattice(dims) as G:
        G.set_force_form(form)
        try:
            mom = np.zeros((L.V, 4, 10))
            G.eo_fermion_force_multi(eps, res, xs, U, mom)
            assert rel(mom, ref_h) < 1e-12
            mom = np.zeros((L.V, 4, 10))
            G.fermion_force_fn_multi(eps, res, xs, c, U, mom)
            assert rel(mom, ref_f) < 1e-12
        finally:
            G.set_force_form(1)


@pytest.mark.parametrize("form", [0, 1])
def test_non_unitary_links_exact_by_default(form):
    """links 3% off U(3): the forms that use the staple form only where it is
    exact (0, 1) still reproduce the path-table definition of the force"""
    dims = (4, 4, 4, 6)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1024))
    U = U * (1.0 + 0.03 * np.random.default_rng(5).standard_normal(U.shape))
    xs = [L.gaussian_vector(330, ko.EVENANDODD)]
    c = ko.asqtad_coeffs(0.4)
    ref_f = pack_ah(L.fn_force(c, U, xs, [1.1], eps=0.5))
    with mq.Lattice(dims) as G:
        G.set_force_form(form)
        try:
            mom = np.zeros((L.V, 4, 10))
            G.fermion_force_fn_multi(0.5, [1.1], xs, c, U, mom)
            assert rel(mom, ref_f) < 1e-12
        finally:
            G.set_force_form(1)


EPS_NAIK = -0.2


def test_load_hisq_links_second_naik_set_vs_oracle():
    dims = (8, 8, 8, 8)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1023))
    fo, lo = L.hisq_links_naik(U, EPS_NAIK)
    with mq.Lattice(dims) as G:
        gf, gl, gW, _ = G.load_hisq_links(U, eps_naik=EPS_NAIK)
        gf0, gl0, _, _ = G.load_hisq_links(U)
    assert rel(gf, fo) < 1e-13 and rel(gl, lo) < 1e-13
    f0, l0, _ = L.hisq_links(U)
    assert rel(gf0, f0) < 1e-13 and rel(gl0, l0) < 1e-13
    assert rel(gf, f0) > 1e-3


def test_hisq_force_second_naik_set_vs_oracle():
    dims = (8, 8, 8, 8)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1024))
    _, _, _, W = L.hisq_links(U, want_W=True)
    Vf = L.fat_links(C1, U)
    xs = [L.gaussian_vector(330 + t, ko.EVENANDODD) for t in range(4)]
    res = [0.6, 0.9, -0.4, 1.1]
    eps = 0.2
    ref = pack_ah(L.hisq_force(U, xs, res, eps=eps, n_naik=2, eps_naik=EPS_NAIK))
    coeff = [(2 * r, C2[1] * 2 * r) for r in res]
    coeff += [(EPS_NAIK * 2 * r / 8, -EPS_NAIK * 2 * r / 24) for r in res[2:]]
    with mq.Lattice(dims) as G:
        mom = np.zeros((L.V, 4, 10))
        G.hisq_force(eps, coeff, xs, C2, C1, W, Vf, U, mom, num_naik_terms=2)
    e = rel(mom, ref)
    print(f"second Naik set force vs oracle: {e:.2e}")
    assert e < 1e-12
    plain = pack_ah(L.hisq_force(U, xs, res, eps=eps))
    assert rel(plain, ref) > 1e-3
