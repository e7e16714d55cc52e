"""FN (fat + Naik) thin-link fermion force on the GPU through the C-ABI:
milc_hip_fermion_force_fn_multi against MILC's golden momentum and the FP64
oracle, milc_hip_asqtad_force against it from the outer products, and the
accumulate-into-mom contract.

Workloads: MILC's 8^4 sample lattice (asqtad, u0 = 0.502, eps = 1, r = 1,
X = ksvector.src.sample.l8) and a random 8^4 configuration (seed 1003, two
terms, eps = 0.37).
"""
import os

import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

GDIR = os.path.join(os.path.dirname(__file__), "golden")


def golden(name, lex):
    return np.load(os.path.join(GDIR, name)).astype(np.float64)[lex].copy()


def pack_ah(F):
    """full anti-Hermitian matrices (V, 4, 3, 3, 2) -> MILC anti_hermitmat (V, 4, 10)"""
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


@pytest.fixture(scope="module")
def l8():
    L = ko.Lattice((8, 8, 8, 8))
    lex = L.lex_of_site()
    U = L.rephase(golden("lat_l8.npy", lex))
    src = golden("ksvector_src_l8.npy", lex)
    return L, U, src, lex


def test_fn_force_golden_and_oracle(l8):
    L, U, src, lex = l8
    c = ko.asqtad_coeffs(0.502)
    ref = pack_ah(L.fn_force(c, U, [src], [1.0], eps=1.0))
    gold = pack_ah(golden("colormatrix_mom_l8.npy", lex))
    with mq.Lattice(L.dims) as G:
        mom = np.zeros((L.V, 4, 10))
        G.fermion_force_fn_multi(1.0, [1.0], [src], c, U, mom)
    print(f"GPU FN force: rel vs oracle {rel(mom, ref):.2e}, vs MILC golden {rel(mom, gold):.2e}")
    assert rel(mom, ref) < 1e-12
    assert rel(mom, gold) < 1e-6


def test_fn_force_random_two_terms_accumulates(l8):
    L = l8[0]
    U = L.rephase(L.gauge_random(1003))
    c = ko.asqtad_coeffs(0.85)
    xs = [L.gaussian_vector(510 + t, ko.EVENANDODD) for t in range(2)]
    res = [0.61, -0.23]
    eps = 0.37
    inc = pack_ah(L.fn_force(c, U, xs, res, eps=eps))
    mom0 = pack_ah(L.fn_force(c, U, [L.gaussian_vector(99, ko.EVENANDODD)], [1.0], eps=1.0))
    mom0[..., 9] = 0.0
    with mq.Lattice(L.dims) as G:
        mom = mom0.copy()
        G.fermion_force_fn_multi(eps, res, xs, c, U, mom)
    assert rel(mom - mom0, inc) < 1e-12
    assert np.all(mom[..., 9] == 0.0)


def test_asqtad_force_from_outer_products(l8):
    """qudaAsqtadForce form: outer products carry 2 eps * weight
    (fermion_force_asqtad_gpu.c:53-60) and the path chain adds TA(U G)."""
    L, U, src, lex = l8
    c = ko.asqtad_coeffs(0.502)
    b = L.gaussian_vector(612, ko.EVENANDODD)
    eps, w1, w2 = 0.05, 0.8, 1.3
    ref = pack_ah(L.fn_force(c, U, [src, b], [w1, w2], eps=eps))
    with mq.Lattice(L.dims) as G:
        O1, O3 = G.compute_oprod([src, b], [(2 * eps * w1, 2 * eps * w1), (2 * eps * w2, 2 * eps * w2)])
        mom = np.zeros((L.V, 4, 10))
        G.asqtad_force(c, O1, O3, U, mom)
    assert rel(mom, ref) < 1e-12
