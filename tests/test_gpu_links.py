"""The 16-bit long-link format through the C-ABI: HISQ Naik links (a real
multiple of a U(3) matrix) go to the two-row form and the 16-bit stencil built
on it matches the FP64 oracle to the 16-bit tolerance, like the nine-element
form; long links that are not scaled-unitary keep the nine-element form.

Workload: 6x4x8x10 lattice (non-cubic), HISQ links from gauge seed 1030,
Gaussian sources from seeds 1400+.  Tolerances: FP64 1e-14, FP32 1e-6,
16-bit 5e-4 relative (16-bit fixed-point links and vectors).
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

DIMS = (6, 4, 8, 10)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def hisq():
    L = ko.Lattice(DIMS)
    U = L.rephase(L.gauge_random(1030))
    fat, lng, _ = L.hisq_links(U)
    return L, fat, lng


def dslash_all(G, L, fat, lng, prec, seed):
    """(GPU, oracle) Dslash outputs for both output parities"""
    out = []
    for out_par, in_par in ((ko.EVEN, ko.ODD), (ko.ODD, ko.EVEN)):
        v = L.gaussian_vector(seed, in_par)
        ref = L.dslash(fat, lng, v, out_par)
        G.dev_write(0, in_par, v)
        G.dev_dslash(prec, 0, out_par, 1)
        sl = L.parity_slice(out_par)
        out.append((G.dev_read(1, out_par)[sl], ref[sl]))
    return out


def test_hisq_long_links_two_row_form(hisq):
    L, fat, lng = hisq
    with mq.Lattice(DIMS) as G:
        G.dev_load_links(fat, lng)
        assert G.link_compression(2) == 1
        assert G.link_compression(1) == 0 and G.link_compression(0) == 0
        for prec, tol in ((0, 1e-14), (1, 1e-6), (2, 5e-4)):
            for got, ref in dslash_all(G, L, fat, lng, prec, 1400):
                e = rel(got, ref)
                print(f"prec {prec}: rel {e:.2e}")
                assert e < tol, (prec, e)
        two_row = dslash_all(G, L, fat, lng, 2, 1401)
        G.set_link_compression(0)
        assert G.link_compression(2) == 0
        nine = dslash_all(G, L, fat, lng, 2, 1401)
        G.set_link_compression(1)
        assert G.link_compression(2) == 1
        for (a, ref), (b, _) in zip(two_row, nine):
            assert rel(b, ref) < 5e-4
            # the two forms differ by the quantisation of the rebuilt row only
            assert rel(a, b) < 5e-4


def test_non_unitary_long_links_keep_nine_elements(hisq):
    L, fat, lng = hisq
    lng2 = lng * (1.0 + 0.01 * np.random.default_rng(7).standard_normal(lng.shape))
    with mq.Lattice(DIMS) as G:
        G.dev_load_links(fat, lng2)
        assert G.link_compression(2) == 0
        for got, ref in dslash_all(G, L, fat, lng2, 2, 1402):
            assert rel(got, ref) < 5e-4


def test_mixed_cg_on_two_row_links_reaches_target(hisq):
    """the mixed 16-bit CG on two-row long links: FP64 true residual 1e-10 and
    the solution within 1e-8 of the oracle's"""
    L, fat, lng = hisq
    m = 0.05
    src = L.gaussian_vector(1403, ko.EVEN)
    ref, _, _ = L.congrad(fat, lng, src, m, 1e-14, parity=ko.EVEN, max_iters=5000, nrestart=5)
    sl = L.parity_slice(ko.EVEN)
    with mq.Lattice(DIMS) as G:
        G.dev_load_links(fat, lng)
        assert G.link_compression(2) == 1
        G.dev_write(0, ko.EVEN, src)
        it, st = G.dev_invert(0, 1, m, mq.EVEN, 1e-10, 2)
        print(it, st)
        x = G.dev_read(1, ko.EVEN)
    assert st["converged"] == 1 and st["resid"] < 1e-10
    assert rel(x[sl], ref[sl]) < 1e-8
