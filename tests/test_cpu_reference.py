"""The oracle against MILC's own compiled CPU code (oracle/_ref, built by
oracle/Makefile.ref from the reference sources).  Skipped when the compiled
reference is not present.  MILC keeps its lattice in process globals, so every
test here uses the same 8^4 lattice."""
import numpy as np
import pytest

import ks_oracle as ko
import milc_ref as mr

pytestmark = pytest.mark.skipif(not mr.available(), reason="oracle/_ref/libmilc_ref.so not built")

DIMS = (8, 8, 8, 8)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def setup():
    L = ko.Lattice(DIMS)
    V = mr.init(DIMS)
    U = L.gauge_random(1000)
    mr.set_thin_links(U)
    f_ref, l_ref = mr.hisq_links(V)
    fat, lng, nsvd = L.hisq_links(L.rephase(U))
    return L, V, fat, lng, f_ref, l_ref


def test_hisq_links_match_milc(setup):
    L, V, fat, lng, f_ref, l_ref = setup
    assert rel(fat, f_ref) < 1e-14
    assert rel(lng, l_ref) < 1e-14


@pytest.mark.parametrize("parity", [ko.EVEN, ko.ODD])
def test_dslash_matches_milc(setup, parity):
    L, V, fat, lng, f_ref, l_ref = setup
    mr.set_fn_links(fat, lng)
    src = L.gaussian_vector(5, ko.EVENANDODD)
    s = L.parity_slice(parity)
    assert rel(L.dslash(fat, lng, src, parity)[s], mr.dslash(src, parity)[s]) < 1e-14


def test_cg_matches_milc_iteration_for_iteration(setup):
    L, V, fat, lng, f_ref, l_ref = setup
    mr.set_fn_links(fat, lng)
    b = L.gaussian_vector(7, ko.EVEN)
    x1, it1, q1 = mr.congrad(b, 0.05, 1e-10, ko.EVEN, 40, 10)  # restarts every 40 iterations
    x2, it2, q2 = L.congrad(fat, lng, b, 0.05, 1e-10, parity=ko.EVEN, max_iters=40, nrestart=10)
    s = L.parity_slice(ko.EVEN)
    assert it1 == it2
    assert q1["restarts"] == q2.final_restart
    assert abs(q1["final_rsq"] - q2.final_rsq) < 1e-6 * q1["final_rsq"]
    assert rel(x2[s], x1[s]) < 1e-13


def test_multicg_matches_milc(setup):
    L, V, fat, lng, f_ref, l_ref = setup
    mr.set_fn_links(fat, lng)
    b = L.gaussian_vector(8, ko.EVEN)
    offs = [4 * 0.05 ** 2 + x for x in (0.0, 0.01, 0.1, 1.0)]
    p1, i1, r1 = mr.multicg(b, offs, 1e-10, ko.EVEN, 2000)
    p2, i2, q = L.multicg_offset(fat, lng, b, offs, 1e-10, parity=ko.EVEN, max_iters=2000, nrestart=1)
    s = L.parity_slice(ko.EVEN)
    assert i1 == i2
    for j in range(len(offs)):
        assert rel(p2[j][s], p1[j][s]) < 1e-13


def test_force_oprod_matches_milc_smearing0(setup):
    L, V, fat, lng, f_ref, l_ref = setup
    xs = [L.gaussian_vector(100 + t, ko.EVENANDODD) for t in range(3)]
    res = [0.3, -1.2, 0.7]
    o1a, o3a = mr.force_oprod(xs, res, V)
    o1b, o3b = L.force_oprod(xs, res)
    assert rel(o1b, o1a) < 1e-15 and rel(o3b, o3a) < 1e-15


def _pack_ah(F):
    out = np.zeros(F.shape[:2] + (10,))
    out[..., 0:2] = F[..., 0, 1, :]
    out[..., 2:4] = F[..., 0, 2, :]
    out[..., 4:6] = F[..., 1, 2, :]
    out[..., 6] = F[..., 0, 0, 1]
    out[..., 7] = F[..., 1, 1, 1]
    out[..., 8] = F[..., 2, 2, 1]
    return out


def test_hisq_force_matches_milc(setup):
    """The full HISQ thin-link force (eo_fermion_force_multi, FNMAT wrapper):
    level 0, level-2 smearing on W, U(3) reunitarisation derivative, fat7
    smearing on U, TA(2 eps U F) -- the oracle against MILC's own code."""
    L, V, fat, lng, f_ref, l_ref = setup
    mr.hisq_links(V)  # MILC's own V, W, U auxiliary links for the force
    U = L.gauge_random(1000)
    xs = [L.gaussian_vector(200 + t, ko.EVENANDODD) for t in range(3)]
    res = [0.25, -0.8, 1.3]
    eps = 0.35
    m_milc = mr.hisq_force(xs, res, eps, V)
    m_orc = _pack_ah(L.hisq_force(L.rephase(U), xs, res, eps=eps))
    e = rel(m_orc, m_milc)
    print(f"HISQ force oracle vs MILC: {e:.2e}")
    assert e < 1e-12


def test_second_naik_set_links_and_force_match_milc(setup):
    """A second Naik set (n_naiks = 2, eps_naik = {0, -0.2}: a charm quark of a
    2+1+1 run): MILC's set-1 links (X_2 + eps X_3) and its force with the last
    two of four terms in that set, against the oracle"""
    L, V, fat, lng, f_ref, l_ref = setup
    U = L.gauge_random(1000)
    epsn = -0.2
    f0, l0, f1, l1 = mr.hisq_links_naik(V, epsn)
    assert rel(f0, f_ref) < 1e-14 and rel(l0, l_ref) < 1e-14
    fo, lo = L.hisq_links_naik(L.rephase(U), epsn)
    assert rel(fo, f1) < 1e-14 and rel(lo, l1) < 1e-14
    assert rel(lo, (1 + epsn) * l_ref) < 1e-14
    xs = [L.gaussian_vector(300 + t, ko.EVENANDODD) for t in range(4)]
    res = [0.25, -0.8, 1.3, 0.6]
    eps = 0.35
    m_milc = mr.hisq_force(xs, res, eps, V, n_naik=2)
    m_orc = _pack_ah(L.hisq_force(L.rephase(U), xs, res, eps=eps, n_naik=2, eps_naik=epsn))
    e = rel(m_orc, m_milc)
    print(f"HISQ force with a second Naik set, oracle vs MILC: {e:.2e}")
    assert e < 1e-12
    # the second set changes the force: the eps_naik term is really there
    m_plain = _pack_ah(L.hisq_force(L.rephase(U), xs, res, eps=eps))
    assert rel(m_plain, m_milc) > 1e-3
