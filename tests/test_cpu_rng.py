"""Oracle of MILC's per-site random numbers (oracle/ks_oracle.c: kso_prn_*,
kso_ranmom_prn, kso_gaussian_prn) and of reunit_su3 (kso_reunit_su3), against
MILC's own printed numbers and the mathematics.

MILC's regression run ks_imp_rhmc/test/su3_rhmc_hisq.1 (6^4, iseed 86658, a
MILC_PRECISION=1 build) draws the momenta (ranmom) and then one Gaussian
source per pseudofermion (grsource_imp_rhmc) from the site generators, and
prints
  GRSOURCE: sum = 1.9636750663e+03 / 1.9027332576e+03 / 1.9328576741e+03 /
                  1.9363357361e+03                (sample-out lines 526-532)
  ACTION: g,h,f = ... -8.03776340931654e+01 ...    (line 539, h = hmom_action)
The oracle's draws, summed with MILC's single-precision accumulation
(ahmat_mag_sq of d_action_rhmc.c:112-124, the FAST magsq_su3vec of
libraries/msq_su3vec.c:38-48), reproduce every printed digit: the momenta and
sources are MILC's, draw for draw.
"""
import numpy as np
import pytest

import ks_oracle as ko

ISEED = 86658
GRSOURCE = [1.9636750663e+03, 1.9027332576e+03, 1.9328576741e+03, 1.9363357361e+03]
H0 = -8.03776340931654e+01
f32 = np.float32


def milc_hmom_action_single(mom):
    """hmom_action with Real = float: ahmat_mag_sq rounded as MILC's code does"""
    p = mom.reshape(-1, 10).astype(np.float64)
    tot = 0.0
    for row in p:
        x = row[6]
        s = f32(0.5 * x * x)
        for k in (7, 8):
            s = f32(float(s) + 0.5 * row[k] * row[k])
        for k in range(6):
            s = f32(s + f32(f32(row[k]) * f32(row[k])))
        tot += float(s) - 4.0
    return tot


def milc_grsource_sum_single(v, nsites):
    """sum over sites of (double) magsq_su3vec, float accumulation per site"""
    a = v.reshape(-1, 6)[:nsites].astype(np.float32)
    s = np.zeros(nsites, dtype=np.float32)
    for k in range(6):
        s = (s + a[:, k] * a[:, k]).astype(np.float32)
    return float(np.sum(s.astype(np.float64)))


@pytest.fixture(scope="module")
def draws():
    L = ko.Lattice((6, 6, 6, 6))
    st = L.prn_seed(ISEED)
    mom = L.ranmom_prn(st, single=True)
    src = [L.gaussian_prn(st, ko.EVEN, single=True) for _ in range(4)]
    return L, mom, src


def test_momenta_reproduce_milc_hmom_action(draws):
    L, mom, _ = draws
    assert milc_hmom_action_single(mom) == pytest.approx(H0, rel=1e-14)
    # in FP64 the same momenta give it to MILC's float rounding
    h = float(np.sum(0.5 * np.sum(mom[..., 6:9] ** 2, -1) + np.sum(mom[..., 0:6] ** 2, -1) - 4.0))
    assert abs(h - H0) < 1e-4
    # every value is a float32 number (the single-precision build's rounding)
    assert np.array_equal(mom.astype(np.float32).astype(np.float64), mom)


def test_sources_reproduce_milc_grsource_sums(draws):
    L, _, src = draws
    E = L.parity_slice(ko.EVEN)
    for (v, s), ref in zip(src, GRSOURCE):
        assert milc_grsource_sum_single(v, L.Vh) == pytest.approx(ref, rel=3e-11)
        assert s == pytest.approx(ref, rel=1e-8)
        assert not v[L.Vh:].any()            # odd sites untouched
        assert float(np.sum(v[E] ** 2)) == pytest.approx(s, rel=1e-13)


def test_generator_streams_are_per_site():
    """each site's generator depends on its lexicographic index only: the
    draws at a site equal a stand-alone generator seeded with (iseed, index)"""
    L = ko.Lattice((4, 4, 4, 4))
    st = L.prn_seed(1234)
    lex = L.lex_of_site()
    for i in (0, 5, L.Vh + 3, L.V - 1):
        ref = ko.prn_stream(1234, int(lex[i]), 20)
        one = st[i].copy()
        got = np.array([ko.lib().kso_prn_next(one.ctypes.data_as(ko.ctypes.c_void_p)) for _ in range(20)])
        assert np.array_equal(got, ref)
    # 24-bit values, no obvious bias
    x = ko.prn_stream(7, 11, 200000).astype(np.float64) / 2 ** 24
    assert x.max() < 1.0 and x.min() >= 0.0
    assert abs(x.mean() - 0.5) < 3e-3 and abs(x.var() - 1 / 12) < 2e-3


@pytest.mark.parametrize("single", [True, False])
def test_gaussian_and_momentum_statistics(single):
    """weight exp(-x^2): <x^2> = 1/2 per real part; momenta with weight
    exp(-|P|^2/2): <|m01|^2> = 1, <m00im^2> = 2/3, traceless"""
    L = ko.Lattice((8, 8, 8, 8))
    st = L.prn_seed(99)
    mom = L.ranmom_prn(st, single=single)
    v, s = L.gaussian_prn(st, ko.EVENANDODD, single=single)
    assert np.mean(v ** 2) == pytest.approx(0.5, abs=0.01)
    assert np.mean(mom[..., 0:6] ** 2) == pytest.approx(0.5, abs=0.01)
    assert np.mean(mom[..., 6] ** 2) == pytest.approx(2 / 3, abs=0.02)
    assert np.abs(np.sum(mom[..., 6:9], -1)).max() < 1e-6
    # the single-precision draws are the double ones rounded (same stream)
    st2 = L.prn_seed(99)
    other = L.ranmom_prn(st2, single=not single)
    assert np.abs(other - mom).max() < 1e-5


def test_reunit_su3_gram_schmidt():
    """rows 0, 1 = Gram-Schmidt of the input rows (numpy QR of U^T, phases
    fixed), row 2 = conj(row0 x row1): det 1, unitary; SU(3) input unchanged;
    the deviations are the ones MILC's check_deviation counts"""
    rng = np.random.default_rng(3)
    L = ko.Lattice((2, 2, 2, 2))
    U = L.gauge_random(77, 0.5)
    Uc = U[..., 0] + 1j * U[..., 1]
    W = U.copy()
    n, md = L.reunitarize(W)
    assert n == 0 and md < 1e-13
    assert np.abs(W - U).max() < 1e-14
    noise = 1e-3 * (rng.standard_normal(U.shape))
    P = np.ascontiguousarray(U + noise)
    Pc = P[..., 0] + 1j * P[..., 1]
    R = P.copy()
    n, md = L.reunitarize(R, tol=1e-4)
    Rc = R[..., 0] + 1j * R[..., 1]
    for idx in np.ndindex(Pc.shape[:2]):
        q, r = np.linalg.qr(Pc[idx].T)
        ph = np.diag(r) / np.abs(np.diag(r))
        q = q * ph[None, :]
        np.testing.assert_allclose(Rc[idx][0], q[:, 0], atol=1e-14)
        np.testing.assert_allclose(Rc[idx][1], q[:, 1], atol=1e-14)
        np.testing.assert_allclose(np.linalg.det(Rc[idx]), 1.0, atol=1e-14)
        np.testing.assert_allclose(Rc[idx] @ Rc[idx].conj().T, np.eye(3), atol=1e-14)
    # deviations of size ~1e-3 on every link: some reach the 1e-4 tolerance
    assert 1e-5 < md < 1e-1 and n > 0
    n2, md2 = L.reunitarize(np.ascontiguousarray(P.copy()), tol=1.0)
    assert n2 == 0 and md2 == pytest.approx(md)
    assert np.abs(Rc - Pc).max() < 1e-2
