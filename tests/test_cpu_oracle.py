"""Pins of the FP64 oracle (oracle/ks_oracle.c), all on the CPU.

Each function of the oracle is checked against something that does not share
its code: MILC's own golden files (tests/golden, from binary_samples/ of the
reference), closed forms, symmetries, or an independent numpy computation.
"""
import os

import numpy as np
import pytest

import ks_oracle as ko

GDIR = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def golden(name, lex):
    return np.load(os.path.join(GDIR, name)).astype(np.float64)[lex].copy()


@pytest.fixture(scope="module")
def milc_l8():
    L = ko.Lattice((8, 8, 8, 8))
    lex = L.lex_of_site()
    U = golden("lat_l8.npy", lex)
    return L, U, lex


# --------------------------------------------------------------------------
# MILC golden data (ks_imp_utilities/test/check_ks_invert, asqtad, u0 = 0.502)
# --------------------------------------------------------------------------
def test_golden_asqtad_fat_links(milc_l8):
    L, U, lex = milc_l8
    fat, _ = L.asqtad_links(L.rephase(U), 0.502)
    ref = golden("fat_l8.npy", lex)
    # the golden file is FP32: agreement to its rounding
    assert np.abs(fat - ref).max() < 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("parity", [ko.EVEN, ko.ODD])
def test_golden_mdaggerm_inverse(milc_l8, parity):
    L, U, lex = milc_l8
    fat = golden("fat_l8.npy", lex)
    _, lng = L.asqtad_links(L.rephase(U), 0.502)
    src = golden("ksvector_src_l8.npy", lex)
    ans = golden("ksvector_ans_MdaggerMinv_l8.npy", lex)
    x, it, q = L.congrad(fat, lng, src, 0.1, 1e-9, parity=parity, max_iters=300, nrestart=5)
    s = L.parity_slice(parity)
    assert q.converged and q.final_rsq < 1e-18
    assert 55 <= it <= 70  # MILC reports 61 iterations, 1 restart (check_ks_invert.2.sample-out:82)
    assert rel(x[s], ans[s]) < 1e-6


def test_golden_m_inverse(milc_l8):
    """M^-1 src via the even-odd reduction of mat_invert.c:328-407:
    x_e = (M^dag M)^-1 (2m b_e - D_eo b_o), x_o = (b_o - D_oe x_e) / 2m."""
    L, U, lex = milc_l8
    m = 0.1
    fat = golden("fat_l8.npy", lex)
    _, lng = L.asqtad_links(L.rephase(U), 0.502)
    b = golden("ksvector_src_l8.npy", lex)
    ans = golden("ksvector_ans_Minv_l8.npy", lex)
    E, O = L.parity_slice(ko.EVEN), L.parity_slice(ko.ODD)
    t = L.dslash(fat, lng, b, ko.EVEN)
    rhs = np.zeros_like(b)
    rhs[E] = 2 * m * b[E] - t[E]
    x, it, q = L.congrad(fat, lng, rhs, m, 1e-10, parity=ko.EVEN, max_iters=300, nrestart=5)
    y = L.dslash(fat, lng, x, ko.ODD)
    x[O] = (b[O] - y[O]) / (2 * m)
    assert rel(x, ans) < 1e-6


def test_golden_fn_force(milc_l8):
    """fermion_force_fn_multi with the asqtad path table (u0 = 0.502), eps = 1,
    one term r = 1, X = ksvector.src: MILC's colormatrix.mom.sample.l8
    (ks_imp_utilities/test/check_fermion_force; FP32 file)."""
    L, U, lex = milc_l8
    src = golden("ksvector_src_l8.npy", lex)
    ans = golden("colormatrix_mom_l8.npy", lex)
    mom = L.fn_force(ko.asqtad_coeffs(0.502), L.rephase(U), [src], [1.0], eps=1.0)
    assert rel(mom, ans) < 1e-6


def test_fn_force_linear_in_residues_and_eps(milc_l8):
    """the force is linear in eps and in each term's residue"""
    L, U, lex = milc_l8
    Up = L.rephase(U)
    c = ko.asqtad_coeffs(0.502)
    a = golden("ksvector_src_l8.npy", lex)
    b = L.gaussian_vector(77, ko.EVENANDODD)
    fa = L.fn_force(c, Up, [a], [1.0], eps=1.0)
    fb = L.fn_force(c, Up, [b], [1.0], eps=1.0)
    fab = L.fn_force(c, Up, [a, b], [0.3, -1.7], eps=0.25)
    assert rel(fab, 0.25 * (0.3 * fa - 1.7 * fb)) < 1e-13


# --------------------------------------------------------------------------
# closed forms and symmetries
# --------------------------------------------------------------------------
def free_links(L, c1, c3):
    """fat = c1 eta_mu(x); Naik = c3 eta_mu(x) eta_mu(x+mu) eta_mu(x+2mu) (the
    antiperiodic sign is picked up by every 3-hop that crosses t = nt-1)."""
    one = np.zeros((L.V, 4, 3, 3, 2))
    for a in range(3):
        one[:, :, a, a, 0] = 1.0
    ph = L.rephase(one)
    eta = ph[:, :, 0, 0, 0]  # (V, 4)
    nbr = np.ctypeslib.as_array(L._g.nbr, shape=(16, L.V))
    eta3 = np.empty_like(eta)
    for mu in range(4):
        x1 = nbr[mu]
        x2 = nbr[mu][x1]
        eta3[:, mu] = eta[:, mu] * eta[x1, mu] * eta[x2, mu]
    return c1 * ph, c3 * one * eta3[:, :, None, None, None]


def test_hisq_free_field_coefficients():
    """Unit thin links: HISQ fat = 9/8 eta, Naik = -1/24 eta exactly."""
    L = ko.Lattice((4, 4, 4, 4))
    one = np.zeros((L.V, 4, 3, 3, 2))
    for a in range(3):
        one[:, :, a, a, 0] = 1.0
    U = L.rephase(one)
    fat, lng, nsvd = L.hisq_links(U)
    f_ref, l_ref = free_links(L, 9 / 8, -1 / 24)
    assert nsvd == 0
    assert np.abs(fat - f_ref).max() < 1e-14
    assert np.abs(lng - l_ref).max() < 1e-15
    # asqtad without tadpole factors: also 9/8 and -1/24
    fa, la = L.asqtad_links(U, 1.0)
    assert np.abs(fa - f_ref).max() < 1e-14 and np.abs(la - l_ref).max() < 1e-15


def plane_wave_even(L, n):
    """Even-site plane wave with momentum 2 pi n_i / L_i (antiperiodic t)."""
    c = L.coords.astype(float)
    p = np.array([2 * np.pi * n[0] / L.dims[0], 2 * np.pi * n[1] / L.dims[1], 2 * np.pi * n[2] / L.dims[2],
                  (2 * n[3] + 1) * np.pi / L.dims[3]])
    ph = np.exp(1j * (c @ p))
    chi = np.array([0.3 + 0.1j, -0.7j, 0.2 - 0.4j])
    v = np.zeros((L.V, 3), complex)
    v[:L.Vh] = ph[:L.Vh, None] * chi[None, :]
    return ko.from_complex(v), p


@pytest.mark.parametrize("n", [(0, 0, 0, 0), (1, 0, 2, 1), (3, 1, 1, 2), (2, 2, 3, 5)])
def test_free_field_normal_operator_closed_form(n):
    """(4m^2 - D_eo D_oe) on an even plane wave = (4m^2 + sum_mu (2c1 sin p + 2c3 sin 3p)^2)."""
    L = ko.Lattice((4, 4, 4, 8))
    c1, c3, m = 9 / 8, -1 / 24, 0.07
    fat, lng = free_links(L, c1, c3)
    v, p = plane_wave_even(L, n)
    lam = 4 * m * m + np.sum((2 * c1 * np.sin(p) + 2 * c3 * np.sin(3 * p)) ** 2)
    out = L.mdagm(fat, lng, 4 * m * m, v, ko.EVEN)
    E = L.parity_slice(ko.EVEN)
    assert rel(out[E], lam * v[E]) < 1e-13
    x, it, q = L.congrad(fat, lng, v, m, 1e-13, parity=ko.EVEN, max_iters=100, nrestart=5)
    assert rel(x[E], v[E] / lam) < 1e-12


@pytest.fixture(scope="module")
def hisq44():
    L = ko.Lattice((4, 4, 4, 4))
    U = L.rephase(L.gauge_random(1000))
    fat, lng, nsvd = L.hisq_links(U)
    return L, U, fat, lng


def test_dslash_antihermitian(hisq44):
    L, U, fat, lng = hisq44
    a = L.gaussian_vector(1, ko.ODD)
    b = L.gaussian_vector(2, ko.EVEN)
    Db = L.dslash(fat, lng, b, ko.ODD)      # odd output
    Da = L.dslash(fat, lng, a, ko.EVEN)     # even output
    lhs = np.vdot(ko.to_complex(a), ko.to_complex(Db))
    rhs = -np.vdot(ko.to_complex(Da), ko.to_complex(b))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def random_su3(rng, n):
    A = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    q, r = np.linalg.qr(A)
    q = q * (np.diagonal(r, axis1=1, axis2=2) / np.abs(np.diagonal(r, axis1=1, axis2=2)))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1 / 3)


def test_gauge_covariance_of_smearing_and_dslash(hisq44):
    L, U, fat, lng = hisq44
    rng = np.random.default_rng(7)
    g = random_su3(rng, L.V)
    nbr = np.ctypeslib.as_array(L._g.nbr, shape=(16, L.V))
    Uc = ko.to_complex(U)
    Ug = np.empty_like(Uc)
    for mu in range(4):
        Ug[:, mu] = g @ Uc[:, mu] @ np.conj(np.transpose(g[nbr[mu]], (0, 2, 1)))
    fat2, lng2, _ = L.hisq_links(ko.from_complex(Ug))
    fc, lc = ko.to_complex(fat), ko.to_complex(lng)
    for mu in range(4):
        gd1 = np.conj(np.transpose(g[nbr[mu]], (0, 2, 1)))
        gd3 = np.conj(np.transpose(g[nbr[8 + mu]], (0, 2, 1)))
        assert rel(ko.to_complex(fat2)[:, mu], g @ fc[:, mu] @ gd1) < 1e-13
        assert rel(ko.to_complex(lng2)[:, mu], g @ lc[:, mu] @ gd3) < 1e-13
    s = L.gaussian_vector(3, ko.ODD)
    gs = ko.from_complex(np.einsum("xab,xb->xa", g, ko.to_complex(s)))
    out = L.dslash(fat2, lng2, gs, ko.EVEN)
    ref = np.einsum("xab,xb->xa", g, ko.to_complex(L.dslash(fat, lng, s, ko.EVEN)))
    E = L.parity_slice(ko.EVEN)
    assert rel(ko.to_complex(out)[E], ref[E]) < 1e-13


def test_cg_against_dense_solve(hisq44):
    L, U, fat, lng = hisq44
    m = 0.1
    E = L.parity_slice(ko.EVEN)
    n = L.Vh * 3
    A = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros((L.V, 3), complex)
        e[k // 3, k % 3] = 1.0
        A[:, k] = ko.to_complex(L.mdagm(fat, lng, 4 * m * m, ko.from_complex(e), ko.EVEN))[E].reshape(-1)
    assert np.allclose(A, A.conj().T, atol=1e-13)
    b = L.gaussian_vector(11, ko.EVEN)
    xd = np.linalg.solve(A, ko.to_complex(b)[E].reshape(-1))
    x, it, q = L.congrad(fat, lng, b, m, 1e-13, parity=ko.EVEN, max_iters=500, nrestart=5)
    assert rel(ko.to_complex(x)[E].reshape(-1), xd) < 1e-11


def test_multishift_equals_single_shift_solves(hisq44):
    L, U, fat, lng = hisq44
    m = 0.05
    offsets = [4 * m * m + s for s in (0.0, 0.01, 0.3, 2.0)]
    b = L.gaussian_vector(12, ko.EVEN)
    xs, it, qics = L.multicg_offset(fat, lng, b, offsets, 1e-12, parity=ko.EVEN, max_iters=2000, nrestart=1)
    E = L.parity_slice(ko.EVEN)
    for x, o in zip(xs, offsets):
        r = b - L.mdagm(fat, lng, o, x, ko.EVEN)
        assert np.sqrt(L.norm2(r, ko.EVEN) / L.norm2(b, ko.EVEN)) < 1e-10
        xs1, _, _ = L.congrad(fat, lng, b, 0.5 * np.sqrt(o), 1e-13, parity=ko.EVEN, max_iters=500, nrestart=5)
        assert rel(x[E], xs1[E]) < 1e-9


def test_gauge_generator_is_exp_of_traceless_hermitian():
    L = ko.Lattice((4, 4, 4, 4))
    U = ko.to_complex(L.gauge_random(1000, 0.4)).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.abs(U @ np.conj(np.transpose(U, (0, 2, 1))) - eye).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13
    w, v = np.linalg.eig(U)
    th = np.angle(w)
    H = (v * (th / 0.4)[:, None, :]) @ np.linalg.inv(v)
    assert np.abs(H - np.conj(np.transpose(H, (0, 2, 1)))).max() < 1e-10
    assert np.abs(np.trace(H, axis1=1, axis2=2)).max() < 1e-10
    # H = (A + A^dag)/2 - Tr/3 with standard complex Gaussian A: Var Re H_01 = 1/2
    assert abs(np.var(H[:, 0, 1].real) - 0.5) < 0.05
    # deterministic in the seed, different across seeds
    assert np.array_equal(L.gauge_random(1000), L.gauge_random(1000))
    assert not np.array_equal(L.gauge_random(1000), L.gauge_random(1001))


def test_gaussian_source_statistics():
    L = ko.Lattice((8, 8, 8, 8))
    v = L.gaussian_vector(5, ko.EVEN)
    E, O = L.parity_slice(ko.EVEN), L.parity_slice(ko.ODD)
    assert np.all(v[O] == 0)
    assert abs(np.mean(v[E] ** 2) - 0.5) < 0.02  # <re^2> = <im^2> = 1/2 (grsource_imp.c:169-195)


def test_u3_unitarize_is_polar_factor():
    rng = np.random.default_rng(3)
    for _ in range(20):
        V = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        W, svd = ko.u3_unitarize(ko.from_complex(V))
        u, s, vh = np.linalg.svd(V)
        assert np.abs(ko.to_complex(W) - u @ vh).max() < 1e-12
        assert svd == 0


def test_force_oprod_against_finite_differences(hisq44):
    """dS = sum_x,mu 2 eps(x) Re Tr[dF_mu(x) O1_mu(x) + dL_mu(x) O3_mu(x)],
    S = sum_t r_t phi^dag (M^dag M + sigma_t)^-1 phi, checked by central differences."""
    L, U, fat, lng = hisq44
    m = 0.1
    sig = [0.01, 0.1, 1.0]
    res = [0.2, 0.5, 0.9]
    offs = [4 * m * m + s for s in sig]
    phi = L.gaussian_vector(21, ko.EVEN)
    E = L.parity_slice(ko.EVEN)

    def action(f, l):
        xs, _, _ = L.multicg_offset(f, l, phi, offs, 1e-14, parity=ko.EVEN, max_iters=3000, nrestart=1)
        return sum(r * np.vdot(ko.to_complex(phi)[E], ko.to_complex(x)[E]).real for r, x in zip(res, xs))

    xs, _, _ = L.multicg_offset(fat, lng, phi, offs, 1e-14, parity=ko.EVEN, max_iters=3000, nrestart=1)
    Z = []
    for x in xs:
        z = x.copy()
        y = L.dslash(fat, lng, x, ko.ODD)
        z[L.parity_slice(ko.ODD)] = y[L.parity_slice(ko.ODD)]
        Z.append(z)
    O1, O3 = L.force_oprod(Z, res)
    rng = np.random.default_rng(9)
    h = 1e-5
    for trial in range(6):
        which = trial % 2
        i = int(rng.integers(L.V))
        mu, a, b, part = (int(v) for v in rng.integers(0, [4, 3, 3, 2]))
        base = fat if which == 0 else lng
        O = O1 if which == 0 else O3
        dp, dm = base.copy(), base.copy()
        dp[i, mu, a, b, part] += h
        dm[i, mu, a, b, part] -= h
        fd = (action(dp, lng) - action(dm, lng)) / (2 * h) if which == 0 else \
             (action(fat, dp) - action(fat, dm)) / (2 * h)
        eps = 1.0 if i < L.Vh else -1.0
        Oc = ko.to_complex(O)[i, mu]
        pred = 2 * eps * (np.conj(Oc[b, a]).real if part == 0 else np.conj(Oc[b, a]).imag)
        assert abs(fd - pred) < 1e-6 * max(1.0, abs(pred)), (trial, fd, pred)


# --------------------------------------------------------------------------
# thin-link forces: chain rule checked by central differences of the action
# --------------------------------------------------------------------------
def test_u3_adjoint_finite_differences():
    """2 Re Tr[dW F_W] = 2 Re Tr[dV F_V] for W = V (V^dag V)^-1/2, by central
    differences of the oracle's own projection in all 18 real directions."""
    rng = np.random.default_rng(31)
    for trial in range(4):
        Vc = 0.6 * np.eye(3) + 0.25 * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
        if trial == 3:  # nearly degenerate singular values
            Vc = 0.8 * np.eye(3) + 1e-4 * (rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))
        Fc = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        V, F = ko.from_complex(Vc), ko.from_complex(Fc)
        FV = ko.to_complex(ko.u3_adjoint(V, F))
        h = 1e-6
        for a in range(3):
            for b in range(3):
                for part in (0, 1):
                    dp, dm = V.copy(), V.copy()
                    dp[a, b, part] += h
                    dm[a, b, part] -= h
                    Wp = ko.to_complex(ko.u3_unitarize(dp)[0])
                    Wm = ko.to_complex(ko.u3_unitarize(dm)[0])
                    fd = 2 * np.trace((Wp - Wm) / (2 * h) @ Fc).real
                    dV = np.zeros((3, 3), complex)
                    dV[a, b] = 1.0 if part == 0 else 1j
                    pred = 2 * np.trace(dV @ FV).real
                    assert abs(fd - pred) < 1e-7 * max(1.0, abs(pred)), (trial, a, b, part, fd, pred)


def _su3_generator(rng):
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    T = (A + A.conj().T) / 2
    return T - np.trace(T) / 3 * np.eye(3)


def _expm_herm(T, t):
    w, v = np.linalg.eigh(T)
    return (v * np.exp(1j * t * w)) @ v.conj().T


def _thin_link_fd(L, U, links_fn, force_fn, seed, ntrial=4):
    """S(U) = sum_t r_t Re<phi, (M^dag M + sigma_t)^-1 phi> on even sites; for
    U_mu(x) -> exp(i t T) U_mu(x): dS/dt = Re Tr[i T mom_mu(x)] / eps."""
    m = 0.1
    sig = [0.02, 0.3]
    res = [0.4, 0.9]
    offs = [4 * m * m + s for s in sig]
    phi = L.gaussian_vector(41, ko.EVEN)
    E, O = L.parity_slice(ko.EVEN), L.parity_slice(ko.ODD)

    def solve(Ul):
        fat, lng = links_fn(Ul)
        xs, _, _ = L.multicg_offset(fat, lng, phi, offs, 1e-14, parity=ko.EVEN, max_iters=4000, nrestart=1)
        return fat, lng, xs

    def action(Ul):
        _, _, xs = solve(Ul)
        return sum(r * float(np.sum(phi[E] * x[E])) for r, x in zip(res, xs))

    fat, lng, xs = solve(U)
    Z = []
    for x in xs:
        z = x.copy()
        z[O] = L.dslash(fat, lng, x, ko.ODD)[O]
        Z.append(z)
    eps = 0.7
    mom = ko.to_complex(force_fn(U, Z, res, eps))
    rng = np.random.default_rng(seed)
    h = 1e-5
    worst = 0.0
    for trial in range(ntrial):
        i = int(rng.integers(L.V))
        mu = int(rng.integers(4))
        T = _su3_generator(rng)
        Uc = ko.to_complex(U)
        vals = []
        for sg in (1, -1):
            Up = Uc.copy()
            Up[i, mu] = _expm_herm(T, sg * h) @ Uc[i, mu]
            vals.append(action(ko.from_complex(Up)))
        fd = (vals[0] - vals[1]) / (2 * h)
        pred = np.trace(1j * T @ mom[i, mu]).real / eps
        worst = max(worst, abs(fd - pred) / max(1e-3, abs(pred)))
        assert abs(fd - pred) < 2e-6 * max(1.0, abs(pred)), (trial, i, mu, fd, pred)
    return worst


def test_fn_force_thin_link_finite_differences():
    """asqtad links (u0 = 0.9) and kso_fn_force against central differences
    of the action through the link construction."""
    L = ko.Lattice((4, 4, 4, 4))
    U = L.rephase(L.gauge_random(77, 0.3))
    c = ko.asqtad_coeffs(0.9)
    _thin_link_fd(L, U, lambda Ul: L.asqtad_links(Ul, 0.9),
                  lambda Ul, Z, r, e: L.fn_force(c, Ul, Z, r, eps=e), seed=5)


def test_hisq_force_thin_link_finite_differences():
    """HISQ chain (fat7 -> U(3) projection -> level 2 + Naik) and
    kso_hisq_force against central differences of the action."""
    L = ko.Lattice((4, 4, 4, 4))
    U = L.rephase(L.gauge_random(78, 0.3))
    _thin_link_fd(L, U, lambda Ul: L.hisq_links(Ul)[:2],
                  lambda Ul, Z, r, e: L.hisq_force(Ul, Z, r, eps=e), seed=6)


def test_hisq_force_second_naik_set_finite_differences():
    """Links of a second Naik set (X_2 + eps_naik X_3, eps_naik = -0.3) and
    kso_hisq_force_naik with every term in that set, against central
    differences of the action built on those links."""
    L = ko.Lattice((4, 4, 4, 4))
    U = L.rephase(L.gauge_random(79, 0.3))
    epsn = -0.3
    _thin_link_fd(L, U, lambda Ul: L.hisq_links_naik(Ul, epsn),
                  lambda Ul, Z, r, e: L.hisq_force(Ul, Z, r, eps=e, n_naik=len(Z), eps_naik=epsn), seed=7)


def test_staple_form_force_equals_path_table_form():
    """The reverse-mode staple form of the smearing adjoint gives the same
    thin-link momentum as the path-table form (asqtad FN and HISQ chains) on a
    non-cubic lattice."""
    L = ko.Lattice((4, 6, 4, 8))
    U = L.rephase(L.gauge_random(12, 0.4))
    xs = [L.gaussian_vector(40 + t, ko.EVENANDODD) for t in range(2)]
    res = [0.5, -0.3]
    try:
        for f in (lambda: L.fn_force(ko.asqtad_coeffs(0.8), U, xs, res, eps=0.4),
                  lambda: L.hisq_force(U, xs, res, eps=0.4)):
            ko.set_force_form(0)
            a = f()
            ko.set_force_form(1)
            b = f()
            assert rel(b, a) < 1e-13
    finally:
        ko.set_force_form(0)
