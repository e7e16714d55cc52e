"""One RHMC trajectory on the device (milc_hip_dev_rhmc_trajectory, the
update() of ks_imp_rhmc/update_rhmc.c): gauge force + HISQ fermion force,
link updates with the fat and Naik links rebuilt, reunitarisation, actions,
accept/reject.

* MILC's own regression trajectories: the HISQ one, ks_imp_rhmc/test/su3_rhmc_hisq.1
  (lat.sample.l6666.hisq, iseed 86658, four pseudofermions of
  rationals.sample.su3_rhmc_hisq, INT_3G1F, 10 steps of 0.02, beta 7.4,
  u0 0.89): with MILC's per-site generators the heatbath noise reproduces the
  printed GRSOURCE sums to every digit and the start actions the printed
  ACTION line; dS and the final plaquette agree with the printed ones within
  the tolerances of MILC's own check (su3_rhmc_hisq.1.errtol: delta S 0.02,
  PLAQ 4e-5); and the asqtad one, su3_omelyan_rhmc.1 (one pseudofermion,
  Omelyan), whose start and end actions, delta S and plaquette agree within
  su3_omelyan_rhmc.1.errtol;
* the device generators against the oracle's restatement of MILC's: momenta
  identical, sources identical;
* the starting actions against the oracle: beta/3 times the gauge-loop sum of
  the device links, and phi^dag R_MD phi with phi = R_GR chi, chi the
  device's Gaussian source of the trajectory's seed (FP64 solves, 1e-9);
* energy conservation: with the fermion action the MD force's own
  (action_from_md), |dH| falls by about 4 when eps halves at fixed
  trajectory length (leapfrog); Omelyan and the nG1F integrators do better;
* accept = 2 restores the starting links exactly; accept = 0 keeps links
  that are SU(3) after the reunitarisation.

Workload of the small tests: 4x4x4x8 lattice, gauge seed 1200 (scale 0.4),
m = 0.1, beta = 5, loop coefficients for u0 = 0.86 with 2+1+1 flavours, a
two-pole MD and a one-pole GR rational (positive poles and residues).
"""
import os

import numpy as np
import pytest

import ks_io as oio
import ks_oracle as ko
import milc_qcd_amd as mq
from milc_qcd_amd import lattice_io as lio

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DIMS = (4, 4, 4, 8)
MASS, BETA = 0.1, 5.0
MD = ([0.5, 0.3], [0.01, 0.3], 0.1)   # residues, poles, constant
GR = ([0.2], [0.5], 1.0)
COEFF = ko.gauge_coeffs(0.86, 4, hisq=True)


def rat(r):
    return mq.ks.Rational.make(r[2], r[0], r[1])


def splitmix(z):
    M = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def pf(resid_md=1e-12):
    return mq.Pseudofermion(mass=MASS, md=rat(MD), gr=rat(GR), fa=rat(MD), resid_md=resid_md, resid_gr=1e-12,
                            resid_fa=1e-12)


def traj(G, steps, eps, integrator=mq.INT_LEAPFROG, accept=2, seed=7, mixed_md=0, resid_md=1e-12):
    return G.dev_rhmc_trajectory(steps, eps, BETA, COEFF, [pf(resid_md)], integrator=integrator, mixed_md=mixed_md,
                                 action_from_md=True, accept=accept, seed=seed)


# ---- MILC's su3_rhmc_hisq.1 regression run (sample-out lines cited) --------
MILC_ISEED = 86658
MILC_GRSOURCE = [1.9636750663e+03, 1.9027332576e+03, 1.9328576741e+03, 1.9363357361e+03]  # 526-532
MILC_ACTION0 = (1.91689577889074e+04, -8.03776340931654e+01, 7.73560159669816e+03)       # 539
MILC_ACTION1 = (1.89814624700873e+04, 1.10639809951186e+02, 7.73207863940299e+03)        # 608
MILC_DELTA_S = -8.320708e-04                                                               # 610
MILC_PLAQ = (1.859026, 1.864015)                                                           # 611
# cgresid_md_fa_gr of su3_rhmc_hisq.1.sample-in, per pseudofermion
MILC_RESID = [(5e-4, 1e-4, 1e-4)] + [(5e-6, 1e-6, 1e-6)] * 3


@pytest.mark.parametrize("mixed", [0, 1])
def test_milc_su3_rhmc_hisq_trajectory(mixed):
    links, info = lio.reload_lattice(os.path.join(HERE, "golden", "io", "lat.sample.l6666.hisq"))
    assert info["checksum_ok"]
    dims = (6, 6, 6, 6)
    pfs = mq.pseudofermions_from_file(os.path.join(HERE, "golden", "rationals.sample.su3_rhmc_hisq"),
                                      [r[0] for r in MILC_RESID], [r[2] for r in MILC_RESID],
                                      [r[1] for r in MILC_RESID])
    assert len(pfs) == 4
    coeff = mq.gauge_loop_coeffs(0.89, 3)
    with mq.Lattice(dims) as G:
        G.dev_load_thin_links(links, phases_in=False)
        G.dev_seed_rng(MILC_ISEED, precision=1)
        st = G.dev_rhmc_trajectory(10, 0.02, 7.4, coeff, pfs, integrator=mq.INT_3G1F, mixed_md=mixed,
                                   mixed_action=mixed, accept=1, site_rng=True)
        U1 = G.dev_read_thin_links(with_phases=False)
    print({k: v for k, v in st.items()})
    np.testing.assert_allclose(st["grsource"], MILC_GRSOURCE, rtol=3e-11)
    assert st["g0"] == pytest.approx(MILC_ACTION0[0], rel=2e-7)
    assert st["h0"] == pytest.approx(MILC_ACTION0[1], abs=1e-4)   # MILC sums the float momenta in float
    assert st["f0"] == pytest.approx(MILC_ACTION0[2], abs=0.05)
    assert st["fermion_forces"] == 10 and st["gauge_forces"] == 30 and st["link_rebuilds"] == 11
    assert st["reunit_errors"] == 0
    assert st["accepted"] == 1
    assert abs(st["dH"] - MILC_DELTA_S) < 0.02                     # su3_rhmc_hisq.1.errtol
    ss, st_ = oio.plaquette(U1[oio.node_index_of_lex(list(dims))], list(dims))
    assert abs(ss - MILC_PLAQ[0]) < 4e-5 and abs(st_ - MILC_PLAQ[1]) < 4e-5
    for k, ref in zip(("g1", "h1", "f1"), MILC_ACTION1):
        assert abs(st[k] - ref) < 0.05, (k, st[k], ref)


# ---- MILC's su3_omelyan_rhmc.1 regression run (asqtad; sample-out lines cited) ----
ASQ_ISEED = 3546789
ASQ_GRSOURCE = 1.9073263760e+03                                                     # 237
ASQ_ACTION0 = (1.67770842255353e+04, -9.36446401476860e+01, 1.90732694119215e+03)   # 241
ASQ_ACTION1 = (1.68225951988426e+04, -1.32490214347839e+02, 1.90066177399457e+03)   # 257
ASQ_DELTA_S = 2.319096e-04                                                           # 259
ASQ_PLAQ = (1.649426, 1.643181)                                                      # 260
# su3_omelyan_rhmc.1.errtol: ACTION g, h, f at the start and the end, delta S, PLAQ
ASQ_TOL0, ASQ_TOL1, ASQ_TOL_DS, ASQ_TOL_PLAQ = (0.03, 0.0002, 0.004), (0.03, 0.004, 0.004), 0.01, 3e-6


@pytest.mark.parametrize("mixed", [0, 1])
def test_milc_su3_omelyan_rhmc_asqtad_trajectory(mixed):
    """MILC's asqtad RHMC regression run ks_imp_rhmc/test/su3_omelyan_rhmc.1
    (lat.sample.l6666b650m010m050, iseed 3546789, one pseudofermion of
    rationals.sample.su3_omelyan_rhmc, INT_OMELYAN, 12 steps of 0.01,
    beta 6.5, asqtad with u0 0.862, residuals 2e-6): the first trajectory
    within every tolerance of MILC's own check"""
    links, info = lio.reload_lattice(os.path.join(HERE, "golden", "io", "lat.sample.l6666b650m010m050"))
    assert info["checksum_ok"]
    dims = (6, 6, 6, 6)
    ss, st0 = oio.plaquette(links[oio.node_index_of_lex(list(dims))], list(dims))
    assert abs(ss - 1.658265) < 1e-6 and abs(st0 - 1.639891) < 1e-6                  # CHECK PLAQ, line 217
    pfs = mq.pseudofermions_from_file(os.path.join(HERE, "golden", "rationals.sample.su3_omelyan_rhmc"),
                                      2e-6, 2e-6, 2e-6)
    coeff = mq.gauge_loop_coeffs(0.862, 3, hisq=False)
    np.testing.assert_allclose(coeff[1:], [-7.355018e-02, -8.663654e-03], rtol=1e-6)  # lines 232-234
    with mq.Lattice(dims) as G:
        G.dev_load_thin_links(links, phases_in=False)
        G.dev_seed_rng(ASQ_ISEED, precision=1)
        st = G.dev_rhmc_trajectory(12, 0.01, 6.5, coeff, pfs, integrator=mq.INT_OMELYAN, mixed_md=mixed,
                                   mixed_action=mixed, accept=1, site_rng=True, asqtad_u0=0.862)
        U1 = G.dev_read_thin_links(with_phases=False)
    print({k: v for k, v in st.items()})
    assert st["grsource"][0] == pytest.approx(ASQ_GRSOURCE, rel=3e-11)
    for k, ref, tol in zip(("g0", "h0", "f0"), ASQ_ACTION0, ASQ_TOL0):
        assert abs(st[k] - ref) < tol, (k, st[k], ref)
    for k, ref, tol in zip(("g1", "h1", "f1"), ASQ_ACTION1, ASQ_TOL1):
        assert abs(st[k] - ref) < tol, (k, st[k], ref)
    assert abs(st["dH"] - ASQ_DELTA_S) < ASQ_TOL_DS
    assert st["accepted"] == 1 and st["fermion_forces"] == 12 and st["reunit_errors"] == 0
    ss, st_ = oio.plaquette(U1[oio.node_index_of_lex(list(dims))], list(dims))
    print("PLAQ", ss, st_)
    assert abs(ss - ASQ_PLAQ[0]) < ASQ_TOL_PLAQ and abs(st_ - ASQ_PLAQ[1]) < ASQ_TOL_PLAQ


# ---- MILC's su3_leapfrog.1 run (asqtad RHMD, two pseudofermions; sample-out lines) ----
LF_GRSOURCE = [1.9073263760e+03, 1.9029227323e+03]                                  # 306, 308
LF_ACTION0 = (1.67770842255353e+04, -9.36446401476860e+01, 3.81024987246096e+03)    # 313
LF_ACTION1 = (1.68258343205252e+04, -1.56867322713137e+02, 3.82467601203918e+03)    # 358
LF_DELTA_S = -4.644800e-02                                                            # 360
LF_PLAQ = (1.649295, 1.642828)                                                        # 361
# su3_leapfrog.1.errtol: start and end ACTION g, h, f; delta S; PLAQ
LF_TOL0, LF_TOL1, LF_TOL_DS, LF_TOL_PLAQ = (3e-4, 2e-4, 0.01), (0.01, 0.008, 0.007), 0.008, (2e-6, 1e-7)


@pytest.mark.parametrize("mixed", [0, 1])
def test_milc_su3_leapfrog_asqtad_rhmd_trajectory(mixed):
    """MILC's asqtad RHMD regression run ks_imp_rhmc/test/su3_leapfrog.1
    (lat.sample.l6666b650m010m050, iseed 3546789, two pseudofermions of
    rationals.sample.su3_leapfrog, INT_LEAPFROG, 12 steps of 0.01, beta 6.5,
    asqtad with u0 0.862, residuals 2e-6, no accept/reject): the first
    trajectory within every tolerance of MILC's own check"""
    links, info = lio.reload_lattice(os.path.join(HERE, "golden", "io", "lat.sample.l6666b650m010m050"))
    dims = (6, 6, 6, 6)
    pfs = mq.pseudofermions_from_file(os.path.join(HERE, "golden", "rationals.sample.su3_leapfrog"),
                                      2e-6, 2e-6, 2e-6)
    assert len(pfs) == 2
    coeff = mq.gauge_loop_coeffs(0.862, 3, hisq=False)
    with mq.Lattice(dims) as G:
        G.dev_load_thin_links(links, phases_in=False)
        G.dev_seed_rng(ASQ_ISEED, precision=1)
        st = G.dev_rhmc_trajectory(12, 0.01, 6.5, coeff, pfs, integrator=mq.INT_LEAPFROG, mixed_md=mixed,
                                   mixed_action=mixed, accept=0, site_rng=True, asqtad_u0=0.862)
        U1 = G.dev_read_thin_links(with_phases=False)
    print({k: v for k, v in st.items()})
    np.testing.assert_allclose(st["grsource"], LF_GRSOURCE, rtol=3e-11)
    for k, ref, tol in zip(("g0", "h0", "f0"), LF_ACTION0, LF_TOL0):
        assert abs(st[k] - ref) < tol, (k, st[k], ref)
    for k, ref, tol in zip(("g1", "h1", "f1"), LF_ACTION1, LF_TOL1):
        assert abs(st[k] - ref) < tol, (k, st[k], ref)
    assert abs(st["dH"] - LF_DELTA_S) < LF_TOL_DS
    assert st["fermion_forces"] == 12
    ss, st_ = oio.plaquette(U1[oio.node_index_of_lex(list(dims))], list(dims))
    print("PLAQ", ss, st_)
    assert abs(ss - LF_PLAQ[0]) < LF_TOL_PLAQ[0]
    # MILC prints 6 decimals; its 1e-7 tolerance is below that rounding
    assert abs(st_ - LF_PLAQ[1]) < 5e-7 + LF_TOL_PLAQ[1]


# ---- MILC's su3_rhmc_hisq.f211.1 run (2+1+1, the charm pseudofermion on a second Naik set) ----
F211_ISEED = 4563421
F211_GRSOURCE = [1.9606615161e+03, 1.9419062139e+03, 1.9778946018e+03, 2.0146820996e+03,
                 1.9699683368e+03]                                                               # 635-643
F211_ACTION0 = (2.08587740172183e+04, -2.05282869219780e+02, 9.86511273326725e+03)            # 651
F211_ACTION1 = (1.98409032770018e+04, 8.18777145951986e+02, 9.85889961154759e+03)             # 770
F211_DELTA_S = -2.384676e-02                                                                    # 772
F211_PLAQ = (1.746504, 1.737158)                                                                # 773
F211_RESID = [(5e-4, 1e-4, 1e-4)] + [(5e-6, 1e-6, 1e-6)] * 4


@pytest.mark.parametrize("mixed", [0, 1])
def test_milc_su3_rhmc_hisq_f211_trajectory(mixed):
    """MILC's 2+1+1 HISQ regression run ks_imp_rhmc/test/su3_rhmc_hisq.f211.1
    (lat.sample.hisq.f211b670.l6666, iseed 4563421, five pseudofermions of
    rationals.sample.su3_rhmc_hisq_f211, the last with naik_term_epsilon
    -0.16875, INT_3G1F, 12 steps of 0.04, beta 6.7, u0 0.89 with four
    flavours): GRSOURCE to every digit, start actions, delta S and the final
    plaquette within su3_rhmc_hisq.f211.1.errtol (delta S 0.01, PLAQ 4e-6)"""
    links, info = lio.reload_lattice(os.path.join(HERE, "golden", "io", "lat.sample.hisq.f211b670.l6666"))
    assert info["checksum_ok"]
    dims = (6, 6, 6, 6)
    pfs = mq.pseudofermions_from_file(os.path.join(HERE, "golden", "rationals.sample.su3_rhmc_hisq_f211"),
                                      [r[0] for r in F211_RESID], [r[2] for r in F211_RESID],
                                      [r[1] for r in F211_RESID])
    assert len(pfs) == 5 and pfs[4].naik_epsilon == -0.16875 and pfs[0].naik_epsilon == 0.0
    coeff = mq.gauge_loop_coeffs(0.89, 4)
    with mq.Lattice(dims) as G:
        G.dev_load_thin_links(links, phases_in=False)
        G.dev_seed_rng(F211_ISEED, precision=1)
        st = G.dev_rhmc_trajectory(12, 0.04, 6.7, coeff, pfs, integrator=mq.INT_3G1F, mixed_md=mixed,
                                   mixed_action=mixed, accept=1, site_rng=True)
        U1 = G.dev_read_thin_links(with_phases=False)
    print({k: v for k, v in st.items()})
    np.testing.assert_allclose(st["grsource"], F211_GRSOURCE, rtol=3e-11)
    assert st["g0"] == pytest.approx(F211_ACTION0[0], rel=2e-7)
    assert st["h0"] == pytest.approx(F211_ACTION0[1], abs=1e-4)
    assert st["f0"] == pytest.approx(F211_ACTION0[2], abs=0.05)
    assert st["fermion_forces"] == 12 and st["reunit_errors"] == 0
    assert st["accepted"] == 1
    assert abs(st["dH"] - F211_DELTA_S) < 0.01                      # su3_rhmc_hisq.f211.1.errtol
    ss, st_ = oio.plaquette(U1[oio.node_index_of_lex(list(dims))], list(dims))
    print("PLAQ", ss, st_)
    assert abs(ss - F211_PLAQ[0]) < 4e-6 + 5e-7 and abs(st_ - F211_PLAQ[1]) < 4e-6 + 5e-7
    for k, ref in zip(("g1", "h1", "f1"), F211_ACTION1):
        assert abs(st[k] - ref) < 0.05, (k, st[k], ref)


# ---- MILC's su3_rhmc_hisq.f2111.1 run (six pseudofermions on three Naik sets) ----
F2111_GRSOURCE = F211_GRSOURCE + [1.9529017883e+03]                                           # 741-751
F2111_ACTION0 = (2.16927145638990e+04, -2.05282869219780e+02, 1.18180142822787e+04)          # 760
F2111_ACTION1 = (2.02631810822886e+04, 1.23237197795510e+03, 1.18098377667516e+04)           # 892
F2111_DELTA_S = -5.514996e-02                                                                  # 894
F2111_PLAQ = (1.769942, 1.759599)                                                              # 895


@pytest.mark.parametrize("mixed", [0, 1])
def test_milc_su3_rhmc_hisq_f2111_trajectory(mixed):
    """MILC's ks_imp_rhmc/test/su3_rhmc_hisq.f2111.1 (the f211 lattice and
    seed, six pseudofermions of rationals.sample.su3_rhmc_hisq_f2111, the last
    two with naik_term_epsilon -0.1520 and -0.2712: three Naik sets; five
    flavours in the gauge action): within su3_rhmc_hisq.f2111.1.errtol
    (delta S 0.01, PLAQ 2e-5)"""
    links, info = lio.reload_lattice(os.path.join(HERE, "golden", "io", "lat.sample.hisq.f211b670.l6666"))
    dims = (6, 6, 6, 6)
    res = [(5e-4, 1e-4, 1e-4)] + [(5e-6, 1e-6, 1e-6)] * 5
    pfs = mq.pseudofermions_from_file(os.path.join(HERE, "golden", "rationals.sample.su3_rhmc_hisq_f2111"),
                                      [r[0] for r in res], [r[2] for r in res], [r[1] for r in res])
    assert len(pfs) == 6 and len({p.naik_epsilon for p in pfs}) == 3
    coeff = mq.gauge_loop_coeffs(0.89, 5)
    with mq.Lattice(dims) as G:
        G.dev_load_thin_links(links, phases_in=False)
        G.dev_seed_rng(F211_ISEED, precision=1)
        st = G.dev_rhmc_trajectory(12, 0.04, 6.7, coeff, pfs, integrator=mq.INT_3G1F, mixed_md=mixed,
                                   mixed_action=mixed, accept=1, site_rng=True)
        U1 = G.dev_read_thin_links(with_phases=False)
    print({k: v for k, v in st.items()})
    np.testing.assert_allclose(st["grsource"], F2111_GRSOURCE, rtol=3e-11)
    assert st["g0"] == pytest.approx(F2111_ACTION0[0], rel=2e-7)
    assert st["f0"] == pytest.approx(F2111_ACTION0[2], abs=0.05)
    assert st["accepted"] == 1 and st["reunit_errors"] == 0
    assert abs(st["dH"] - F2111_DELTA_S) < 0.01                     # su3_rhmc_hisq.f2111.1.errtol
    ss, st_ = oio.plaquette(U1[oio.node_index_of_lex(list(dims))], list(dims))
    print("PLAQ", ss, st_)
    assert abs(ss - F2111_PLAQ[0]) < 2e-5 and abs(st_ - F2111_PLAQ[1]) < 2e-5
    for k, ref in zip(("g1", "h1", "f1"), F2111_ACTION1):
        assert abs(st[k] - ref) < 0.05, (k, st[k], ref)


def test_device_generators_match_oracle():
    """milc_hip_dev_ranmom / dev_gaussian_source draw MILC's numbers: the
    oracle's restatement (pinned to MILC's printed sums) gives the same values"""
    dims = (6, 4, 8, 10)
    L = ko.Lattice(dims)
    E = L.parity_slice(ko.EVEN)
    O = L.parity_slice(ko.ODD)
    for single in (True, False):
        st = L.prn_seed(4321)
        mom_ref = L.ranmom_prn(st, single=single)
        v_ref, s_ref = L.gaussian_prn(st, ko.EVEN, single=single)
        w_ref, t_ref = L.gaussian_prn(st, ko.ODD, single=single)
        with mq.Lattice(dims) as G:
            G.dev_seed_rng(4321, precision=1 if single else 2)
            mom = G.dev_ranmom()
            s = G.dev_gaussian_source(mq.EVEN, 3)
            t = G.dev_gaussian_source(mq.ODD, 4)
            v = G.dev_read(3, mq.EVEN)
            w = G.dev_read(4, mq.ODD)
        assert np.abs(mom - mom_ref).max() < 1e-15
        assert np.abs(v[E] - v_ref[E]).max() < 1e-15 and np.abs(w[O] - w_ref[O]).max() < 1e-15
        # single: the sum is MILC's float accumulation, the oracle's is FP64
        assert s == pytest.approx(s_ref, rel=1e-7 if single else 1e-13)
        assert t == pytest.approx(t_ref, rel=1e-7 if single else 1e-13)


def test_reunitarize_vs_oracle():
    """milc_hip_reunitarize with the KS phases in = the oracle's reunit_su3 of
    the bare links, phases put back"""
    L = ko.Lattice(DIMS)
    rng = np.random.default_rng(11)
    U = L.gauge_random(1201, 0.4)
    P = np.ascontiguousarray(U + 2e-3 * rng.standard_normal(U.shape))
    ref = P.copy()
    n_ref, md_ref = L.reunitarize(ref, tol=1e-4)
    with mq.Lattice(DIMS) as G:
        X = L.rephase(P.copy())
        n, md = G.reunitarize(X, tolerance=1e-4, phases_in=True)
        Y = P.copy()
        n2, md2 = G.reunitarize(Y, tolerance=1e-4, phases_in=False)
    assert np.abs(L.rephase(X) - ref).max() < 1e-14
    assert np.abs(Y - ref).max() < 1e-14
    assert n == n_ref == n2 and md == pytest.approx(md_ref, rel=1e-10) and md2 == pytest.approx(md_ref, rel=1e-10)


def test_start_actions_vs_oracle():
    L = ko.Lattice(DIMS)
    U = L.rephase(L.gauge_random(1200))
    fat, lng, _ = L.hisq_links(U)
    chi = L.gaussian_vector(splitmix(7 + 1), ko.EVEN)
    m4 = 4 * MASS * MASS
    E = L.parity_slice(ko.EVEN)
    xs, _, _ = L.multicg_offset(fat, lng, chi, [m4 + p for p in GR[1]], 1e-14, parity=ko.EVEN, max_iters=5000,
                                nrestart=1)
    phi = GR[2] * chi + sum(r * x for r, x in zip(GR[0], xs))
    ys, _, _ = L.multicg_offset(fat, lng, phi, [m4 + p for p in MD[1]], 1e-14, parity=ko.EVEN, max_iters=5000,
                                nrestart=1)
    f_ref = MD[2] * float(np.sum(phi[E] ** 2)) + sum(r * float(np.sum(phi[E] * y[E])) for r, y in zip(MD[0], ys))
    g_ref = BETA / 3 * L.gauge_action(COEFF, np.ascontiguousarray(L.gauge_random(1200)))
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        st = traj(G, 2, 0.02)
        print(st)
    assert abs(st["g0"] - g_ref) < 1e-12 * abs(g_ref)
    assert abs(st["f0"] - f_ref) < 1e-9 * abs(f_ref)
    assert st["grsource"][0] == pytest.approx(float(np.sum(chi[E] ** 2)), rel=1e-12)


def test_energy_conservation_and_restore():
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        U0 = G.dev_read_thin_links()
        a = traj(G, 4, 0.05)
        assert np.array_equal(G.dev_read_thin_links(), U0)  # accept = 2 restored the links
        b = traj(G, 8, 0.025)
        o = traj(G, 4, 0.05, integrator=mq.INT_OMELYAN)
        print("leapfrog dH", a["dH"], b["dH"], "omelyan", o["dH"])
        assert 3.0 < abs(a["dH"]) / abs(b["dH"]) < 5.5
        assert abs(o["dH"]) < abs(a["dH"])
        assert a["fermion_forces"] == 4 and o["fermion_forces"] == 4
        # the links are rebuilt for each fermion kick, the final actions and,
        # with accept = 2, once more for the restored links
        assert a["link_rebuilds"] == 6 and o["link_rebuilds"] == 6
        # accept = 0 keeps the evolved links, which are SU(3) after the reunitarisation
        k = traj(G, 4, 0.05, accept=0)
        assert k["accepted"] == 1 and k["reunit_errors"] == 0 and 0 < k["reunit_max_deviation"] < 1e-12
        U1 = G.dev_read_thin_links()
        assert not np.array_equal(U1, U0)
        Uc = U1[..., 0] + 1j * U1[..., 1]
        dev = np.abs(np.conj(np.swapaxes(Uc, -1, -2)) @ Uc - np.eye(3)).max()
        assert dev < 1e-14
        assert np.abs(np.linalg.det(Uc) - 1).max() < 1e-14


def test_failed_trajectory_leaves_starting_links_of_its_action():
    """Error contract of milc_hip_dev_rhmc_trajectory: a failure at any stage
    (injected from the host: after the heatbath, the start actions, the MD)
    leaves the starting thin links and the fat and Naik links of that quark
    action in place, and a following trajectory of the other action rebuilds
    its own links."""
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        U0 = G.dev_read_thin_links(with_phases=True)
        hisq = G.dev_read_links()
        # an asqtad trajectory that is rejected leaves the asqtad links of U0
        G.dev_rhmc_trajectory(2, 0.02, BETA, COEFF, [pf()], integrator=mq.INT_LEAPFROG, mixed_md=0, accept=2,
                              seed=3, asqtad_u0=0.86)
        asq = G.dev_read_links()
        assert np.array_equal(G.dev_read_thin_links(with_phases=True), U0)
        assert not np.allclose(asq[0], hisq[0])
        for stage in (1, 2, 3):
            # the HISQ links of U0 first (rejected HISQ trajectory)
            traj(G, 2, 0.02)
            assert np.array_equal(G.dev_read_links()[0], hisq[0])
            G.debug_fail_trajectory(stage)
            with pytest.raises(mq.MilcHipError):
                G.dev_rhmc_trajectory(2, 0.02, BETA, COEFF, [pf()], integrator=mq.INT_LEAPFROG, mixed_md=0,
                                      accept=0, seed=3, asqtad_u0=0.86)
            assert np.array_equal(G.dev_read_thin_links(with_phases=True), U0), stage
            f, l = G.dev_read_links()
            assert np.array_equal(f, asq[0]) and np.array_equal(l, asq[1]), stage
        # the next HISQ trajectory rebuilds the HISQ links of the starting links
        traj(G, 2, 0.02)
        f, l = G.dev_read_links()
        assert np.array_equal(f, hisq[0]) and np.array_equal(l, hisq[1])


@pytest.mark.parametrize("alg,ngauge", [(mq.INT_2G1F, 2), (mq.INT_3G1F, 3), (mq.INT_5G1F, 5), (mq.INT_6G1F, 6)])
def test_multi_timescale_integrators(alg, ngauge):
    """INT_nG1F: n gauge kicks per fermion kick; dH falls about 4x when eps
    halves, and stays below the leapfrog's at the same step"""
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        a = traj(G, 4, 0.05, integrator=alg)
        b = traj(G, 8, 0.025, integrator=alg)
        lf = traj(G, 4, 0.05)
    print(alg, a["dH"], b["dH"], lf["dH"])
    assert a["fermion_forces"] == 4 and a["gauge_forces"] == 4 * ngauge
    assert a["link_rebuilds"] == 6
    assert 2.5 < abs(a["dH"]) / abs(b["dH"]) < 6.0
    assert abs(a["dH"]) < abs(lf["dH"])


def test_mixed_md_solves_track_fp64():
    """mixed-precision MD solves (FP32 multi-shift + 16-bit refinement, links
    rebuilt each step) give the FP64-solve trajectory's dH to the solver
    tolerance"""
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        a = traj(G, 4, 0.05)
        b = traj(G, 4, 0.05, mixed_md=1, resid_md=1e-10)
    assert abs(a["dH"] - b["dH"]) < 1e-5 * max(1.0, abs(a["dH"]))


def test_batched_pseudofermion_solves_match_one_source_solves():
    """three pseudofermions with different masses and rational orders: the
    batched multi-shift with a shift set per source (heatbath, MD and action
    solves) reproduces the one-source FP64 solves' actions and dH"""
    def pfs(resid):
        return [mq.Pseudofermion(mass=m, md=rat(MD), gr=rat(g), fa=rat(MD), resid_md=resid, resid_gr=resid,
                                 resid_fa=resid)
                for m, g in ((0.1, GR), (0.3, ([0.2, 0.4], [0.5, 1.5], 1.0)), (0.05, GR))]
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        a = G.dev_rhmc_trajectory(4, 0.05, BETA, COEFF, pfs(1e-12), integrator=mq.INT_LEAPFROG, mixed_md=0,
                                  mixed_action=0, action_from_md=False, accept=2, seed=11)
        b = G.dev_rhmc_trajectory(4, 0.05, BETA, COEFF, pfs(1e-10), integrator=mq.INT_LEAPFROG, mixed_md=1,
                                  mixed_action=1, action_from_md=False, accept=2, seed=11)
    print(a["f0"], b["f0"], a["dH"], b["dH"])
    assert a["grsource"] == b["grsource"]
    assert abs(a["f0"] - b["f0"]) < 1e-9 * abs(a["f0"])
    assert abs(a["f1"] - b["f1"]) < 1e-9 * abs(a["f1"])
    assert abs(a["dH"] - b["dH"]) < 1e-6 * max(1.0, abs(a["dH"]))
