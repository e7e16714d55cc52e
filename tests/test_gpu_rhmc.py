"""One RHMC trajectory on the device (milc_hip_dev_rhmc_trajectory, the
update() of ks_imp_rhmc/update_rhmc.c): gauge force + HISQ fermion force,
link updates with the fat and Naik links rebuilt, actions, accept/reject.

* the starting actions against the oracle: beta/3 times the gauge-loop sum of
  the device links, and phi^dag R_MD phi with phi = R_GR chi, chi the
  device's Gaussian source of the trajectory's seed (FP64 solves, 1e-9);
* energy conservation: with the fermion action the MD force's own
  (action_from_md), |dH| falls by about 4 when eps halves at fixed
  trajectory length (leapfrog), and Omelyan's dH is smaller than leapfrog's;
* accept = 2 restores the starting links exactly; accept = 0 keeps links
  that stay unitary.

Workload: 4x4x4x8 lattice, gauge seed 1200 (scale 0.4), m = 0.1, beta = 5,
loop coefficients for u0 = 0.86 with 2+1+1 flavours, a two-pole MD and a
one-pole GR rational (positive poles and residues).
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

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


def traj(G, steps, eps, integrator=0, accept=2, seed=7):
    return G.dev_rhmc_trajectory(steps, eps, BETA, COEFF, MASS, rat(MD), rat(GR), rat(MD), integrator=integrator,
                                 resid_md=1e-12, resid_gr=1e-12, resid_fa=1e-12, mixed_md=0, action_from_md=True,
                                 accept=accept, seed=seed)


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


def test_energy_conservation_and_restore():
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        U0 = G.dev_read_thin_links()
        a = traj(G, 4, 0.05)
        assert np.array_equal(G.dev_read_thin_links(), U0)  # accept = 2 restored the links
        b = traj(G, 8, 0.025)
        o = traj(G, 4, 0.05, integrator=1)
        print("leapfrog dH", a["dH"], b["dH"], "omelyan", o["dH"])
        assert 3.0 < abs(a["dH"]) / abs(b["dH"]) < 5.5
        assert abs(o["dH"]) < abs(a["dH"])
        assert a["fermion_forces"] == 4 and o["fermion_forces"] == 4
        # accept = 0 keeps the evolved links, which stay unitary
        k = traj(G, 4, 0.05, accept=0)
        assert k["accepted"] == 1
        U1 = G.dev_read_thin_links()
        assert not np.array_equal(U1, U0)
        Uc = U1[..., 0] + 1j * U1[..., 1]
        dev = np.abs(np.conj(np.swapaxes(Uc, -1, -2)) @ Uc - np.eye(3)).max()
        assert dev < 1e-12


def test_mixed_md_solves_track_fp64():
    """mixed-precision MD solves (FP32 multi-shift + 16-bit refinement, links
    rebuilt each step) give the FP64-solve trajectory's dH to the solver
    tolerance"""
    with mq.Lattice(DIMS) as G:
        G.dev_gauge_hisq(1200, 0.4)
        a = traj(G, 4, 0.05)
        b = G.dev_rhmc_trajectory(4, 0.05, BETA, COEFF, MASS, rat(MD), rat(GR), rat(MD), integrator=0,
                                  resid_md=1e-10, resid_gr=1e-12, resid_fa=1e-12, mixed_md=1, action_from_md=True,
                                  accept=2, seed=7)
    assert abs(a["dH"] - b["dH"]) < 1e-5 * max(1.0, abs(a["dH"]))
