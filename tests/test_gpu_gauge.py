"""Gauge action, gauge force, link update and momentum action on the GPU
through the C-ABI, against the FP64 oracle (itself pinned to MILC's printed
action and coefficients and to central differences; tests/test_cpu_gauge.py):
force 1e-13, actions 1e-13, update 1e-14 relative; with and without the KS
phases in the links; a pure-gauge leapfrog trajectory on the device's entry
points against the same trajectory in the oracle.

Workload: 6x4x8x10 lattice, thin links from gauge seed 1110 (scale 0.4),
momenta from seed 13, coefficients for u0 = 0.86 and 2+1+1 flavours.
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

DIMS = (6, 4, 8, 10)


def pack_ah(F):
    out = np.zeros(F.shape[:2] + (10,))
    out[..., 0:2] = F[..., 0, 1, :]
    out[..., 2:4] = F[..., 0, 2, :]
    out[..., 4:6] = F[..., 1, 2, :]
    out[..., 6] = F[..., 0, 0, 1]
    out[..., 7] = F[..., 1, 1, 1]
    out[..., 8] = F[..., 2, 2, 1]
    return out


def unpack_ah(p):
    F = np.zeros(p.shape[:2] + (3, 3, 2))
    F[..., 0, 1, :] = p[..., 0:2]
    F[..., 0, 2, :] = p[..., 2:4]
    F[..., 1, 2, :] = p[..., 4:6]
    F[..., 1, 0, 0], F[..., 1, 0, 1] = -p[..., 0], p[..., 1]
    F[..., 2, 0, 0], F[..., 2, 0, 1] = -p[..., 2], p[..., 3]
    F[..., 2, 1, 0], F[..., 2, 1, 1] = -p[..., 4], p[..., 5]
    F[..., 0, 0, 1], F[..., 1, 1, 1], F[..., 2, 2, 1] = p[..., 6], p[..., 7], p[..., 8]
    return F


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def setup():
    L = ko.Lattice(DIMS)
    U = np.ascontiguousarray(L.gauge_random(1110))
    rng = np.random.default_rng(13)
    A = rng.standard_normal((L.V, 4, 3, 3)) + 1j * rng.standard_normal((L.V, 4, 3, 3))
    A = 0.5 * (A - np.conj(np.swapaxes(A, -1, -2)))
    A -= np.eye(3) * np.trace(A, axis1=-2, axis2=-1)[..., None, None] / 3
    P = np.ascontiguousarray(np.stack([A.real, A.imag], -1))
    return L, U, P, ko.gauge_coeffs(0.86, 4, hisq=True)


def test_gauge_force_action_vs_oracle(setup):
    L, U, P, c = setup
    eb3 = 0.037
    ref = P.copy()
    L.gauge_force(c, eb3, U, ref)
    S_ref = L.gauge_action(c, U)
    with mq.Lattice(DIMS) as G:
        for phased in (False, True):
            links = L.rephase(U) if phased else U
            mom = pack_ah(P)
            G.gauge_force(c, eb3, links, mom, phases_in=phased)
            e = rel(mom, pack_ah(ref))
            print(f"gauge force phases_in={phased}: {e:.2e}")
            assert e < 1e-13
            S = G.gauge_action(c, links, phases_in=phased)
            assert abs(S - S_ref) < 1e-13 * abs(S_ref)
        assert abs(G.mom_action(pack_ah(P)) - L.mom_action(P)) < 1e-12 * abs(L.mom_action(P))


def test_update_u_vs_oracle(setup):
    L, U, P, c = setup
    ref = U.copy()
    L.update_u(0.03, ref, P)
    with mq.Lattice(DIMS) as G:
        for phased in (False, True):
            links = np.ascontiguousarray(L.rephase(U) if phased else U.copy())
            G.update_u(0.03, links, pack_ah(P), phases_in=phased)
            want = L.rephase(ref) if phased else ref
            assert rel(links, want) < 1e-14


def test_leapfrog_trajectory_vs_oracle(setup):
    L, U, P, c = setup
    beta, eps, n = 6.0, 0.02, 6
    eb3 = beta / 3.0

    def run(force, update, U0, P0):
        Ux, Px = U0.copy(), P0.copy()
        force(0.5 * eps * eb3, Ux, Px)
        for s in range(n):
            update(eps, Ux, Px)
            force((eps if s < n - 1 else 0.5 * eps) * eb3, Ux, Px)
        return Ux, Px

    U_ref, P_ref = run(lambda e, Ux, Px: L.gauge_force(c, e, Ux, Px), lambda e, Ux, Px: L.update_u(e, Ux, Px),
                       U, P)
    with mq.Lattice(DIMS) as G:
        U_g, P_g = run(lambda e, Ux, Px: G.gauge_force(c, e, Ux, Px), lambda e, Ux, Px: G.update_u(e, Ux, Px),
                       U, pack_ah(P))
        H0 = beta / 3 * G.gauge_action(c, U) + G.mom_action(pack_ah(P))
        H1 = beta / 3 * G.gauge_action(c, U_g) + G.mom_action(P_g)
    assert rel(U_g, U_ref) < 1e-12 and rel(P_g, pack_ah(P_ref)) < 1e-12
    H1_ref = beta / 3 * L.gauge_action(c, U_ref) + L.mom_action(P_ref)
    assert abs((H1 - H0) - (H1_ref - (beta / 3 * L.gauge_action(c, U) + L.mom_action(P)))) < 1e-8
