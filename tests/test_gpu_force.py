"""Fermion force level 0 on the GPU: the hand-written adjoint (outer-product
kernel) wrapped in a torch.autograd.Function, checked against central finite
differences of the action computed by the GPU solver, and the whole pipeline
against the FP64 oracle.

Workload of the finite-difference check: 16^3 x 32, m = 0.05, 9 poles
sigma_i = 1e-3 * 10^(4i/8) (geometric in [1e-3, 10]), residues r_i = 0.1 (i+1),
S = sum_i r_i phi^dag (M^dag M + sigma_i)^-1 phi, eps = 1e-5 in FP64 on 64
seeded link elements (seed 3000) of the fat and Naik links.
"""
import numpy as np
import pytest
import torch

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

M = 0.05
SIGMAS = [1e-3 * 10 ** (4 * i / 8) for i in range(9)]
RES = [0.1 * (i + 1) for i in range(9)]
OFFSETS = [4 * M * M + s for s in SIGMAS]


def grad_from_oprod(O, Vh):
    """dS/d(Re, Im) of a link field from its outer product: 2 eps(x) (O^dag)_ab."""
    Ot = O.transpose(2, 3)  # (V, 4, b, a, 2) -> element [a][b] holds O_ba
    g = torch.empty_like(O)
    g[..., 0] = 2.0 * Ot[..., 0]
    g[..., 1] = -2.0 * Ot[..., 1]
    eps = torch.ones(O.shape[0], dtype=O.dtype)
    eps[Vh:] = -1.0
    return g * eps[:, None, None, None, None]


class HisqLevel0Action(torch.autograd.Function):
    """S(fat, long) evaluated by the GPU multi-shift solver; backward is the
    hand-written adjoint (milc_hip_dev_force_level0 outer products)."""

    @staticmethod
    def forward(ctx, fat, lng, G, slot, resid):
        G.dev_load_links(fat.detach().numpy(), lng.detach().numpy())
        S, O1, O3, _ = G.dev_force_level0(slot, OFFSETS, RES, resid)
        ctx.save_for_backward(torch.from_numpy(O1), torch.from_numpy(O3))
        ctx.Vh = G.Vh
        return torch.tensor(S, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        O1, O3 = ctx.saved_tensors
        return g * grad_from_oprod(O1, ctx.Vh), g * grad_from_oprod(O3, ctx.Vh), None, None, None


@pytest.fixture(scope="module")
def c4():
    dims = (16, 16, 16, 32)
    G = mq.Lattice(dims)
    G.dev_gauge_hisq(1004)
    fat, lng = G.dev_read_links()
    G.dev_source(2400, mq.EVEN, 0)
    yield G, fat, lng
    G.close()


def test_force_adjoint_vs_central_finite_differences(c4):
    G, fat, lng = c4
    fat_t = torch.tensor(fat, requires_grad=True)
    lng_t = torch.tensor(lng, requires_grad=True)
    S = HisqLevel0Action.apply(fat_t, lng_t, G, 0, 1e-14)
    S.backward()
    gF, gL = fat_t.grad.numpy(), lng_t.grad.numpy()
    rng = np.random.default_rng(3000)
    h = 1e-5
    worst = 0.0
    for k in range(64):
        which = k % 2
        i = int(rng.integers(G.V))
        mu, a, b, part = (int(v) for v in rng.integers(0, [4, 3, 3, 2]))
        base = fat if which == 0 else lng
        vals = []
        for sgn in (+1, -1):
            pert = base.copy()
            pert[i, mu, a, b, part] += sgn * h
            if which == 0:
                G.dev_load_links(pert, lng)
            else:
                G.dev_load_links(fat, pert)
            s, _, _, _ = G.dev_force_level0(0, OFFSETS, RES, 1e-14, want_oprod=False)
            vals.append(s)
        fd = (vals[0] - vals[1]) / (2 * h)
        an = (gF if which == 0 else gL)[i, mu, a, b, part]
        err = abs(fd - an) / max(1.0, abs(an))
        worst = max(worst, err)
        assert err < 1e-5, (k, which, i, mu, a, b, part, fd, an)
    print(f"S = {S.item():.10e}; worst relative FD error over 64 elements {worst:.2e}")


def test_force_mixed_precision_error_vs_fp64(c4):
    G, fat, lng = c4
    G.dev_load_links(fat, lng)
    S64, A1, A3, it64 = G.dev_force_level0(0, OFFSETS, RES, 1e-12, mixed=0)
    Smx, B1, B3, itmx = G.dev_force_level0(0, OFFSETS, RES, 1e-10, mixed=1)
    e1 = np.linalg.norm(B1 - A1) / np.linalg.norm(A1)
    e3 = np.linalg.norm(B3 - A3) / np.linalg.norm(A3)
    print(f"mixed force: rel err O1 {e1:.2e} O3 {e3:.2e}; S rel err {abs(Smx - S64) / abs(S64):.2e}; "
          f"iters fp64 {it64} mixed {itmx}")
    assert e1 < 1e-8 and e3 < 1e-8
    assert abs(Smx - S64) < 1e-9 * abs(S64)


def test_force_pipeline_vs_oracle():
    """multi-shift -> D_oe -> outer products on 8^4, GPU against the oracle."""
    dims = (8, 8, 8, 8)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1000))
    fat, lng, _ = L.hisq_links(U)
    phi = L.gaussian_vector(2400, ko.EVEN)
    xs, _, _ = L.multicg_offset(fat, lng, phi, OFFSETS, 1e-14, parity=ko.EVEN, max_iters=5000, nrestart=1)
    Z = []
    for x in xs:
        z = x.copy()
        z[L.parity_slice(ko.ODD)] = L.dslash(fat, lng, x, ko.ODD)[L.parity_slice(ko.ODD)]
        Z.append(z)
    O1, O3 = L.force_oprod(Z, RES)
    E = L.parity_slice(ko.EVEN)
    S_ref = sum(r * np.vdot(phi[E].ravel(), x[E].ravel()) for r, x in zip(RES, xs))
    with mq.Lattice(dims) as G:
        G.dev_load_links(fat, lng)
        G.dev_write(0, mq.EVEN, phi)
        S, G1, G3, _ = G.dev_force_level0(0, OFFSETS, RES, 1e-14)
    assert abs(S - S_ref) < 1e-12 * abs(S_ref)
    assert np.linalg.norm(G1 - O1) / np.linalg.norm(O1) < 1e-11
    assert np.linalg.norm(G3 - O3) / np.linalg.norm(O3) < 1e-11


def test_rhmc_action_vs_oracle():
    """fermion_action: sum_even |a0 phi + sum_k a_k x_k|^2 against the oracle."""
    dims = (8, 8, 8, 8)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1001))
    fat, lng, _ = L.hisq_links(U)
    phi = L.gaussian_vector(2500, ko.EVEN)
    roots = [0.0] + OFFSETS[:6]
    resid_ = [0.7] + RES[:6]
    xs, _, _ = L.multicg_offset(fat, lng, phi, roots[1:], 1e-14, parity=ko.EVEN, max_iters=5000, nrestart=1)
    E = L.parity_slice(ko.EVEN)
    s = resid_[0] * phi
    for a, x in zip(resid_[1:], xs):
        s = s + a * x
    ref = float(np.sum(s[E] ** 2))
    with mq.Lattice(dims) as G:
        G.dev_load_links(fat, lng)
        G.dev_write(0, mq.EVEN, phi)
        act, it = G.dev_rhmc_action(0, roots, resid_, 1e-14, mixed=0)
        act_mx, it2 = G.dev_rhmc_action(0, roots, resid_, 1e-10, mixed=1)
    assert abs(act - ref) < 1e-12 * ref
    assert abs(act_mx - ref) < 1e-9 * ref
