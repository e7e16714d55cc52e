"""The larger workloads at their real sizes, checked through properties that
do not need a CPU solve at that size, plus the same entry points against the
oracle on a small lattice.

* 24^3 x 48, m = 0.02, 12 shifts sigma_i = 1e-4 10^(4i/11), 16 sources
  (seeds 2200+j, gauge 1002): batched multi-shift + refinement; every shift's
  FP64 true residual below 1e-9 and source 0 within 1e-6 of an FP64 solve.
* 48^3 x 96, m = 0.01, same shifts, residues 0.1 (i+1), source seed 2400,
  gauge 1004: one RHMC force (mixed) against the FP64 force on the device,
  relative error below 1e-7 for the outer products and for the HISQ
  thin-link momentum; the action (mixed solve, FP64 sum) within 1e-9 of the
  action from an FP64 solve.
* 8^4: milc_hip_dev_hmc_step's sum_t r_t <phi, X_t> and action against the oracle.
"""
import math

import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

SIGMAS = [1e-4 * 10 ** (4 * i / 11) for i in range(12)]


def test_batched_multishift_24x24x24x48_16_sources():
    m = 0.02
    offs = [4 * m * m + s for s in SIGMAS]
    with mq.Lattice((24, 24, 24, 48)) as G:
        G.dev_gauge_hisq(1002, 0.4)
        for j in range(16):
            G.dev_source(2200 + j, mq.EVEN, j)
        st = G.dev_multishift_block(list(range(16)), offs, 5e-6, 1e-9, out_slot0=16, out_nsrc=1)
        print(st)
        assert st["converged"] == 1
        assert max(st["true_rel"]) < 1e-9
        G.dev_multishift(0, offs, 1e-12, 0, out_slot0=40)
        errs = [G.dev_diff(16 + j, 40 + j) for j in range(len(offs))]
        print("rel err vs FP64 per shift", errs)
        assert max(errs) < 1e-6


def test_hmc_step_48x48x48x96_mixed_vs_fp64_force():
    m = 0.01
    offs = [4 * m * m + s for s in SIGMAS]
    res = [0.1 * (i + 1) for i in range(12)]
    with mq.Lattice((48, 48, 48, 96)) as G:
        G.dev_gauge_hisq(1004, 0.4)
        G.dev_source(2400, mq.EVEN, 0)
        st = G.dev_hmc_step(0, offs, res, 1e-9, mixed=1, compare_fp64=True, thin_force=True)
        print(st)
        assert st["rel_err_O1"] < 1e-7 and st["rel_err_O3"] < 1e-7
        assert 0 <= st["rel_err_mom"] < 1e-7
        assert math.isfinite(st["action"]) and st["action"] > 0
        assert 0 <= st["rel_err_action"] < 1e-9
        assert st["S"] > 0


def test_hmc_step_vs_oracle_8x8x8x8():
    dims = (8, 8, 8, 8)
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(1010))
    fat, lng, _ = L.hisq_links(U)
    phi = L.gaussian_vector(2500, ko.EVEN)
    m = 0.05
    offs = [4 * m * m + s for s in SIGMAS[::3]]
    res = [0.3, 0.2, 0.5, 0.7]
    xs, _, _ = L.multicg_offset(fat, lng, phi, offs, 1e-14, parity=ko.EVEN, max_iters=5000, nrestart=1)
    E = L.parity_slice(ko.EVEN)
    S_ref = sum(r * float(np.sum(phi[E] * x[E])) for r, x in zip(res, xs))
    tot = phi.copy()
    for r, x in zip(res, xs):
        tot = tot + r * x
    act_ref = float(np.sum(tot[E] ** 2))
    with mq.Lattice(dims) as G:
        G.dev_load_links(fat, lng)
        G.dev_write(0, ko.EVEN, phi)
        st = G.dev_hmc_step(0, offs, res, 1e-11, mixed=1, compare_fp64=True)
    assert abs(st["S"] - S_ref) < 1e-9 * abs(S_ref)
    assert abs(st["action"] - act_ref) < 1e-10 * act_ref
    assert abs(st["action_fp64"] - act_ref) < 1e-10 * act_ref
    assert st["rel_err_O1"] < 1e-8 and st["rel_err_O3"] < 1e-8
