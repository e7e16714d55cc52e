"""Many right-hand sides on the GPU: the batched stencil (every VALU and MFMA
variant the autotuner can pick), the batched multi-shift solver with deferred
shift updates and per-shift refinement, and the block CG, all against the FP64
oracle through the C-ABI.

Workloads: 8x6x4x10 lattice (non-cubic, exercises every direction's wrap) and
8^4, HISQ links from gauge seeds 1005/1006, Gaussian sources from seeds 700+.
Tolerances: FP64 stencil 1e-13 relative, FP32 stencil 2e-6 relative, 16-bit
stencil 2e-4 relative; solvers
to their stated true-residual targets, solutions within 1e-7 relative of the
oracle's 1e-14 solves.
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def hisq_setup(dims, seed):
    L = ko.Lattice(dims)
    U = L.rephase(L.gauge_random(seed))
    fat, lng, _ = L.hisq_links(U)
    return L, fat, lng


@pytest.mark.parametrize("prec,nrhs", [(2, 4), (2, 8), (2, 16), (1, 4), (1, 8), (1, 16), (3, 4), (3, 8), (3, 16)])
def test_block_dslash_every_variant_vs_oracle(prec, nrhs):
    """prec 2 FP64, 1 FP32, 3 the 16-bit records and links of the batched CG's
    sloppy operator (input and output records rounded to 16 bits: 2e-4)."""
    dims = (8, 6, 4, 10)
    L, fat, lng = hisq_setup(dims, 1005)
    tol = {2: 1e-13, 1: 2e-6, 3: 2e-4}[prec]
    with mq.Lattice(dims) as G:
        G.dev_load_links(fat, lng)
        nvar, chosen, us, desc = G.dev_block_tune(prec, nrhs)
        print(f"prec {prec} nrhs {nrhs}: {nvar} variants, tuned -> {chosen} ({desc})")
        for out_par, in_par in ((ko.EVEN, ko.ODD), (ko.ODD, ko.EVEN)):
            srcs = [L.gaussian_vector(700 + s, in_par) for s in range(nrhs)]
            refs = [L.dslash(fat, lng, v, out_par) for v in srcs]
            for s, v in enumerate(srcs):
                G.dev_write(s, in_par, v)
            sl = L.parity_slice(out_par)
            for var in range(nvar):
                G.dev_block_dslash(prec, nrhs, var, list(range(nrhs)), out_par, list(range(20, 20 + nrhs)))
                for s in range(nrhs):
                    got = G.dev_read(20 + s, out_par)
                    e = rel(got[sl], refs[s][sl])
                    assert e < tol, (prec, nrhs, var, out_par, s, e)


@pytest.mark.parametrize("prec,nrhs", [(2, 4), (1, 4), (1, 16), (3, 4), (3, 16)])
def test_block_dslash_site_orders_vs_oracle(prec, nrhs):
    """Every variant on a lattice where every blocked site order applies
    (16x16x24x4: blocks of 8x8x8x4, 8x4x4x4 and 4x4x4x4 sites)."""
    dims = (16, 16, 24, 4)
    L, fat, lng = hisq_setup(dims, 1007)
    tol = {2: 1e-13, 1: 2e-6, 3: 2e-4}[prec]
    with mq.Lattice(dims) as G:
        G.dev_load_links(fat, lng)
        nvar, chosen, us, desc = G.dev_block_tune(prec, nrhs)
        descs = [G.dev_block_describe(prec, nrhs, v) for v in range(nvar)]
        print(f"prec {prec} nrhs {nrhs}: {nvar} variants, tuned -> {chosen} ({desc})")
        assert all(any(b in d for d in descs) for b in ("8x8x8x4", "8x4x4x4", "4x4x4x4"))
        srcs = [L.gaussian_vector(760 + s, ko.ODD) for s in range(nrhs)]
        refs = [L.dslash(fat, lng, v, ko.EVEN) for v in srcs]
        for s, v in enumerate(srcs):
            G.dev_write(s, ko.ODD, v)
        sl = L.parity_slice(ko.EVEN)
        for var in range(nvar):
            G.dev_block_dslash(prec, nrhs, var, list(range(nrhs)), ko.EVEN, list(range(20, 20 + nrhs)))
            for s in range(nrhs):
                e = rel(G.dev_read(20 + s, ko.EVEN)[sl], refs[s][sl])
                assert e < tol, (prec, nrhs, var, descs[var], s, e)


def test_block_dslash_16bit_nine_element_naik_vs_oracle():
    """The 16-bit batched stencil with the Naik links in the nine-element form
    (two-row form switched off), 16 rhs, every variant: 2e-4 of the oracle."""
    dims = (8, 6, 4, 10)
    L, fat, lng = hisq_setup(dims, 1005)
    with mq.Lattice(dims) as G:
        G.set_link_compression(0)  # process-wide setting: restored below
        try:
            G.dev_load_links(fat, lng)
            nvar, chosen, us, desc = G.dev_block_tune(3, 16)
            assert G.link_compression(2) == 0
            srcs = [L.gaussian_vector(800 + s, ko.ODD) for s in range(16)]
            refs = [L.dslash(fat, lng, v, ko.EVEN) for v in srcs]
            for s, v in enumerate(srcs):
                G.dev_write(s, ko.ODD, v)
            sl = L.parity_slice(ko.EVEN)
            for var in range(nvar):
                G.dev_block_dslash(3, 16, var, list(range(16)), ko.EVEN, list(range(20, 36)))
                for s in range(16):
                    e = rel(G.dev_read(20 + s, ko.EVEN)[sl], refs[s][sl])
                    assert e < 2e-4, (var, s, e)
        finally:
            G.set_link_compression(1)


@pytest.fixture(scope="module")
def l8():
    dims = (8, 8, 8, 8)
    L, fat, lng = hisq_setup(dims, 1006)
    return L, fat, lng


OFFS = [4 * 0.05 ** 2 + s for s in (1e-3, 1e-2, 0.1, 1.0)]


def oracle_multishift(L, fat, lng, srcs, parity):
    out = []
    for v in srcs:
        xs, _, _ = L.multicg_offset(fat, lng, v, OFFS, 1e-14, parity=parity, max_iters=5000, nrestart=1)
        out.append(xs)
    return out


@pytest.mark.parametrize("parity", [ko.EVEN, ko.ODD])
def test_multishift_block_refined_vs_oracle(l8, parity):
    """3 sources (padded to a batch of 4), 4 shifts; FP32 multi-shift then
    per-shift refinement to a 1e-10 FP64 true residual."""
    L, fat, lng = l8
    srcs = [L.gaussian_vector(800 + s, parity) for s in range(3)]
    ref = oracle_multishift(L, fat, lng, srcs, parity)
    sl = L.parity_slice(parity)
    with mq.Lattice(L.dims) as G:
        G.dev_load_links(fat, lng)
        for s, v in enumerate(srcs):
            G.dev_write(s, parity, v)
        st = G.dev_multishift_block([0, 1, 2], OFFS, 1e-5, 1e-10, parity=parity, out_slot0=8)
        print(st)
        assert st["converged"] == 1
        assert max(st["true_rel"]) < 1e-10
        for s in range(3):
            for j in range(len(OFFS)):
                got = G.dev_read(8 + s * len(OFFS) + j, parity)
                assert rel(got[sl], ref[s][j][sl]) < 1e-7, (s, j)


def test_multishift_block_fp32_stage_vs_oracle(l8):
    """The FP32 multi-shift stage alone (no refinement): deferred-window updates
    reproduce the FP64 solutions to FP32 accuracy, 16 sources in one batch."""
    L, fat, lng = l8
    srcs = [L.gaussian_vector(900 + s, ko.EVEN) for s in range(16)]
    ref = oracle_multishift(L, fat, lng, srcs[:2], ko.EVEN)
    sl = L.parity_slice(ko.EVEN)
    with mq.Lattice(L.dims) as G:
        G.dev_load_links(fat, lng)
        for s, v in enumerate(srcs):
            G.dev_write(s, ko.EVEN, v)
        st = G.dev_multishift_block(list(range(16)), OFFS, 1e-6, 0.0, parity=ko.EVEN, out_slot0=16, out_nsrc=2)
        assert st["converged"] == 1
        # the oracle's base-shift iteration count on the same source
        _, it_ref, _ = L.multicg_offset(fat, lng, srcs[0], OFFS, 1e-6, parity=ko.EVEN, max_iters=5000, nrestart=1)
        assert abs(st["ms_iters"] - it_ref) <= 8 + 2, (st["ms_iters"], it_ref)
        for s in range(2):
            for j in range(len(OFFS)):
                got = G.dev_read(16 + s * len(OFFS) + j, ko.EVEN)
                assert rel(got[sl], ref[s][j][sl]) < 5e-5, (s, j)
        # a second identical solve gives bit-identical results (deterministic reductions)
        a = G.dev_read(16, ko.EVEN)
        G.dev_multishift_block(list(range(16)), OFFS, 1e-6, 0.0, parity=ko.EVEN, out_slot0=16, out_nsrc=2)
        assert np.array_equal(a, G.dev_read(16, ko.EVEN))


def test_multishift_block_seven_shifts_refined_from_1e4_vs_oracle(l8):
    """2 sources (a batch of 4) and 7 shifts, the FP32 stage stopped at 1e-4 so
    that every shift's refinement runs sloppy phases and reliable updates down to
    a 1e-10 FP64 true residual."""
    L, fat, lng = l8
    offs = [4 * 0.05 ** 2 + s for s in (3e-4, 1e-3, 3e-3, 1e-2, 3e-2, 0.1, 1.0)]
    srcs = [L.gaussian_vector(1200 + s, ko.EVEN) for s in range(2)]
    ref = []
    for v in srcs:
        xs, _, _ = L.multicg_offset(fat, lng, v, offs, 1e-14, parity=ko.EVEN, max_iters=5000, nrestart=1)
        ref.append(xs)
    sl = L.parity_slice(ko.EVEN)
    with mq.Lattice(L.dims) as G:
        G.dev_load_links(fat, lng)
        for s, v in enumerate(srcs):
            G.dev_write(s, ko.EVEN, v)
        st = G.dev_multishift_block([0, 1], offs, 1e-4, 1e-10, parity=ko.EVEN, out_slot0=8)
        print(st)
        assert st["converged"] == 1
        assert max(st["true_rel"]) < 1e-10
        for s in range(2):
            for j in range(len(offs)):
                got = G.dev_read(8 + s * len(offs) + j, ko.EVEN)
                assert rel(got[sl], ref[s][j][sl]) < 1e-7, (s, j)


def test_multishift_invert_block_host_abi(l8):
    """milc_hip_multishift_invert_block with MILC-layout host arrays."""
    L, fat, lng = l8
    srcs = [L.gaussian_vector(1100 + s, ko.ODD) for s in range(5)]
    ref = oracle_multishift(L, fat, lng, srcs, ko.ODD)
    sl = L.parity_slice(ko.ODD)
    with mq.Lattice(L.dims) as G:
        G.set_links(fat, lng)
        psim, res, it = G.ks_multicg_offset_block(srcs, OFFS, ko.ODD, 1e-5, 1e-10)
    for s in range(5):
        for j in range(len(OFFS)):
            assert res[s][j] < 1e-10
            assert rel(psim[s][j][sl], ref[s][j][sl]) < 1e-7
            other = psim[s][j][L.parity_slice(ko.EVEN)]
            assert not other.any()  # only the solve parity is written


def test_invert_block_vs_oracle(l8):
    """ks_congrad_block_parity form: 6 sources, one with a nonzero initial guess."""
    L, fat, lng = l8
    m = 0.05
    srcs = [L.gaussian_vector(1200 + s, ko.EVEN) for s in range(6)]
    dests = [L.new_vector() for _ in range(6)]
    dests[2][:] = 0.3 * L.gaussian_vector(1300, ko.EVEN)
    sl = L.parity_slice(ko.EVEN)
    with mq.Lattice(L.dims) as G:
        G.set_links(fat, lng)
        q = mq.QuarkInvertControl(max=2000, nrestart=5, parity=ko.EVEN, resid=1e-10)
        it, fres = G.ks_congrad_block_parity(srcs, dests, q, m)
    assert max(fres) < 1e-10 and q.converged == 1
    for s in range(6):
        x, _, _ = L.congrad(fat, lng, srcs[s], m, 1e-14, parity=ko.EVEN, max_iters=5000, nrestart=5)
        assert rel(dests[s][sl], x[sl]) < 1e-8, s


def test_batched_solvers_stop_at_max_iters(l8):
    """The iteration cap holds inside the captured windows: the batched
    multi-shift stops at max_iters (ks_multicg_offset.c counts the set-up as
    one), the batched CG at max_iters plus the closing true-residual
    evaluation, which MILC also counts (d_congrad5_fn_milc.c:223-234); both
    report no convergence.  Caps that are not multiples of the 8-iteration
    window."""
    L, fat, lng = l8
    srcs = [L.gaussian_vector(1400 + s, ko.EVEN) for s in range(4)]
    with mq.Lattice(L.dims) as G:
        G.dev_load_links(fat, lng)
        for s, v in enumerate(srcs):
            G.dev_write(s, ko.EVEN, v)
        for cap in (13, 30):
            st = G.dev_multishift_block(list(range(4)), OFFS, 1e-20, 0.0, parity=ko.EVEN, max_iters=cap)
            assert st["ms_iters"] == cap and st["converged"] == 0, (cap, st["ms_iters"])
        G.set_links(fat, lng)
        for cap in (11, 21):
            dests = [L.new_vector() for _ in range(4)]
            q = mq.QuarkInvertControl(max=cap, nrestart=1, parity=ko.EVEN, resid=1e-14)
            it, fres = G.ks_congrad_block_parity(srcs, dests, q, 0.05)
            assert it <= cap + 1 and q.converged == 0, (cap, it)
            assert max(fres) > 1e-14
