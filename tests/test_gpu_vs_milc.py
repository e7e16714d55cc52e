"""GPU path against MILC's own compiled CPU code (oracle/_ref), FP64, 8^4.

The north-star bar: FP64 results match the reference CPU path to <= 1e-12
relative on the solution and on the force.  Skipped when the compiled
reference is absent.
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq
import milc_ref as mr

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not mr.available(), reason="oracle/_ref not built")]

DIMS = (8, 8, 8, 8)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def env():
    L = ko.Lattice(DIMS)
    V = mr.init(DIMS)
    U = L.gauge_random(1000)
    mr.set_thin_links(U)
    f_ref, l_ref = mr.hisq_links(V)   # MILC's HISQ links; also the links MILC's solvers use
    G = mq.Lattice(DIMS)
    yield L, V, U, f_ref, l_ref, G
    G.close()


def test_hisq_links_gpu_vs_milc(env):
    L, V, U, f_ref, l_ref, G = env
    gfat, glng, gW, nsvd = G.load_hisq_links(L.rephase(U))
    assert rel(gfat, f_ref) < 1e-13 and rel(glng, l_ref) < 1e-13
    assert nsvd == 0


@pytest.mark.parametrize("parity", [mq.EVEN, mq.ODD])
def test_dslash_gpu_vs_milc(env, parity):
    L, V, U, f_ref, l_ref, G = env
    G.set_links(f_ref, l_ref)
    src = L.gaussian_vector(5, ko.EVENANDODD)
    s = L.parity_slice(parity)
    assert rel(G.dslash_fn_field(src, parity)[s], mr.dslash(src, parity)[s]) < 1e-14


@pytest.mark.parametrize("mass,resid", [(0.5, 1e-14), (0.05, 1e-12)])
def test_cg_gpu_vs_milc(env, mass, resid):
    L, V, U, f_ref, l_ref, G = env
    G.set_links(f_ref, l_ref)
    b = L.gaussian_vector(7, ko.EVEN)
    x_ref, it_ref, q_ref = mr.congrad(b, mass, resid, mr.EVEN, 50, 20)
    x = np.zeros_like(b)
    qic = mq.QuarkInvertControl(max=50, nrestart=20, parity=mq.EVEN, resid=resid)
    it = G.ks_congrad_parity(b, x, qic, mass)
    s = L.parity_slice(ko.EVEN)
    err = rel(x[s], x_ref[s])
    print(f"m={mass}: iters gpu {it} milc {it_ref}; rel {err:.3e}")
    assert abs(it - it_ref) <= 1
    assert err < (1e-12 if mass >= 0.5 else 1e-10)


def test_multishift_gpu_vs_milc(env):
    L, V, U, f_ref, l_ref, G = env
    G.set_links(f_ref, l_ref)
    b = L.gaussian_vector(8, ko.EVEN)
    offs = [4 * 0.5 ** 2 + x for x in (0.0, 0.01, 0.1, 1.0)]
    p_ref, i_ref, _ = mr.multicg(b, offs, 1e-14, mr.EVEN, 2000)
    ksp = [mq.KsParam(offset=o) for o in offs]
    qic = [mq.QuarkInvertControl(max=2000, nrestart=1, parity=mq.EVEN, resid=1e-14) for _ in offs]
    xs = G.ks_multicg_offset_field(b, ksp, qic, mixed=0, prec=2)
    s = L.parity_slice(ko.EVEN)
    assert abs(qic[0].final_iters - i_ref) <= 1
    for j in range(len(offs)):
        assert rel(xs[j][s], p_ref[j][s]) < 1e-12


def test_force_oprod_gpu_vs_milc(env):
    L, V, U, f_ref, l_ref, G = env
    xs = [L.gaussian_vector(100 + t, ko.EVENANDODD) for t in range(4)]
    res = [0.3, -1.2, 0.7, 0.05]
    o1a, o3a = mr.force_oprod(xs, res, V)
    g1, g3 = G.compute_oprod(xs, [(r, r) for r in res])
    assert rel(g1, o1a) < 1e-14 and rel(g3, o3a) < 1e-14
