"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle.

Every test builds its inputs with the oracle (seeded), runs the same call
through libmilc_hip.so and compares.  Tolerances are stated per precision:
FP64 results to <= 1e-12 relative where the comparison is well conditioned,
FP32 to 1e-5, 16-bit storage to 2e-3.
"""
import numpy as np
import pytest

import ks_oracle as ko
import milc_qcd_amd as mq

pytestmark = pytest.mark.gpu

DIMS_SMALL = (4, 4, 4, 8)
DIMS_8 = (8, 8, 8, 8)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def hisq8():
    L = ko.Lattice(DIMS_8)
    U = L.rephase(L.gauge_random(1000))
    fat, lng, nsvd, W = L.hisq_links(U, want_W=True)
    return L, U, fat, lng, W, nsvd


@pytest.fixture(scope="module")
def gpu8(hisq8):
    L, U, fat, lng, W, _ = hisq8
    G = mq.Lattice(DIMS_8)
    G.set_links(fat, lng)
    yield G
    G.close()


def test_library_exports():
    lib = mq.load_library()
    for name in mq.EXPORTS:
        assert hasattr(lib, name)


@pytest.mark.parametrize("parity", [mq.EVEN, mq.ODD])
def test_dslash_f64(hisq8, gpu8, parity):
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(77, ko.EVENANDODD)
    ref = L.dslash(fat, lng, src, parity)
    out = gpu8.dslash_fn_field(src, parity, prec=2)
    s = L.parity_slice(parity)
    print(f"dslash f64 parity {parity}: rel {rel(out[s], ref[s]):.3e}")
    assert rel(out[s], ref[s]) < 1e-14
    # other parity untouched (zero in the returned field)
    o = L.parity_slice(ko.ODD if parity == ko.EVEN else ko.EVEN)
    assert np.all(out[o] == 0)


def test_dslash_f32(hisq8, gpu8):
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(78, ko.EVENANDODD)
    ref = L.dslash(fat, lng, src, ko.EVEN)
    out = gpu8.dslash_fn_field(src, mq.EVEN, prec=1)
    s = L.parity_slice(ko.EVEN)
    assert rel(out[s], ref[s]) < 1e-5


def test_cg_f64_matches_oracle_iteration_for_iteration(hisq8, gpu8):
    """Same restart cadence, same iteration count, solution to 1e-12 (m = 0.5, well conditioned)."""
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(2000, ko.EVEN)
    ref, it_ref, q_ref = L.congrad(fat, lng, src, 0.5, 1e-14, parity=ko.EVEN, max_iters=40, nrestart=10)
    x = np.zeros_like(src)
    qic = mq.QuarkInvertControl(max=40, nrestart=10, parity=mq.EVEN, resid=1e-14)
    it = gpu8.ks_congrad_parity(src, x, qic, 0.5)
    s = L.parity_slice(ko.EVEN)
    print(f"cg f64: iters gpu {it} oracle {it_ref}; rel {rel(x[s], ref[s]):.3e}")
    assert abs(it - it_ref) <= 1
    assert rel(x[s], ref[s]) < 1e-12
    assert np.all(x[L.parity_slice(ko.ODD)] == 0)


def test_cg_f64_light_mass(hisq8, gpu8):
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(2001, ko.ODD)
    ref, it_ref, q_ref = L.congrad(fat, lng, src, 0.05, 1e-12, parity=ko.ODD, max_iters=300, nrestart=5)
    x = np.zeros_like(src)
    qic = mq.QuarkInvertControl(max=300, nrestart=5, parity=mq.ODD, resid=1e-12)
    it = gpu8.ks_congrad_parity(src, x, qic, 0.05)
    s = L.parity_slice(ko.ODD)
    # FP64 true residual recomputed by the oracle
    r = src - L.mdagm(fat, lng, 4 * 0.05 ** 2, x, ko.ODD)
    tr = np.sqrt(L.norm2(r, ko.ODD) / L.norm2(src, ko.ODD))
    print(f"cg f64 m=0.05: iters {it}/{it_ref} true resid {tr:.3e} rel {rel(x[s], ref[s]):.3e}")
    assert abs(it - it_ref) <= 2
    assert tr < 1e-12
    assert rel(x[s], ref[s]) < 1e-9


@pytest.mark.parametrize("mixed,tol", [(1, 1e-10), (2, 1e-10)])
def test_cg_mixed_reaches_fp64_residual(hisq8, gpu8, mixed, tol):
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(2002, ko.EVEN)
    ref, it_ref, _ = L.congrad(fat, lng, src, 0.05, 1e-13, parity=ko.EVEN, max_iters=1000, nrestart=5)
    x = np.zeros_like(src)
    qic = mq.QuarkInvertControl(max=4000, nrestart=1, parity=mq.EVEN, resid=tol, mixed=mixed)
    it = gpu8.ks_congrad_parity(src, x, qic, 0.05)
    r = src - L.mdagm(fat, lng, 4 * 0.05 ** 2, x, ko.EVEN)
    tr = np.sqrt(L.norm2(r, ko.EVEN) / L.norm2(src, ko.EVEN))
    s = L.parity_slice(ko.EVEN)
    print(f"mixed={mixed}: iters {it} (fp64 {it_ref}); true resid {tr:.3e}; err vs oracle {rel(x[s], ref[s]):.3e}")
    assert qic.converged
    assert tr < tol * 1.01
    assert rel(x[s], ref[s]) < 1e-7


def test_hisq_links_gpu_vs_oracle(hisq8, gpu8):
    L, U, fat, lng, W, nsvd = hisq8
    gfat, glng, gW, gsvd = gpu8.load_hisq_links(U)
    print(f"hisq links: fat {rel(gfat, fat):.3e} lng {rel(glng, lng):.3e} W {rel(gW, W):.3e} svd {gsvd}/{nsvd}")
    assert rel(gW, W) < 1e-13
    assert rel(gfat, fat) < 1e-13
    assert rel(glng, lng) < 1e-13
    assert gsvd == nsvd


def test_device_gauge_generator_matches_oracle():
    L = ko.Lattice(DIMS_SMALL)
    U = L.rephase(L.gauge_random(1234))
    fat, lng, _ = L.hisq_links(U)
    with mq.Lattice(DIMS_SMALL) as G:
        G.dev_gauge_hisq(1234)
        gfat, glng = G.dev_read_links()
    assert rel(gfat, fat) < 1e-12
    assert rel(glng, lng) < 1e-12


def test_device_source_matches_oracle():
    L = ko.Lattice(DIMS_SMALL)
    with mq.Lattice(DIMS_SMALL) as G:
        G.dev_source(99, mq.ODD, 0)
        v = G.dev_read(0, mq.ODD)
    ref = L.gaussian_vector(99, ko.ODD)
    s = L.parity_slice(ko.ODD)
    assert rel(v[s], ref[s]) < 1e-14


def test_multishift_f64_matches_oracle(hisq8, gpu8):
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(3000, ko.EVEN)
    m = 0.05
    offsets = [4 * m * m + s for s in (1e-3, 1e-2, 1e-1, 1.0)]
    ref, it_ref, _ = L.multicg_offset(fat, lng, src, offsets, 1e-12, parity=ko.EVEN, max_iters=2000, nrestart=1)
    ksp = [mq.KsParam(offset=o) for o in offsets]
    qic = [mq.QuarkInvertControl(max=2000, nrestart=1, parity=mq.EVEN, resid=1e-12) for _ in offsets]
    xs = gpu8.ks_multicg_offset_field(src, ksp, qic, mixed=0, prec=2)
    s = L.parity_slice(ko.EVEN)
    for j, o in enumerate(offsets):
        e = rel(xs[j][s], ref[j][s])
        print(f"shift {o:.4g}: rel {e:.3e}")
        assert e < 1e-9
    assert abs(qic[0].final_iters - it_ref) <= 2


@pytest.mark.parametrize("mixed,parity", [(1, ko.EVEN), (2, ko.EVEN), (2, ko.ODD)])
def test_multishift_mixed_per_shift_residuals(hisq8, gpu8, mixed, parity):
    """FP32 bulk, then every shift refined to its own FP64 target with an FP32
    (mixed 1) or 16-bit (mixed 2) sloppy operator; six shifts are refined as a
    batched group of four and one of two (two unused slots)."""
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(3001, parity)
    m = 0.02
    offsets = [4 * m * m + s for s in np.geomspace(1e-4, 1.0, 6)]
    targets = [1e-9, 1e-9, 3e-9, 1e-8, 1e-8, 3e-8]
    ksp = [mq.KsParam(offset=o) for o in offsets]
    qic = [mq.QuarkInvertControl(max=3000, nrestart=1, parity=parity, resid=t) for t in targets]
    xs = gpu8.ks_multicg_offset_field(src, ksp, qic, mixed=mixed, prec=1)
    bn = np.sqrt(L.norm2(src, parity))
    for j, o in enumerate(offsets):
        r = src - L.mdagm(fat, lng, o, xs[j], parity)
        tr = np.sqrt(L.norm2(r, parity)) / bn
        print(f"shift {o:.4g}: true resid {tr:.3e} (target {targets[j]:g}, reported {qic[j].final_rsq ** 0.5:.3e})")
        assert tr < 1.01 * targets[j]
        assert qic[j].converged == 1
        assert abs(qic[j].final_rsq ** 0.5 - tr) <= 1e-3 * tr


@pytest.mark.parametrize("nshift", [1, 5, 9])
def test_multishift_mixed_group_edges(hisq8, gpu8, nshift):
    """Shift counts at the edges of the grouped refinement: one shift (the
    one-source mixed CG alone), five (a group of four and a lone shift) and
    nine (two groups and a lone shift); every shift to its FP64 target."""
    L, U, fat, lng, W, _ = hisq8
    src = L.gaussian_vector(3100 + nshift, ko.EVEN)
    m = 0.03
    offsets = [4 * m * m + s for s in np.geomspace(3e-4, 0.5, nshift)] if nshift > 1 else [4 * m * m + 1e-3]
    ksp = [mq.KsParam(offset=o) for o in offsets]
    qic = [mq.QuarkInvertControl(max=3000, nrestart=1, parity=ko.EVEN, resid=1e-9) for _ in offsets]
    xs = gpu8.ks_multicg_offset_field(src, ksp, qic, mixed=2, prec=1)
    bn = np.sqrt(L.norm2(src, ko.EVEN))
    for j, o in enumerate(offsets):
        tr = np.sqrt(L.norm2(src - L.mdagm(fat, lng, o, xs[j], ko.EVEN), ko.EVEN)) / bn
        assert tr < 1.01e-9, (nshift, j, tr)
        assert qic[j].converged == 1


def test_oprod_matches_oracle(hisq8, gpu8):
    L, U, fat, lng, W, _ = hisq8
    xs = [L.gaussian_vector(4000 + t, ko.EVENANDODD) for t in range(3)]
    res = [0.3, -1.2, 0.7]
    o1, o3 = L.force_oprod(xs, res)
    g1, g3 = gpu8.compute_oprod(xs, [(r, r) for r in res])
    assert rel(g1, o1) < 1e-14
    assert rel(g3, o3) < 1e-14
    # independent one-hop / three-hop coefficients
    g1b, g3b = gpu8.compute_oprod(xs, [(r, 2 * r) for r in res])
    assert rel(g1b, o1) < 1e-14 and rel(g3b, 2 * o3) < 1e-14


def test_golden_asqtad_mdaggerm_inverse_on_gpu(oracle):
    """MILC's own check_ks_invert golden: (M^dag M)^-1 src, asqtad u0=0.502, m=0.1, 8^4."""
    import os
    gdir = os.path.join(os.path.dirname(__file__), "golden")
    L = ko.Lattice(DIMS_8)
    lex = L.lex_of_site()
    U = np.load(os.path.join(gdir, "lat_l8.npy")).astype(np.float64)[lex].copy()
    fat = np.load(os.path.join(gdir, "fat_l8.npy")).astype(np.float64)[lex].copy()
    src = np.load(os.path.join(gdir, "ksvector_src_l8.npy")).astype(np.float64)[lex].copy()
    ans = np.load(os.path.join(gdir, "ksvector_ans_MdaggerMinv_l8.npy")).astype(np.float64)[lex].copy()
    _, lng = L.asqtad_links(L.rephase(U), 0.502)
    with mq.Lattice(DIMS_8) as G:
        G.set_links(fat, lng)
        for par in (mq.EVEN, mq.ODD):
            x = np.zeros_like(src)
            qic = mq.QuarkInvertControl(max=300, nrestart=5, parity=par, resid=1e-9)
            G.ks_congrad_parity(src, x, qic, 0.1)
            s = L.parity_slice(par)
            e = rel(x[s], ans[s])
            print(f"golden parity {par}: rel {e:.3e}")
            assert e < 1e-6  # the golden answer is stored in FP32


def test_cg_iteration_bench_times_every_kernel(hisq8):
    """milc_hip_dev_cg_iteration_bench: the 16-bit CG iteration's four kernels
    timed in their own sequence, and a whole iteration without events; all
    positive, and the four parts (each with its event gap) account for at
    least the time of an iteration without events, give or take launch jitter."""
    L, U, fat, lng, W, _ = hisq8
    with mq.Lattice(DIMS_8) as G:
        G.dev_load_links(fat, lng)
        G.dev_write(0, ko.EVEN, L.gaussian_vector(91, ko.EVEN))
        t = G.dev_cg_iteration_bench(0, 0.05, mq.EVEN, 20)
        print(t)
        assert all(v > 0 for v in t.values())
        parts = t["dslash"] + t["dslash_mdagm"] + t["cg_scalars"] + t["cg_update"]
        assert parts > 0.5 * t["iteration"]
        with pytest.raises(mq.MilcHipError):
            G.dev_cg_iteration_bench(0, 0.05, mq.EVEN, 0)
