"""The MILC glue harness on the CPU (oracle/Makefile.glue).

* the -DUSE_CG_GPU builds take ks_congrad_parity_gpu, ks_congrad_block_parity_gpu,
  ks_multicg_offset_field_gpu, get_fn_last / set_fn_last and
  eo_fermion_force_multi from the glue of integration/ and reach the library
  through milc_hip_invert, milc_hip_invert_block, milc_hip_multishift_invert
  and milc_hip_hisq_force (undefined in the harness, defined in libmilc_hip.so);
* MILC's CPU builds of the harness in both precisions solve the same system
  as the FP64 oracle restatement: the harness feeds MILC what it claims.
The GPU comparison of the two builds is tests/test_gpu_milc_glue.py.
"""
import os
import subprocess

import numpy as np
import pytest

import ks_oracle as ko
import milc_glue as mg

pytestmark = pytest.mark.skipif(not mg.available(), reason="oracle/_ref/glue not built")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def symbols(path, undefined):
    out = subprocess.run(["nm", "-D", "--undefined-only" if undefined else "--defined-only", path],
                         capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@pytest.mark.parametrize("p", [1, 2])
def test_hip_build_routes_through_the_library(p):
    path = mg.lib_path(f"hip_p{p}")
    defined, undefined = symbols(path, False), symbols(path, True)
    for name in ("ks_congrad_parity_gpu", "ks_congrad_block_parity_gpu", "ks_multicg_offset_field_gpu",
                 "get_fn_last", "set_fn_last", "eo_fermion_force_multi", "ks_congrad_field", "ks_multicg_field"):
        assert name in defined, name
    for name in ("milc_hip_init", "milc_hip_invert", "milc_hip_invert_block", "milc_hip_multishift_invert",
                 "milc_hip_hisq_force"):
        assert name in undefined, name
    # the library's defined entry points cover what the harness needs
    lib = symbols(os.path.join(ROOT, "milc_qcd_amd", "libmilc_hip.so"), False)
    assert {n for n in undefined if n.startswith("milc_hip_")} <= lib
    cpu = symbols(mg.lib_path(f"cpu_p{p}"), False)
    assert "ks_congrad_parity_gpu" not in cpu and "eo_fermion_force_multi" in cpu


def test_cpu_builds_solve_the_oracle_system():
    dims = (8, 8, 8, 8)
    L = ko.Lattice(dims)
    U = L.gauge_random(1000)
    fat, lng, _ = L.hisq_links(L.rephase(U))
    b = L.gaussian_vector(7, ko.EVEN)
    x_ref, _, _ = L.congrad(fat, lng, b, 0.05, 1e-12, parity=ko.EVEN, max_iters=1000, nrestart=5)
    s = L.parity_slice(ko.EVEN)
    for p, bar in ((2, 1e-9), (1, 1e-4)):
        B = mg.Build(f"cpu_p{p}")
        B.init(dims)
        B.set_thin_links(U)
        f, l = B.hisq_links()
        assert mg.rel(f, fat) < (1e-13 if p == 2 else 1e-6)
        x, it, q = B.congrad_field(b, np.zeros_like(b), mg.EVEN, 0.05, 1000, 5, 1e-10 if p == 2 else 2e-6)
        assert q["converged"] == 1
        assert mg.rel(x[s], x_ref[s]) < bar
