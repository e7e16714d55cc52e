"""MILC's own callers through the glue of integration/ against MILC's CPU build.

oracle/Makefile.glue compiles MILC 7.8.1's ks_congrad_field (d_congrad5_fn.c:16-36),
ks_multicg_field (ks_multicg.c:65-102, HYBRID) and eo_fermion_force_multi
(HISQ, FNMAT) twice per precision: as MILC's CPU build and with -DUSE_CG_GPU,
where the routing macros of include/imp_ferm_links.h:73-77,238-239 send the
solves through integration/d_congrad5_fn_hip.c and ks_multicg_offset_hip.c
into libmilc_hip.so, and integration/fermion_force_hisq_hip.c takes the place
of fermion_force_hisq_multi.o.  Both builds get the same thin links and build
their HISQ links with MILC's own code; the same calls then run through both.

MILC_PRECISION 2: the solutions agree to 1e-9 relative at resid 1e-10 (each
solve's error against the exact solution is bounded by its residual times the
condition number; in practice both land well inside that), iteration counts
agree to +-1, the relative residual (Fermilab, d_congrad5_fn_milc.c:37-56,92)
stops both at the same iteration with the same final_relrsq, and the force
agrees to 1e-12.  MILC_PRECISION 1: the host arrays are float; the CPU build
computes in float, the library in FP32/FP64, so the bars are those of float:
solutions 1e-4 at resid 2e-6, force 1e-5.
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import milc_glue  # noqa: E402

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not milc_glue.available(), reason="oracle/_ref/glue not built")]


@pytest.fixture(scope="module", params=[2, 1], ids=["MILC_PRECISION2", "MILC_PRECISION1"])
def result(request, tmp_path_factory):
    p = request.param
    out = tmp_path_factory.mktemp("glue") / f"p{p}.json"
    print(f"glue harness, MILC_PRECISION {p}", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "oracle", "milc_glue.py"), "--precision", str(p),
                        "--out", str(out)], timeout=240)
    assert r.returncode == 0, f"harness failed ({r.returncode})"
    with open(out) as f:
        return json.load(f)


def bars(p):
    return {"sol": 1e-9, "force": 1e-12} if p == 2 else {"sol": 1e-4, "force": 1e-5}


def test_links_same_in_both_builds(result):
    # MILC builds the links on the CPU in both builds
    assert result["links_cpu_vs_hip_build"] == 0.0


@pytest.mark.parametrize("case", ["cg_even", "cg_evenandodd"])
def test_ks_congrad_field(result, case):
    r = result[case]
    b = bars(result["precision"])
    assert r["qic_hip"]["converged"] == 1 and r["qic_cpu"]["converged"] == 1
    assert r["rel_solution"] < b["sol"]
    assert r["untouched_other_parity"]
    if result["precision"] == 2:
        # MILC's restarted FP64 CG, iteration for iteration
        assert abs(r["iters_hip"] - r["iters_cpu"]) <= 1


@pytest.mark.parametrize("case", ["cg_relresid_only", "cg_relresid_and_resid"])
def test_ks_congrad_field_relative_residual(result, case):
    r = result[case]
    b = bars(result["precision"])
    assert r["qic_hip"]["converged"] == 1
    assert r["rel_solution"] < b["sol"] * 10
    if result["precision"] == 2:
        assert abs(r["iters_hip"] - r["iters_cpu"]) <= 1
        # final_relrsq holds my_relative_residue of the solution in both builds
        q_c, q_h = r["qic_cpu"]["final_relrsq"], r["qic_hip"]["final_relrsq"]
        assert abs(q_h - q_c) <= 1e-3 * q_c
    else:
        assert r["qic_hip"]["final_relrsq"] < (3e-3) ** 2


def test_ks_multicg_field(result):
    r = result["multicg"]
    b = bars(result["precision"])
    # the base system stops at the same residual; the CPU build refines every
    # shift (ks_multicg.c:212), the GPU build only shifts on other links
    assert max(r["rel_solution"]) < b["sol"] * 10


def test_eo_fermion_force_multi(result):
    r = result["force"]
    assert r["norm"] > 0
    assert r["rel_mom"] < bars(result["precision"])["force"]
