#!/usr/bin/env python3
"""Benchmark: HISQ CG time to an FP64-verified relative residual of 1e-10 on
one MI355X, mixed precision (16-bit per-site-normed sloppy iterations with
FP32 arithmetic and FP64 reliable updates) against the plain FP64 GPU solver.

Workload: 32^3 x 64 lattice, thin links U = exp(i 0.4 H) from seed 1001,
HISQ fat + Naik links built on the device, mass 0.01, even-site Gaussian
sources with seeds 2100 + j (j = 0..9).  One step = one mixed-precision solve
of one source, inputs resident in HBM.  value = seconds per solve (lower is
better).

Also reported: FP64 s/solve and the speedup, iteration counts, the error of
the mixed solution against the FP64 solution, Dslash algorithmic GB/s per
precision, a STREAM copy reference, the roofline fraction of the dominant
kernel (the 16-bit Dslash) and the CPU baseline (the FP64 oracle CG timed on
this host's cores over a bounded number of iterations).

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 1] [--no-cpu]
       python bench.py --workload c1   (the 8^4, m = 0.05 case: FP64 and mixed GPU solves to 1e-8
           against MILC's compiled CG on the same links; prints its own JSON line)
       python bench.py --workload multishift   (batched RHMC multi-shift: 24^3 x 48,
           m = 0.02, 12 shifts sigma_i = 1e-4 10^(4i/11), 16 sources, per-shift
           FP64 true residual 1e-9; prints its own JSON line)
       python bench.py --workload hmc   (one RHMC force + FP64 action on 48^3 x 96,
           m = 0.01, 12 shifts as above, residues 0.1 (i+1); prints its own JSON line)
       python bench.py --workload force   (the HISQ fermion force alone on 48^3 x 96,
           12 terms from resident vectors; prints its own JSON line)
       python bench.py --workload io   (lattice files at 32^3 x 64: v5 and SciDAC
           write / read throughput, file -> device HISQ links)
       python bench.py --workload trajectory   (one RHMC trajectory at 48^3 x 96 with the
           rational functions of MILC's HISQ RHMC test; FP64 dH)
       --no-sequential (multishift: leave out the one-source comparison solves)
       --no-reference (multishift: leave out the FP64 one-source reference solves too,
           so that only the batched solver runs, e.g. under a profiler)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (32, 32, 32, 64)
MASS = 0.01
RESID = 1e-10
GAUGE_SEED = 1001
SRC_SEEDS = [2100 + j for j in range(10)]
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per output site of one single-parity Dslash:
# 16 link matrices + 1 source + 1 destination vector
DSLASH_BYTES = {0: 16 * 144 + 48 + 48, 1: 16 * 72 + 24 + 24, 2: 16 * 36 + 16 + 16}
# 16-bit stencil with the Naik links in the two-row form (rows 0, 1 and the det phase: 26 B)
DSLASH_BYTES_H16_TWO_ROW = 8 * 36 + 8 * 26 + 16 + 16
DSLASH_FLOPS = 1146


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class stdout_to_stderr:
    """MILC's C code prints to fd 1; keep the benchmark's stdout to the one JSON line."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *a):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)


def host_info(lib=None):
    """nproc, CPU model and, for a compiled CPU baseline, the compiler that
    built it (the GCC identification string in the library's .comment data)."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        with open("/proc/cpuinfo") as f:
            info["cpu_model"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        info["cpu_model"] = None
    if lib:
        info["compiler"] = None
        try:
            with open(os.path.join(ROOT, lib), "rb") as f:
                data = f.read()
            k = data.find(b"GCC: (")
            if k >= 0:
                info["compiler"] = data[k:data.index(b"\0", k)].decode(errors="replace")
        except (OSError, ValueError):
            pass
    return info


def cpu_baseline(G, fp64_iters, budget_s=15.0):
    """MILC's own FP64 CG (compiled reference, ks_congrad_parity_cpu) on this
    host's cores over a bounded number of iterations; falls back to the FP64
    oracle port when the compiled reference is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    import milc_ref as mr
    if mr.available():
        import milc_qcd_amd as mq
        t0 = time.time()
        mr.init(DIMS)
        fat, lng = G.dev_read_links()
        mr.set_fn_links(fat, lng)
        del fat, lng
        src = G.dev_read(0, mq.EVEN)
        log(f"cpu baseline (MILC reference): setup {time.time() - t0:.1f}s")
        t1 = time.perf_counter()
        _, it, _ = mr.congrad(src, MASS, RESID, mr.EVEN, 2, 1)
        per = (time.perf_counter() - t1) / max(it, 1)
        k = max(2, min(200, int(budget_s / max(per, 1e-6))))
        t1 = time.perf_counter()
        _, it, _ = mr.congrad(src, MASS, RESID, mr.EVEN, k, 1)
        dt = time.perf_counter() - t1
        s_per_iter = dt / it
        return {
            "value": s_per_iter * fp64_iters,
            "unit": "s/solve",
            "cores": threads,
            "kind": "reference",
            **host_info("oracle/_ref/libmilc_ref.so"),
            "sample": (f"MILC 7.8.1 ks_congrad_parity_cpu (FP64, FEWSUMS, DBLSTORE_FN, OpenMP {threads} threads), "
                       f"compiled from the reference sources, on the same 32^3x64 HISQ links and source: "
                       f"{it} iterations timed ({dt:.1f} s, {s_per_iter * 1e3:.1f} ms/iter), scaled by the "
                       f"GPU FP64 iteration count {fp64_iters}"),
        }
    import ks_oracle as ko
    t0 = time.time()
    L = ko.Lattice(DIMS)
    fat, lng = G.dev_read_links()
    src = L.gaussian_vector(SRC_SEEDS[0], ko.EVEN)
    log(f"cpu baseline: setup {time.time() - t0:.1f}s")
    # calibrate with 2 iterations, then run a sample that fits the budget
    t1 = time.perf_counter()
    L.congrad(fat, lng, src, MASS, RESID, parity=ko.EVEN, max_iters=2, nrestart=1)
    per = (time.perf_counter() - t1) / 2
    k = max(2, min(200, int(budget_s / max(per, 1e-6))))
    t1 = time.perf_counter()
    _, it, _ = L.congrad(fat, lng, src, MASS, RESID, parity=ko.EVEN, max_iters=k, nrestart=1)
    dt = time.perf_counter() - t1
    s_per_iter = dt / it
    del fat, lng
    return {
        "value": s_per_iter * fp64_iters,
        "unit": "s/solve",
        "cores": threads,
        "kind": "port",
        **host_info("oracle/libks_oracle.so"),
        "sample": (f"FP64 oracle CG (MILC ks_congrad_parity_cpu algorithm, OpenMP) on the same 32^3x64 HISQ links "
                   f"and source: {it} iterations timed ({dt:.1f} s, {s_per_iter * 1e3:.1f} ms/iter), "
                   f"scaled by the GPU FP64 iteration count {fp64_iters}"),
    }


MS_DIMS = (24, 24, 24, 48)
MS_MASS = 0.02
MS_SIGMAS = [1e-4 * 10 ** (4 * i / 11) for i in range(12)]
MS_GAUGE_SEED = 1002
MS_SRC_SEEDS = [2200 + j for j in range(16)]
MS_RESID = 1e-9        # per-shift FP64 true residual
MS_STAGE_RESID = 1e-8  # FP32 multi-shift stage (base system), then per-shift refinement


def cpu_multishift_baseline(G, dims, offs, fp64_iters, nsrc, budget_s=15.0):
    """MILC's own ks_multicg_offset_field_cpu (compiled reference, FP64) on the
    host cores, same links and source 0, over a bounded number of iterations,
    scaled to fp64_iters iterations per source and nsrc sources."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import milc_ref as mr
    if not mr.available():
        return None
    import milc_qcd_amd as mq
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    mr.init(dims)
    fat, lng = G.dev_read_links()
    mr.set_fn_links(fat, lng)
    del fat, lng
    src = G.dev_read(0, mq.EVEN)
    t1 = time.perf_counter()
    _, it, _ = mr.multicg(src, offs, 1e-30, mr.EVEN, 3)
    per = (time.perf_counter() - t1) / max(it, 1)
    k = max(3, min(300, int(budget_s / max(per, 1e-6))))
    t1 = time.perf_counter()
    _, it, _ = mr.multicg(src, offs, 1e-30, mr.EVEN, k)
    dt = time.perf_counter() - t1
    s_per_iter = dt / it
    return {
        "value": s_per_iter * fp64_iters * nsrc,
        "unit": "s/solve",
        "cores": threads,
        "kind": "reference",
        **host_info("oracle/_ref/libmilc_ref.so"),
        "sample": (f"MILC 7.8.1 ks_multicg_offset_field_cpu (FP64, {len(offs)} shifts, OpenMP {threads} threads), "
                   f"compiled from the reference sources, same {'x'.join(map(str, dims))} HISQ links and source 0: "
                   f"{it} iterations timed ({dt:.1f} s, {s_per_iter * 1e3:.1f} ms/iter), scaled by the GPU FP64 "
                   f"multi-shift iteration count {fp64_iters} to the target and by {nsrc} sources"),
    }


def multishift_workload(args):
    """16 sources x 12 shifts in one batched solve (FP32 multi-shift with FP64
    recurrences and deferred shift updates, captured window graph, autotuned
    batched stencil; per-shift mixed refinement with the 16-bit batched
    stencil and FP64 reliable updates), against 16 one-source solves
    of the same solver family and an FP64 reference solve of source 0."""
    import milc_qcd_amd as mq
    offs = [4 * MS_MASS ** 2 + s for s in MS_SIGMAS]
    n = len(offs)
    G = mq.Lattice(MS_DIMS)
    G.dev_gauge_hisq(MS_GAUGE_SEED, 0.4)
    for j, sd in enumerate(MS_SRC_SEEDS):
        G.dev_source(sd, mq.EVEN, j)
    Vh = G.Vh
    # autotuner: every variant of the 16-rhs FP32 stencil, timed
    nvar, chosen, us_best, desc = G.dev_block_tune(1, 16)
    variants = []
    for v in range(nvar):
        us = G.dev_block_dslash_bench(1, 16, v, 20)
        variants.append({"variant": v, "usec": round(us, 2), "desc": G.dev_block_describe(1, 16, v)})
    bytes_site = 16 * 72 + 16 * (24 + 24)
    log(f"batched stencil: {nvar} variants, chosen {chosen} ({desc}) {us_best:.1f} us")
    # the 16-bit form of the refinement CG: 16-bit fat links (36 B), Naik links
    # (26 B in the two-row form, else 36 B), 16-byte records in and out
    nvar16, chosen16, _, desc16 = G.dev_block_tune(3, 16)
    us16 = G.dev_block_dslash_bench(3, 16, chosen16, 20)
    bytes_site16 = 8 * 36 + 8 * (26 if G.link_compression(2) == 1 else 36) + 16 * (16 + 16)
    log(f"16-bit batched stencil: {nvar16} variants, chosen {chosen16} ({desc16}) {us16:.1f} us")
    if args.kernel_only:
        # the tuned stencils alone, for counter runs: no solver is launched
        us_f32_16 = G.dev_block_dslash_bench(1, 16, chosen, 20)
        G.close()
        print(json.dumps({"workload": "batched stencils, 16 rhs, 24^3x48",
                          "f32": {"usec": round(us_f32_16, 2), "alg_bytes_per_site": bytes_site, "desc": desc},
                          "16bit": {"usec": round(us16, 2), "alg_bytes_per_site": bytes_site16, "desc": desc16}}),
              flush=True)
        return
    for i in range(args.warmup):
        G.dev_multishift_block(list(range(16)), offs, MS_STAGE_RESID, MS_RESID)
    t1 = time.perf_counter()
    for i in range(args.steps):
        st = G.dev_multishift_block(list(range(16)), offs, MS_STAGE_RESID, MS_RESID, out_slot0=16, out_nsrc=1)
        if not st["converged"]:
            raise RuntimeError(f"batched multi-shift did not converge: {st}")
    t_batch = (time.perf_counter() - t1) / args.steps
    log(f"batched: {t_batch:.3f} s per 16-source solve; {st}")
    # FP64 reference of source 0 and the batched solutions' error against it
    it64 = t64 = it64_target = None
    errs = [None]
    if not args.no_reference:
        it64, t64, _ = G.dev_multishift(0, offs, 1e-12, 0, out_slot0=40)
        errs = [G.dev_diff(16 + j, 40 + j) for j in range(n)]
        it64_target, _, _ = G.dev_multishift(0, offs, MS_RESID, 0)
    # one-source solves (same FP32 + refinement scheme), all 16 sources
    t_seq, seq_iters = None, None
    if not (args.no_sequential or args.no_reference):
        t1 = time.perf_counter()
        seq_iters = []
        for j in range(16):
            it, _, conv = G.dev_multishift(j, offs, MS_RESID, 1)
            seq_iters.append(it)
        t_seq = time.perf_counter() - t1
    us_f32 = G.dev_dslash_bench(1, mq.EVEN, 20)
    cpu = None
    if not args.no_cpu and it64_target is not None:
        try:
            with stdout_to_stderr():
                cpu = cpu_multishift_baseline(G, MS_DIMS, offs, it64_target, 16)
        except Exception as e:  # the CPU leg must not hide the GPU result
            log(f"cpu baseline failed: {e!r}")
    G.close()
    # the kernel with the larger share of the solve's stencil time (two
    # applications per multi-shift and per refinement iteration) is the one the
    # roofline line reports
    t32 = 2 * st["ms_iters"] * us_best
    t16 = 2 * sum(st["refine_iters"]) * us16
    share16 = t16 / (t16 + t32)
    if share16 >= 0.5:
        roof = {"bound": "hbm", "achieved": round(bytes_site16 * Vh / us16 / 1e3, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(bytes_site16 * Vh / us16 / 1e3 / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"16-bit batched stencil of the refinement, 16 rhs ({desc16})"}
    else:
        roof = {"bound": "hbm", "achieved": round(bytes_site * Vh / us_best / 1e3, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(bytes_site * Vh / us_best / 1e3 / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"batched stencil of the multi-shift stage, 16 rhs ({desc})"}
    out = {
        "metric": "batched multi-shift RHMC solve: 16 sources x 12 shifts to per-shift FP64 true residual 1e-9",
        "value": round(t_batch, 4),
        "unit": "s/solve (16 sources x 12 shifts)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 multi-shift + f64 recurrences, 16-bit/f64 mixed refinement (batched 16-bit stencil)",
        "data": "synthetic",
        "config": {"workload": "hisq_multishift_batched_24x24x24x48_m0.02_12shifts_16src", "lattice": list(MS_DIMS),
                   "mass": MS_MASS, "sigmas": MS_SIGMAS, "sources": 16, "gauge_seed": MS_GAUGE_SEED,
                   "target_true_rel_resid": MS_RESID, "multishift_stage_resid": MS_STAGE_RESID},
        "multishift_iters": st["ms_iters"],
        "multishift_seconds": round(st["ms_seconds"], 4),
        "refine_seconds": round(st["refine_seconds"], 4),
        "refine_iters_per_shift": st["refine_iters"],
        "true_rel_resid_per_shift_max": st["true_rel"],
        "rel_err_vs_fp64_solve_src0": None if errs[0] is None else max(errs),
        "fp64_reference": None if it64 is None else {"iters": it64, "seconds": round(t64, 4)},
        "sequential_16_one_source_solves_s": None if t_seq is None else round(t_seq, 4),
        "speedup_batched_vs_sequential": None if t_seq is None else round(t_seq / t_batch, 3),
        "sequential_iters": seq_iters,
        "autotune": {"chosen": chosen, "desc": desc, "variants": variants},
        "batched_dslash": {"usec": round(us_best, 2), "alg_bytes_per_site": bytes_site,
                           "GBps": round(bytes_site * Vh / us_best / 1e3, 1),
                           "usec_per_rhs": round(us_best / 16, 2),
                           "single_rhs_f32_dslash_usec": round(us_f32, 2)},
        "batched_dslash_16bit": {"usec": round(us16, 2), "alg_bytes_per_site": bytes_site16,
                                 "GBps": round(bytes_site16 * Vh / us16 / 1e3, 1), "desc": desc16,
                                 "share_of_stencil_time": round(share16, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


HMC_DIMS = (48, 48, 48, 96)
HMC_MASS = 0.01
HMC_GAUGE_SEED = 1004
HMC_SRC_SEED = 2400
HMC_RESID = 1e-9


def hmc_workload(args):
    """One molecular-dynamics fermion force (12-shift mixed multi-shift, D_oe,
    outer products, HISQ chain to the thin-link momentum) plus the fermion
    action (its own mixed multi-shift solve to FP64 true residual 1e-10, the
    sum in FP64) on 48^3 x 96; force and action are compared with the same
    computations in FP64 on the device."""
    import milc_qcd_amd as mq
    offs = [4 * HMC_MASS ** 2 + s for s in MS_SIGMAS]
    res = [0.1 * (i + 1) for i in range(len(offs))]
    G = mq.Lattice(HMC_DIMS)
    t0 = time.time()
    G.dev_gauge_hisq(HMC_GAUGE_SEED, 0.4)
    log(f"48^3x96 links: {time.time() - t0:.1f}s")
    G.dev_source(HMC_SRC_SEED, mq.EVEN, 0)
    Vh = G.Vh
    dsl = {}
    two_row = G.link_compression(2) == 1
    for prec, name in ((0, "f64"), (1, "f32"), (2, "h16")):
        us = G.dev_dslash_bench(prec, mq.EVEN, 10)
        nb = DSLASH_BYTES_H16_TWO_ROW if (prec == 2 and two_row) else DSLASH_BYTES[prec]
        dsl[name] = {"usec": round(us, 1), "alg_bytes_per_site": nb, "GBps": round(nb * Vh / us / 1e3, 1),
                     "frac_peak": round(nb * Vh / us / 1e3 / HBM_PEAK_GBS, 4)}
    log(f"dslash {dsl}")
    steps = []
    for i in range(args.warmup + args.steps):
        st = G.dev_hmc_step(0, offs, res, HMC_RESID, mixed=1, compare_fp64=(i == args.warmup), thin_force=True)
        log(f"step {i}: {st}")
        if i >= args.warmup:
            steps.append(st)
    cpu = None
    if not args.no_cpu:
        try:
            it64, _, _ = G.dev_multishift(0, offs, HMC_RESID, 0)
            with stdout_to_stderr():
                cpu = cpu_multishift_baseline(G, HMC_DIMS, offs, it64, 1)
            if cpu:
                cpu["sample"] += "; the multi-shift solve of the step only (the CPU force and action are not timed)"
        except Exception as e:
            log(f"cpu baseline failed: {e!r}")
    G.close()
    t_step = sum(s["force_seconds"] + s["thin_force_seconds"] + s["action_seconds"] for s in steps) / len(steps)
    cmp = [s for s in steps if s["rel_err_O1"] >= 0][0]
    out = {
        "metric": "RHMC fermion force (to thin-link momentum) + FP64 action per MD step (48^3x96 HISQ, 12 shifts)",
        "value": round(t_step, 4),
        "unit": "s/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 multi-shift + f64 recurrences, 16-bit/f64 mixed refinement per shift (force and action "
                 "solves, to f64 true residuals), f64 force, f64 action sum",
        "data": "synthetic",
        "config": {"workload": "hisq_rhmc_force_action_48x48x48x96_m0.01_12shifts", "lattice": list(HMC_DIMS),
                   "mass": HMC_MASS, "sigmas": MS_SIGMAS, "residues": res, "gauge_seed": HMC_GAUGE_SEED,
                   "source_seed": HMC_SRC_SEED, "force_resid": HMC_RESID},
        "force_seconds": round(sum(s["force_seconds"] for s in steps) / len(steps), 4),
        "thin_link_hisq_force_seconds": round(sum(s["thin_force_seconds"] for s in steps) / len(steps), 4),
        "thin_link_momentum_rel_err_vs_fp64": cmp["rel_err_mom"],
        "action_seconds": round(sum(s["action_seconds"] for s in steps) / len(steps), 4),
        "force_iters": steps[-1]["force_iters"],
        "action_iters": steps[-1]["action_iters"],
        "action": steps[-1]["action"],
        "action_rel_err_vs_fp64": cmp["rel_err_action"],
        "force_rel_err_vs_fp64": {"O1": cmp["rel_err_O1"], "O3": cmp["rel_err_O3"]},
        "fp64_force_seconds": round(cmp["fp64_force_seconds"], 4),
        "dslash": dsl,
        "roofline": {"bound": "hbm", "achieved": dsl["f32"]["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dsl["f32"]["frac_peak"], "traffic": None,
                     "kernel": "k_dslash<P_F32> (the multi-shift bulk, 1200 B/site alg.)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


FORCE_TERMS = 12
FORCE_SRC_SEEDS = [2600 + t for t in range(FORCE_TERMS)]
FORCE_CPU_DIMS = (16, 16, 16, 32)


def cpu_force_baseline(res, eps):
    """MILC's own eo_fermion_force_multi (compiled reference, FP64, HISQ) on
    the host cores: one call with the same number of terms on a 16^3 x 32
    sample lattice (same gauge generator), scaled by the volume ratio."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import milc_ref as mr
    if not mr.available():
        return None
    import ks_oracle as ko
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    L = ko.Lattice(FORCE_CPU_DIMS)
    V = mr.init(FORCE_CPU_DIMS)
    mr.set_thin_links(L.gauge_random(HMC_GAUGE_SEED))
    mr.hisq_links(V)
    xs = [L.gaussian_vector(sd, ko.EVENANDODD) for sd in FORCE_SRC_SEEDS]
    t1 = time.perf_counter()
    mr.hisq_force(xs, res, eps, V)
    dt = time.perf_counter() - t1
    scale = (HMC_DIMS[0] * HMC_DIMS[1] * HMC_DIMS[2] * HMC_DIMS[3]) / V
    return {
        "value": dt * scale,
        "unit": "s/force",
        "cores": threads,
        "kind": "reference",
        **host_info("oracle/_ref/libmilc_ref.so"),
        "sample": (f"MILC 7.8.1 eo_fermion_force_multi (HISQ, FP64, {len(res)} terms, OpenMP {threads} threads), "
                   f"compiled from the reference sources, one call on a {'x'.join(map(str, FORCE_CPU_DIMS))} lattice "
                   f"(thin links from gauge seed {HMC_GAUGE_SEED}, Gaussian Z_t): {dt:.2f} s, scaled by the volume "
                   f"ratio {scale:.0f} to {'x'.join(map(str, HMC_DIMS))}"),
    }


def force_workload(args):
    """The HISQ fermion force alone on 48^3 x 96: 12 terms, X_t resident on the
    device (Gaussian even-site vectors; the force's cost does not depend on
    them being solutions), D_oe X_t, outer products, level-2 smearing adjoint
    on W, U(3) projection derivative, level-1 adjoint on U, TA into the
    momentum.  One step = one force."""
    import milc_qcd_amd as mq
    res = [0.1 * (i + 1) for i in range(FORCE_TERMS)]
    eps = 1.0
    G = mq.Lattice(HMC_DIMS)
    G.dev_gauge_hisq(HMC_GAUGE_SEED, 0.4)
    for t, sd in enumerate(FORCE_SRC_SEEDS):
        G.dev_source(sd, mq.EVEN, t)
    steps = []
    for i in range(args.warmup + args.steps):
        st = G.dev_hisq_force(0, res, eps)
        log(f"force {i}: {st}")
        if i >= args.warmup:
            steps.append(st)
    G.close()
    cpu = None
    if not args.no_cpu:
        try:
            with stdout_to_stderr():
                cpu = cpu_force_baseline(res, eps)
        except Exception as e:
            log(f"cpu baseline failed: {e!r}")
    t_oprod = sum(s["oprod_seconds"] for s in steps) / len(steps)
    t_thin = sum(s["thin_force_seconds"] for s in steps) / len(steps)
    t = t_oprod + t_thin
    out = {
        "metric": "HISQ fermion force from resident solutions to the thin-link momentum (48^3x96, 12 terms)",
        "value": round(t, 4),
        "unit": "s/force",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "hisq_fermion_force_48x48x48x96_12terms", "lattice": list(HMC_DIMS),
                   "terms": FORCE_TERMS, "residues": res, "eps": eps, "gauge_seed": HMC_GAUGE_SEED,
                   "source_seeds": FORCE_SRC_SEEDS},
        "oprod_seconds": round(t_oprod, 4),
        "thin_link_chain_seconds": round(t_thin, 4),
        "aux_links_seconds": round(steps[-1]["aux_seconds"], 4),
        "mom_norm2": steps[-1]["mom_norm2"],
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


TRAJ_DIMS = (48, 48, 48, 96)
TRAJ_STEPS = 4         # INT_3G1F: 2 step pairs, 4 fermion and 12 gauge kicks
TRAJ_EPS = 0.02
TRAJ_BETA, TRAJ_U0, TRAJ_NF = 7.4, 0.89, 3   # the couplings of MILC's su3_rhmc_hisq test
TRAJ_GAUGE_SCALE = 0.1  # U = exp(0.1 i H): a smooth start, forces of a physical size
# cgresid_md_fa_gr of MILC's su3_rhmc_hisq.1 input, per pseudofermion (MD, GR, FA)
TRAJ_RESID = [(5e-4, 1e-4, 1e-4)] + [(5e-6, 1e-6, 1e-6)] * 3


def trajectory_workload(args):
    """One RHMC trajectory on the device at 48^3 x 96 as MILC's HISQ build runs
    it (su3_rhmc_hisq: INT_3G1F, four pseudofermions with the rational
    functions of its test, masses 0.0018 / 0.018 folded into the poles, its
    solver residuals): Gaussian momenta, phi_i = R_GR chi_i, actions, the
    3G1F integrator (gauge force 3x per fermion force, HISQ fermion force
    from all 30 MD terms), mixed-precision solves (FP32 multi-shift, every
    shift refined to its FP64 true residual), fat and Naik links
    rebuilt when used, reunitarisation per step pair, FP64 dH.  Each timed
    trajectory starts from the same links (the evolved links are restored)."""
    import milc_qcd_amd as mq
    pfs = mq.pseudofermions_from_file(os.path.join(ROOT, "tests", "golden", "rationals.sample.su3_rhmc_hisq"),
                                      [r[0] for r in TRAJ_RESID], [r[1] for r in TRAJ_RESID],
                                      [r[2] for r in TRAJ_RESID])
    coeff = mq.gauge_loop_coeffs(TRAJ_U0, TRAJ_NF, True)
    G = mq.Lattice(TRAJ_DIMS)
    G.dev_gauge_hisq(HMC_GAUGE_SEED, TRAJ_GAUGE_SCALE)
    runs = []
    for i in range(args.warmup + args.steps):
        st = G.dev_rhmc_trajectory(TRAJ_STEPS, TRAJ_EPS, TRAJ_BETA, coeff, pfs, integrator=mq.INT_3G1F,
                                   mixed_md=1, mixed_action=1, accept=2, seed=100 + i)
        log(f"trajectory {i}: {st}")
        if i >= args.warmup:
            runs.append(st)
    G.close()
    t = sum(r["seconds"] for r in runs) / len(runs)
    last = runs[-1]
    out = {
        "metric": "RHMC trajectory on one GPU (48^3x96 HISQ, 3G1F, 4 pseudofermions, FP64 dH)",
        "value": round(t, 3),
        "unit": "s/trajectory",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 1),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 actions, forces and link updates; f32/16-bit mixed solves to f64 true residuals",
        "data": "synthetic",
        "config": {"workload": "rhmc_trajectory_48x48x48x96_3g1f4", "lattice": list(TRAJ_DIMS),
                   "integrator": "INT_3G1F", "md_steps": TRAJ_STEPS, "epsilon": TRAJ_EPS, "beta": TRAJ_BETA,
                   "u0": TRAJ_U0, "rationals": "tests/golden/rationals.sample.su3_rhmc_hisq, 4 pseudofermions",
                   "resid_md_gr_fa": TRAJ_RESID, "gauge_seed": HMC_GAUGE_SEED, "gauge_scale": TRAJ_GAUGE_SCALE},
        "dH": [r["dH"] for r in runs],
        "actions_start": {"gauge": last["g0"], "momentum": last["h0"], "fermion": last["f0"]},
        "per_fermion_force_s": round(sum(r["fermion_force_seconds"] for r in runs)
                                     / sum(r["fermion_forces"] for r in runs), 4),
        "per_gauge_force_s": round(sum(r["gauge_force_seconds"] for r in runs)
                                   / sum(r["gauge_forces"] for r in runs), 4),
        "link_rebuilds": last["link_rebuilds"],
        "fermion_force_s": round(sum(r["fermion_force_seconds"] for r in runs) / len(runs), 3),
        "gauge_force_s": round(sum(r["gauge_force_seconds"] for r in runs) / len(runs), 3),
        "link_update_rebuild_s": round(sum(r["link_seconds"] for r in runs) / len(runs), 3),
        "action_s": round(sum(r["action_seconds"] for r in runs) / len(runs), 3),
        "md_solver_iters": last["md_iters"],
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_io_baseline(path, nbytes):
    """the FP32 SciDAC file read by the numpy oracle reader (LIME parse, CRC-32
    checksums, byte swap), one pass, on this host"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ks_io as oio
    t1 = time.perf_counter()
    f = oio.read_lime_field(path)
    dt = time.perf_counter() - t1
    ok = f["suma"] == f["calca"] and f["sumb"] == f["calcb"]
    return {"value": round(nbytes / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            **host_info(), "compiler": f"numpy {np.__version__} (Python)",
            "sample": (f"numpy restatement of the SciDAC reader (oracle/ks_io.py) on the same "
                       f"{nbytes / 1e6:.0f} MB file: {dt:.2f} s, checksum {'ok' if ok else 'MISMATCH'}")}


def io_workload(args):
    """Lattice files on the 32^3 x 64 configuration: the device-generated thin
    links written as MILC v5 and as FP32 SciDAC, read back (parse, checksum,
    byte swap, reorder to MILC's node_index order) and loaded onto the device,
    where the HISQ links are rebuilt.  value = SciDAC read throughput (file
    bytes / read time, file in the page cache)."""
    import shutil
    import tempfile
    import milc_qcd_amd as mq
    from milc_qcd_amd import lattice_io as lio
    G = mq.Lattice(DIMS)
    G.dev_gauge_hisq(GAUGE_SEED, 0.4)
    fat0, lng0 = G.dev_read_links()
    U = G.dev_read_thin_links()
    d = tempfile.mkdtemp(prefix="milc_io_")
    res = {}
    try:
        dims = list(DIMS)
        for name, flag in (("scidac_f32", lio.SAVE_SERIAL_SCIDAC), ("v5", lio.SAVE_SERIAL)):
            f = os.path.join(d, name)
            tw, tr = [], []
            for i in range(args.warmup + args.steps):
                t1 = time.perf_counter()
                lio.save_lattice(flag, f, U, dims, precision=1)
                t2 = time.perf_counter()
                U2, info = lio.reload_lattice(f, dims)
                t3 = time.perf_counter()
                if not info["checksum_ok"]:
                    raise RuntimeError(f"{name}: checksum mismatch after round trip")
                if i >= args.warmup:
                    tw.append(t2 - t1)
                    tr.append(t3 - t2)
            nbytes = os.path.getsize(f)
            res[name] = {"file_bytes": nbytes, "write_s": round(min(tw), 4), "read_s": round(min(tr), 4),
                         "write_GBps": round(nbytes / min(tw) / 1e9, 3), "read_GBps": round(nbytes / min(tr) / 1e9, 3)}
            log(f"{name}: {res[name]}")
        # file -> device thin links -> HISQ links on the device
        t1 = time.perf_counter()
        G.dev_load_thin_links(U2)
        t_load = time.perf_counter() - t1
        fat, lng = G.dev_read_links()
        err = float(np.linalg.norm(fat - fat0) / np.linalg.norm(fat0))
        cpu = None
        if not args.no_cpu:
            try:
                cpu = cpu_io_baseline(os.path.join(d, "scidac_f32"), res["scidac_f32"]["file_bytes"])
            except Exception as e:
                log(f"cpu baseline failed: {e!r}")
    finally:
        shutil.rmtree(d, ignore_errors=True)
    G.close()
    sc = res["scidac_f32"]
    out = {
        "metric": "SciDAC gauge-file read (LIME parse, CRC-32 checksums, byte swap, reorder) throughput, 32^3x64",
        "value": sc["read_GBps"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sc["read_s"] * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 file, f64 in memory",
        "data": "synthetic",
        "config": {"workload": "lattice_files_32x32x32x64", "lattice": list(DIMS), "gauge_seed": GAUGE_SEED},
        "formats": res,
        "file_to_device_hisq_links_s": round(t_load, 4),
        "fat_links_rel_diff_after_fp32_file": err,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


C1_DIMS = (8, 8, 8, 8)
C1_MASS = 0.05
C1_RESID = 1e-8


def c1_workload(args):
    """The CPU-runnable case: 8^4, m = 0.05, U = exp(0.4 i H) from seed 1001,
    HISQ links built on the device, even-site Gaussian sources 2100 + j,
    (M^dag M) x = b to 1e-8.  FP64 and 16-bit mixed solves on the GPU, the
    FP64 solution against the oracle-grade FP64 solve of MILC's own CG
    (compiled reference, oracle/_ref) on the same links and source, and that
    CG timed on this host's cores.  The lattice (4096 sites per parity) lives in
    cache: the GPU times are launch latency, not bandwidth."""
    import milc_qcd_amd as mq
    G = mq.Lattice(C1_DIMS)
    G.dev_gauge_hisq(GAUGE_SEED, 0.4)
    for j in range(4):
        G.dev_source(SRC_SEEDS[j], mq.EVEN, j)
    res = {}
    for mixed, name in ((0, "fp64"), (2, "mixed16")):
        for i in range(args.warmup):
            G.dev_invert(0, 10 + (mixed > 0), C1_MASS, mq.EVEN, C1_RESID, mixed)
        its, rr = [], []
        t1 = time.perf_counter()
        for i in range(args.steps):
            it, st = G.dev_invert(i % 4, 10 + (mixed > 0), C1_MASS, mq.EVEN, C1_RESID, mixed)
            its.append(it)
            rr.append(st["resid"])
        res[name] = {"s_per_solve": (time.perf_counter() - t1) / args.steps, "iters": its, "true_rel_resid": max(rr)}
        log(f"{name}: {res[name]}")
    # both GPU solutions of the last source against each other
    err_mixed = G.dev_diff(11, 10)
    cpu, err_fp64 = None, None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import milc_ref as mr
    if mr.available():
        with stdout_to_stderr():
            mr.init(C1_DIMS)
            fat, lng = G.dev_read_links()
            mr.set_fn_links(fat, lng)
            last = (args.steps - 1) % 4
            src = G.dev_read(last, mq.EVEN)
            x_gpu = G.dev_read(10, mq.EVEN)
            t1 = time.perf_counter()
            reps = 5
            for _ in range(reps):
                x_ref, it_ref, _ = mr.congrad(src, C1_MASS, C1_RESID, mr.EVEN, 1000, 5)
            dt = (time.perf_counter() - t1) / reps
        Vh = G.Vh
        err_fp64 = float(np.linalg.norm(x_gpu[:Vh] - x_ref[:Vh]) / np.linalg.norm(x_ref[:Vh]))
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": dt, "unit": "s/solve", "cores": threads, "kind": "reference",
               **host_info("oracle/_ref/libmilc_ref.so"),
               "sample": f"MILC 7.8.1 ks_congrad_parity_cpu, full solve to 1e-8 ({it_ref} iterations), mean of {reps}"}
    G.close()
    out = {
        "metric": "HISQ CG time-to-1e-8 on 8^4 (FP64 and mixed, 1 GPU; cache-resident)",
        "value": round(res["mixed16"]["s_per_solve"], 6),
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["mixed16"]["s_per_solve"] * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "16-bit storage / f32 arithmetic / f64 reliable updates",
        "data": "synthetic",
        "config": {"workload": "hisq_cg_8x8x8x8_m0.05", "lattice": list(C1_DIMS), "mass": C1_MASS,
                   "target_true_rel_resid": C1_RESID, "gauge_seed": GAUGE_SEED},
        "fp64_s_per_solve": round(res["fp64"]["s_per_solve"], 6),
        "iters_fp64": res["fp64"]["iters"],
        "iters_mixed": res["mixed16"]["iters"],
        "true_rel_resid_fp64": res["fp64"]["true_rel_resid"],
        "true_rel_resid_mixed": res["mixed16"]["true_rel_resid"],
        "rel_err_fp64_vs_milc_solution": err_fp64,
        "rel_err_mixed_vs_fp64_solution": err_mixed,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the Dslash and copy bursts, or for multishift the tuned batched stencils "
                         "(for rocprofv3 counter runs)")
    ap.add_argument("--workload", choices=["cg", "c1", "multishift", "hmc", "force", "io", "trajectory"], default="cg")
    ap.add_argument("--no-sequential", action="store_true",
                    help="multishift: leave out the 16 one-source comparison solves")
    ap.add_argument("--no-reference", action="store_true",
                    help="multishift: leave out the FP64 one-source reference solves (implies --no-sequential)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the multi-GPU domain decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c1":
        c1_workload(args)
        return
    if args.workload == "multishift":
        multishift_workload(args)
        return
    if args.workload == "hmc":
        hmc_workload(args)
        return
    if args.workload == "force":
        force_workload(args)
        return
    if args.workload == "io":
        io_workload(args)
        return
    if args.workload == "trajectory":
        trajectory_workload(args)
        return

    import milc_qcd_amd as mq

    G = mq.Lattice(DIMS)
    t0 = time.time()
    nsvd = G.dev_gauge_hisq(GAUGE_SEED, 0.4)
    log(f"gauge + HISQ links on device: {time.time() - t0:.2f}s, svd-fallback links {nsvd}")
    for j, s in enumerate(SRC_SEEDS):
        G.dev_source(s, mq.EVEN, j)
    XM, XD = 10, 11
    if args.kernel_only:
        res = {}
        two_row = G.link_compression(2) == 1
        for prec, name in ((0, "f64"), (1, "f32"), (2, "h16")):
            us = G.dev_dslash_bench(prec, mq.EVEN, 20)
            nb = DSLASH_BYTES_H16_TWO_ROW if (prec == 2 and two_row) else DSLASH_BYTES[prec]
            res[name] = {"usec": round(us, 2), "alg_bytes_per_launch": nb * G.Vh}
        us = G.dev_copy_bench(1 << 30, 5)
        res["copy"] = {"usec": round(us, 2), "bytes_read": 1 << 30, "bytes_written": 1 << 30}
        G.close()
        print(json.dumps({"kernel_only": res}), flush=True)
        return

    for i in range(args.warmup):
        G.dev_invert(i % 10, XM, MASS, mq.EVEN, RESID, 2)
    # ---- timed region: K mixed-precision solves --------------------------
    mixed_iters, mixed_res = [], []
    t1 = time.perf_counter()
    for i in range(args.steps):
        it, st = G.dev_invert(i % 10, XM, MASS, mq.EVEN, RESID, 2)
        mixed_iters.append(it)
        mixed_res.append(st["resid"])
        if not st["converged"]:
            raise RuntimeError(f"mixed solve {i} did not converge: {st}")
    t_mixed = (time.perf_counter() - t1) / args.steps
    log(f"mixed: {t_mixed * 1e3:.1f} ms/solve, iters {mixed_iters}")

    # ---- FP64 reference solves and FP32-sloppy variant (not the headline) ----
    n64 = min(args.steps, 10)
    fp64_iters = []
    t1 = time.perf_counter()
    for i in range(n64):
        it, st = G.dev_invert(i % 10, XD, MASS, mq.EVEN, RESID, 0)
        fp64_iters.append(it)
        if not st["converged"]:
            raise RuntimeError(f"fp64 solve {i} did not converge: {st}")
    t_fp64 = (time.perf_counter() - t1) / n64
    log(f"fp64: {t_fp64 * 1e3:.1f} ms/solve, iters {fp64_iters}")
    t1 = time.perf_counter()
    it32, st32 = G.dev_invert(0, XM, MASS, mq.EVEN, RESID, 1)
    t_f32 = time.perf_counter() - t1
    # error of the mixed solution against the FP64 solution, source 0
    G.dev_invert(0, XM, MASS, mq.EVEN, RESID, 2)
    G.dev_invert(0, XD, MASS, mq.EVEN, RESID, 0)
    err = G.dev_diff(XM, XD)

    # ---- Dslash roofline (HIP events on the library stream) ---------------
    Vh = G.Vh
    two_row = G.link_compression(2) == 1
    dsl = {}
    for prec, name in ((0, "f64"), (1, "f32"), (2, "h16")):
        us = G.dev_dslash_bench(prec, mq.EVEN, 50)
        nb = DSLASH_BYTES_H16_TWO_ROW if (prec == 2 and two_row) else DSLASH_BYTES[prec]
        dsl[name] = {"usec": round(us, 2), "alg_bytes_per_site": nb, "GBps": round(nb * Vh / us / 1e3, 1),
                     "frac_peak": round(nb * Vh / us / 1e3 / HBM_PEAK_GBS, 4),
                     "TFLOPs": round(DSLASH_FLOPS * Vh / us / 1e6, 2)}
    if two_row:
        # the same launch priced at the nine-element byte count, and the
        # nine-element long links themselves (format switched for this probe)
        dsl["h16"]["GBps_at_608_B_per_site"] = round(DSLASH_BYTES[2] * Vh / dsl["h16"]["usec"] / 1e3, 1)
        G.set_link_compression(0)
        us = G.dev_dslash_bench(2, mq.EVEN, 50)
        dsl["h16_nine_element_long_links"] = {"usec": round(us, 2), "alg_bytes_per_site": DSLASH_BYTES[2],
                                              "GBps": round(DSLASH_BYTES[2] * Vh / us / 1e3, 1)}
        t1 = time.perf_counter()
        G.dev_invert(0, XM, MASS, mq.EVEN, RESID, 2)
        dsl["h16_nine_element_long_links"]["mixed_s_per_solve"] = round(time.perf_counter() - t1, 5)
        G.set_link_compression(1)
    copy_bytes = 2 << 30
    us = G.dev_copy_bench(copy_bytes, 20)
    stream = 2 * copy_bytes / us / 1e3
    log(f"dslash {dsl}; stream copy {stream:.0f} GB/s")
    # the mixed CG iteration's kernels inside their own sequence (source 0),
    # beside the bursts above, and the time per iteration of the timed solves
    cgi = {k: round(v, 2) for k, v in G.dev_cg_iteration_bench(0, MASS, mq.EVEN, 50).items()}
    cgi["solve_us_per_iteration"] = round(t_mixed / (sum(mixed_iters) / len(mixed_iters)) * 1e6, 2)
    log(f"cg iteration in sequence: {cgi}")

    # HBM bytes per launch of the 16-bit Dslash from the committed rocprofv3
    # FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or null if absent
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "dslash_32x32x32x64_hbm_traffic.txt")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("hbm_bytes_per_launch k_dslash<2>"):
                traffic = int(line.split("=")[1])

    cpu = None
    if not args.no_cpu:
        try:
            with stdout_to_stderr():
                cpu = cpu_baseline(G, int(round(sum(fp64_iters) / len(fp64_iters))))
        except Exception as e:  # the CPU leg must not hide the GPU result
            log(f"cpu baseline failed: {e!r}")
    G.close()

    h16 = dsl["h16"]
    out = {
        "metric": "HISQ CG time-to-1e-10 (mixed vs FP64, 1 GPU) + Dslash HBM GB/s % of peak",
        "value": round(t_mixed, 5),
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "16-bit storage / f32 arithmetic / f64 reliable updates",
        "data": "synthetic",
        "config": {"workload": "hisq_cg_mixed_vs_fp64_32x32x32x64_m0.01", "lattice": list(DIMS), "mass": MASS,
                   "target_true_rel_resid": RESID, "sources": len(SRC_SEEDS), "gauge_seed": GAUGE_SEED,
                   "parallelism": "1 GPU"},
        "fp64_s_per_solve": round(t_fp64, 5),
        "speedup_mixed_vs_fp64": round(t_fp64 / t_mixed, 3),
        "f32_sloppy_s_per_solve": round(t_f32, 5),
        "iters_mixed": mixed_iters,
        "iters_fp64": fp64_iters,
        "iters_f32_sloppy": it32,
        "true_rel_resid_mixed": max(mixed_res),
        "rel_err_mixed_vs_fp64_solution": err,
        "dslash": dsl,
        "cg_iteration_in_solve_us": cgi,
        "stream_copy_GBps": round(stream, 1),
        "roofline": {"bound": "hbm", "achieved": h16["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": h16["frac_peak"], "traffic": traffic,
                     "kernel": (f"k_dslash<P_H16> (16-bit fat links, "
                                f"{'two-row' if two_row else 'nine-element'} 16-bit long links, 16-bit vectors, "
                                f"{DSLASH_BYTES_H16_TWO_ROW if two_row else DSLASH_BYTES[2]} B/site alg.)")},
        "long_links_two_row": two_row,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
