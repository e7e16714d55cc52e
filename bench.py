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
"""
import argparse
import json
import os
import sys
import time

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
        "sample": (f"FP64 oracle CG (MILC ks_congrad_parity_cpu algorithm, OpenMP) on the same 32^3x64 HISQ links "
                   f"and source: {it} iterations timed ({dt:.1f} s, {s_per_iter * 1e3:.1f} ms/iter), "
                   f"scaled by the GPU FP64 iteration count {fp64_iters}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the Dslash and copy bursts (for rocprofv3 counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the multi-GPU domain decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

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
        for prec, name in ((0, "f64"), (1, "f32"), (2, "h16")):
            us = G.dev_dslash_bench(prec, mq.EVEN, 20)
            res[name] = {"usec": round(us, 2), "alg_bytes_per_launch": DSLASH_BYTES[prec] * G.Vh}
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
    dsl = {}
    for prec, name in ((0, "f64"), (1, "f32"), (2, "h16")):
        us = G.dev_dslash_bench(prec, mq.EVEN, 50)
        dsl[name] = {"usec": round(us, 2), "GBps": round(DSLASH_BYTES[prec] * Vh / us / 1e3, 1),
                     "frac_peak": round(DSLASH_BYTES[prec] * Vh / us / 1e3 / HBM_PEAK_GBS, 4),
                     "TFLOPs": round(DSLASH_FLOPS * Vh / us / 1e6, 2)}
    copy_bytes = 2 << 30
    us = G.dev_copy_bench(copy_bytes, 20)
    stream = 2 * copy_bytes / us / 1e3
    log(f"dslash {dsl}; stream copy {stream:.0f} GB/s")

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
        "stream_copy_GBps": round(stream, 1),
        "roofline": {"bound": "hbm", "achieved": h16["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": h16["frac_peak"], "traffic": traffic,
                     "kernel": "k_dslash<P_H16> (16-bit links+vectors, 608 B/site alg.)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
