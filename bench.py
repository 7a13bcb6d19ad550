#!/usr/bin/env python3
"""Benchmark: k-point solves/s at N = 288 (32-atom fcc supercell, MP 8x8x8).

One step = one pass of the per-k hot path over the whole structure: for
each of the 512 k-points build H(k), S(k), solve H c = e S c with vectors,
form the first-order pressure derivatives, then the Fermi search and the
band-energy / pressure sums (mode 4).  Inputs are synthetic (seeded NRL
parameters, seeded displacements) and already resident on the GPU when the
timed region starts.

Prints one JSON line (see the driver contract), with
  roofline      the dominant kernel's achieved rate vs the MI355X peak,
                timed with HIP events on the plan's stream;
  cpu_baseline  the CPU restatement (oracle) timed on this host's cores on a
                bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tight-binding_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from tbk import structures as S  # noqa: E402
from tbk._lib import TBK_FP64, TBK_MIXED  # noqa: E402
from tbk.api import Plan  # noqa: E402

METRIC = "k-pt solves/s (N=288) + served p50/p99 latency, mixed vs FP64 plain, same GPU"
FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 matrix
FP32_PEAK_TFLOPS = 157.3     # MI355X dense FP32 matrix
HBM_PEAK_GBS = 8000.0


def stage_model(n: int):
    """Algorithmic work per k-point of each stage (canonical LAPACK counts,
    complex arithmetic; bytes for the HBM-bound build)."""
    return {
        "build_hs": ("hbm", 2 * 16 * n * n),                 # write H and S
        "potrf_trtri": ("mfma", (4.0 / 3 + 4.0 / 3) * n ** 3),
        "reduce_gemm": ("mfma", 8.0 * n ** 3),                # two triangular GEMMs
        "hetrd": ("mfma", 16.0 / 3 * n ** 3),
        "stein": ("mfma", 0.0),                               # O(n^2): not a contraction
        "backtransform": ("mfma", 8.0 * n ** 3 + 4.0 * n ** 3),
        "pressure": ("mfma", 16.0 * n ** 3),
        "fermi_sums": ("hbm", 0.0),
    }


def cpu_baseline(st, kpts, weights, seconds=12.0):
    """The oracle on this host's cores over a bounded sample of k-points."""
    import oracle
    threads = min(os.cpu_count() or 1, 64)
    setup = oracle.Setup(st)
    nk = max(threads, 8)
    t0 = time.perf_counter()
    setup.bande(kpts[:threads], weights[:threads], mode=4, nthreads=threads)
    dt = time.perf_counter() - t0
    # scale the sample to about `seconds` of work
    reps = max(1, int(seconds / max(dt, 1e-3)))
    nk = min(len(kpts), threads * reps)
    t0 = time.perf_counter()
    setup.bande(kpts[:nk], weights[:nk], mode=4, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": nk / dt, "unit": "k-point solves/s", "cores": threads, "kind": "port",
            "sample": f"{nk} of the 512 k-points of the same structure (N={st.nv}), "
                      f"full per-k path incl. pressure, OpenMP over k, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    st = S.config_c2()
    kpts = S.monkhorst_pack((8, 8, 8))
    w = np.full(len(kpts), 1.0 / len(kpts))
    prec = TBK_MIXED if args.precision == "mixed" else TBK_FP64
    print(f"# workload C2: natoms={st.natoms} N={st.nv} nk={len(kpts)}", file=sys.stderr, flush=True)

    # ---- headline: all k-points batched -------------------------------------
    plan = Plan(st, kpts, w, batch=0, precision=prec)
    tune = {"shapes": 0, "seconds": 0.0}
    if not args.no_autotune:      # GEMM tile autotuner (outside the timed region)
        t0 = time.perf_counter()
        tune["shapes"] = plan.autotune(3)
        tune["seconds"] = time.perf_counter() - t0
        print(f"# autotuned {tune['shapes']} GEMM shapes in {tune['seconds']:.1f} s", file=sys.stderr,
              flush=True)
    for _ in range(args.warmup):
        plan.run(4, 0.005, 10.0)
    plan.set_timing(True)       # stage events over the timed steps only
    t0 = time.perf_counter()
    results = None
    for _ in range(args.steps):
        results = plan.run(4, 0.005, 10.0)
    t1 = time.perf_counter()
    stages = plan.stage_times()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = len(kpts) / (ms_step / 1e3)
    print(f"# batched: {ms_step:.2f} ms/step, {value:.1f} k-solves/s", file=sys.stderr, flush=True)
    plan.close()

    # ---- plain FP64: one k-point per launch sequence (sample) ----------------
    ns = min(args.plain_sample, len(kpts))
    pplain = Plan(st, kpts[:ns], w[:ns] * len(kpts) / ns, batch=1, precision=TBK_FP64)
    pplain.run(4, 0.005, 10.0)
    t0 = time.perf_counter()
    pplain.run(4, 0.005, 10.0)
    plain_rate = ns / (time.perf_counter() - t0)
    pplain.close()
    print(f"# plain fp64: {plain_rate:.1f} k-solves/s", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel --------------------------------------
    # stage times are summed over the timed steps (one launch of each
    # single-kernel stage per step for the batched plan)
    n = st.nv
    model = stage_model(n)
    per_step = {k: v / args.steps for k, v in stages.items()}
    dom = max(per_step, key=lambda k: per_step[k])
    bound, work_per_k = model.get(dom, ("mfma", 0.0))
    t_launch = per_step[dom] / 1e3
    if bound == "hbm":
        achieved = work_per_k * len(kpts) / t_launch / 1e9
        peak, unit = HBM_PEAK_GBS, "GB/s"
    else:
        achieved = work_per_k * len(kpts) / t_launch / 1e12
        peak = FP32_PEAK_TFLOPS if (prec == TBK_MIXED and dom in ("hetrd", "backtransform")) else FP64_PEAK_TFLOPS
        unit = "TFLOP/s"
    roofline = {"kernel": dom, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                "frac": achieved / peak, "traffic": None,
                "launch_ms": per_step[dom]}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "k-point solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == TBK_FP64 else "f32+f64",
        "data": "synthetic (seeded NRL parameters and displacements)",
        "config": {"workload": "C2: 32-atom fcc 2x2x2 supercell, a=3.61 A, sigma=0.05 A, "
                               "sp3d5 N=288 complex128, MP 8x8x8 = 512 k, mode 4 (E_tot + P), "
                               "kT=0.005 Ry, all k batched",
                   "natoms": st.natoms, "N": n, "nk": len(kpts),
                   "precision": args.precision, "parallelism": "single GPU"},
        "plain_fp64_ksolves_per_s": plain_rate,
        "speedup_vs_plain": value / plain_rate,
        "stages_ms_per_step": per_step,
        "autotune": tune,
        "result": results,
        "roofline": roofline,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st, kpts, w)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
