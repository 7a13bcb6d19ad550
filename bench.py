#!/usr/bin/env python3
"""Benchmark: k-point solves/s at N = 288 (32-atom fcc supercell, MP 8x8x8).

One step = one pass of the per-k hot path over the whole structure: for
each of the 512 k-points build H(k), S(k), solve H c = e S c with vectors,
form the first-order pressure derivatives, then the Fermi search and the
band-energy / pressure sums (mode 4).  Inputs are synthetic (seeded NRL
parameters, seeded displacements) and already resident on the GPU when the
timed region starts.

Both precisions are measured (FP64, and mixed: FP32 tridiagonalisation
and back-transform with FP64 Ogita-Aishima refinement); the headline value
is --precision (default mixed).  Prints one JSON line with
  roofline      the headline step's Householder reduction: k_hetrd_res
                (resident tridiagonalisation, one launch per step) as 16/3 N^3
                flops per k over the launch's event time against the matrix
                peak of the precision, or k_hetrd_panel (one-stage panels) as
                algorithmic bytes per launch over its average duration (HIP
                events on the plan's stream around each launch) against the
                HBM peak; traffic from the committed counter summaries;
  roofline_panel_fp64  k_hetrd_panel of the FP64 step (its dominant kernel);
  roofline_build  the H/S build: 2 * 16 * N^2 bytes per k written over the
                build stage's event time, against the HBM peak;
  roofline_solve  per precision, the eigensolve's LAPACK flop count (22.67 N^3
                per k) over the summed event time of the solve stages,
                against the FP64 (plain) or FP32 (mixed) matrix-core peak;
  c3, c4        the 108-atom cell (N = 972, 64 k) and the 256-structure
                uneven batch, both precisions (--no-extra skips them);
  serving       Poisson request streams (10, 100, 500 req/s) through the
                continuous-batching request loop: p50/p99 latency;
  cpu_baseline  the CPU restatement (oracle) timed on this host's cores (at
                most 16) on a bounded sample of the same workload, and on
                one core; the CPU model.
--profile-pass PREC runs one batched step of PREC (C2, or C3 with
--workload c3) and a 1 GiB device copy (the calibration kernel for
counter-based traffic) and exits; it is the program profiled with rocprofv3.
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


def measured_traffic(precision: str, profile: str = "c2_hetrd_panel_traffic.txt"):
    """HBM bytes per launch of a kernel from its committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the same counter
    reads for a 1 GiB copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        for line in open(path):
            if line.startswith(f"{precision}_traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def panel_roofline(n, nk, stages, precision):
    """k_hetrd_panel (one-stage Householder panels, HBM-bound Hermitian
    products): algorithmic bytes per launch over its average launch time
    (HIP events on the plan's stream around each launch)."""
    csize = 16 if precision == "fp64" else 8
    panel_ms = stages.get("hetrd_panel", float("nan"))
    launches = (n - 2) // 32 + 1
    bytes_step = hetrd_panel_bytes(n, csize) * nk
    achieved = bytes_step / (panel_ms / 1e3) / 1e9
    return {"kernel": "k_hetrd_panel", "precision": precision, "bound": "hbm", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(precision),
            "traffic_source": "profiles/c2_hetrd_panel_traffic.txt (rocprofv3 --pmc FETCH_SIZE, "
                              "WRITE_SIZE of the same step, scaled by a 1 GiB copy); not measured in this run",
            "launches_per_step": launches, "avg_launch_ms": panel_ms / launches,
            "algorithmic_bytes_per_launch": bytes_step / launches}


def resident_roofline(n, nk, stages, precision):
    """k_hetrd_res (the resident tridiagonalisation: the whole reduction of
    each k-point in one workgroup, the matrix held in registers; one launch
    per step): the Householder reduction's 16/3 N^3 real flops per k over
    the launch's event time, against the matrix peak of the precision (the
    rank-2 updates run on the matrix cores, the Hermitian products on the
    vector units)."""
    ms = stages.get("hetrd_resident", float("nan"))
    flops = 16.0 / 3.0 * n ** 3 * nk
    tf = flops / (ms / 1e3) / 1e12
    peak = FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS
    traffic = measured_traffic(precision, "c2_hetrd_res_traffic.txt")
    return {"kernel": "k_hetrd_res", "precision": precision, "bound": "mfma", "achieved": tf, "peak": peak,
            "unit": "TFLOP/s", "frac": tf / peak, "traffic": traffic,
            "traffic_source": "profiles/c2_hetrd_res_traffic.txt (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of "
                              "the same step, scaled by a 1 GiB copy); not measured in this run",
            "hbm_gbs": (traffic / (ms / 1e3) / 1e9) if traffic else None,
            "launches_per_step": 1, "avg_launch_ms": ms, "algorithmic_flops_per_launch": flops}


def hetrd_panel_bytes(n: int, csize: int) -> float:
    """Algorithmic HBM bytes per k-point of all Householder panel launches
    (k_hetrd_panel) at order n with complex elements of csize bytes: per
    column j the lower triangle of the trailing matrix (the Hermitian
    product), the panel's V and W rows twice (dots and the fused update),
    the column itself, the stores of v (A, V, V2) and w, and per panel the
    V T^H block."""
    tot = 0.0
    for j in range(n - 1):
        m = n - j - 1
        i = j % 32
        tot += m * (m + 1) / 2 + 4 * i * m + (m + 1) + 4 * m
    panels = (n - 2) // 32 + 1
    tot += panels * 2 * 32 * n
    return tot * csize


# LAPACK's flop counts for the generalized eigensolve with vectors, per k
# at order N: ZPOTRF 4/3, ZHEGST 4, ZHETRD 16/3, apply Q 8 and the
# triangular back-transform 4 (in N^3 real flops).
SOLVE_FLOP_PER_N3 = 4.0 / 3.0 + 4.0 + 16.0 / 3.0 + 8.0 + 4.0
SOLVE_STAGES = ("potrf_trtri", "reduce_gemm", "hetrd_panel", "hetrd_update", "hetrd_resident", "hetrd_band",
                "hetrd_bulge", "stein",
                "backtransform", "refine")


def solve_roofline(n, nk, stages, precision):
    """The eigensolve against the matrix-core peak of its precision:
    22.67 N^3 flops per k over the summed time of the solve stages (HIP
    events on the plan's stream), FP64 peak for the plain method, FP32 peak
    for the mixed one (whose tridiagonalisation and back-transform run in
    FP32)."""
    ms = sum(stages.get(k, 0.0) for k in SOLVE_STAGES)
    flops = SOLVE_FLOP_PER_N3 * n ** 3 * nk
    tf = flops / (ms / 1e3) / 1e12 if ms > 0 else float("nan")
    peak = FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS
    return {"bound": "mfma", "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak,
            "algorithmic_flops": flops, "solve_ms": ms, "stages": [k for k in SOLVE_STAGES if k in stages]}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(st, kpts, weights, seconds=12.0):
    """The oracle on this host's cores (at most 16, the host share of one
    GPU) over a bounded sample of k-points, and on one core over a smaller
    sample."""
    import oracle
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = max(1, min(avail, 16))
    setup = oracle.Setup(st)
    t0 = time.perf_counter()
    setup.bande(kpts[:threads], weights[:threads], mode=4, nthreads=threads)
    dt = time.perf_counter() - t0
    # scale the sample to about `seconds` of work
    reps = max(1, int(seconds / max(dt, 1e-3)))
    nk = min(len(kpts), threads * reps)
    t0 = time.perf_counter()
    setup.bande(kpts[:nk], weights[:nk], mode=4, nthreads=threads)
    dt = time.perf_counter() - t0
    n1 = 8
    t0 = time.perf_counter()
    setup.bande(kpts[:n1], weights[:n1], mode=4, nthreads=1)
    dt1 = time.perf_counter() - t0
    return {"value": nk / dt, "unit": "k-point solves/s", "cores": threads, "kind": "port",
            "sample": f"{nk} of the 512 k-points of the same structure (N={st.nv}), "
                      f"full per-k path incl. pressure, OpenMP over k, {dt:.1f} s",
            "value_1core": n1 / dt1, "sample_1core": f"{n1} k-points on one thread, {dt1:.1f} s",
            "host_cpus_visible": os.cpu_count(), "cpu_model": cpu_model()}


def run_batched(st, kpts, w, prec, steps, warmup, autotune=True, tridiag=0):
    """All k-points in one batch; returns timing, stage times, results and
    the plan creation time (host set-up, uploads, allocation)."""
    t0 = time.perf_counter()
    plan = Plan(st, kpts, w, batch=0, precision=prec)
    create_ms = (time.perf_counter() - t0) * 1e3
    if tridiag:
        plan.set_tridiag(tridiag)
    tune = {"shapes": 0, "seconds": 0.0}
    if autotune:                  # GEMM tile autotuner (outside the timed region)
        t0 = time.perf_counter()
        tune["shapes"] = plan.autotune(3)
        tune["seconds"] = time.perf_counter() - t0
    for _ in range(warmup):
        plan.run(4, 0.005, 10.0)
    # the timed steps: the captured launch sequence replayed (the capture
    # itself is one more untimed step)
    plan.set_graph(True)
    plan.run(4, 0.005, 10.0)
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = plan.run(4, 0.005, 10.0)
    t1 = time.perf_counter()
    plan.set_graph(False)
    # stage times: the same steps launched eagerly with events between the
    # stages (which a captured sequence does not carry)
    plan.set_timing(True)
    t2 = time.perf_counter()
    for _ in range(steps):
        res_eager = plan.run(4, 0.005, 10.0)
    t3 = time.perf_counter()
    stages = {k: v / steps for k, v in plan.stage_times().items()}
    plan.close()
    if res_eager != res:
        raise RuntimeError(f"graph replay and eager launch disagree: {res} vs {res_eager}")
    ms = (t1 - t0) * 1e3 / steps
    return {"ms_per_step": ms, "ksolves_per_s": len(kpts) / (ms / 1e3), "stages_ms_per_step": stages,
            "eager_timed_ms_per_step": (t3 - t2) * 1e3 / steps,
            "result": res, "autotune": tune, "plan_create_ms": create_ms}


def serve_stream(rates, nreq, precision, seed=20261203, inflight=8):
    """Poisson streams of E_tot + pressure requests: structures as in the
    uneven batch up to 128 atoms (N <= 1152), Monkhorst-Pack grids up to
    6^3.  Per rate: latency percentiles, served rate, the device's busy
    fraction and the work rate (sum of nv^3 nk per second)."""
    from tbk.api import Server
    pool = S.config_c4(seed=seed, count=256, max_atoms=128, max_q=6, nk_budget=9 * 6 ** 3)
    rng = np.random.default_rng(seed)
    out = []
    for rate in rates:
        srv = Server(precision=precision, budget_bytes=96 << 30, max_inflight=inflight)
        picks = rng.integers(len(pool), size=nreq)
        gaps = rng.exponential(1.0 / rate, size=nreq)
        reqs = []
        for i in picks:
            st, q = pool[int(i)]
            k = S.monkhorst_pack((q, q, q))
            reqs.append((st, k, np.full(len(k), 1.0 / len(k))))
        ids = []
        t_start = time.perf_counter()
        t_next = t_start
        for (st, k, w), g in zip(reqs, gaps):
            t_next += g
            while time.perf_counter() < t_next:
                time.sleep(min(0.001, max(0.0, t_next - time.perf_counter())))
            ids.append(srv.submit(st, k, w))
        t_sub = time.perf_counter()
        res = [srv.result(r) for r in ids]
        t_end = time.perf_counter()
        stats = srv.stats()
        tm = srv.timing()
        srv.close()
        lat = np.array([r["latency_ms"] for r in res])
        bad = sum(1 for r in res if r["status"] != 0)
        nk = sum(len(k) for _, k, _ in reqs)
        p99 = float(np.percentile(lat, 99))
        span = tm["span_ms"] / 1e3
        out.append({"rate_req_per_s": rate, "requests": nreq, "failed": bad,
                    "offered_req_per_s": nreq / (t_sub - t_start),
                    "p50_ms": float(np.percentile(lat, 50)), "p90_ms": float(np.percentile(lat, 90)),
                    "p99_ms": p99, "samples_above_p99": int((lat > p99).sum()),
                    "mean_queue_ms": float(np.mean([r["queue_ms"] for r in res])),
                    "req_per_s": nreq / (t_end - t_start), "ksolves_per_s": nk / (t_end - t_start),
                    "device_busy_frac": tm["busy_ms"] / tm["span_ms"] if tm["span_ms"] > 0 else None,
                    "work_n3k_per_s": tm["work_n3k"] / span if span > 0 else None,
                    "work_n3k_per_busy_s": tm["work_n3k"] / (tm["busy_ms"] / 1e3) if tm["busy_ms"] > 0 else None,
                    "host_setup_ms_per_req": tm["setup_ms"] / nreq,
                    "launch_ms_per_group": tm["launch_ms"] / max(1, stats["groups"]),
                    "groups": stats["groups"], "largest_group": stats["max_group"]})
        o = out[-1]
        print(f"# serve {rate:5.0f} req/s: served {o['req_per_s']:.1f} req/s, p50 {o['p50_ms']:.1f} ms "
              f"p99 {o['p99_ms']:.1f} ms, device busy {o['device_busy_frac']:.2f}, "
              f"{o['work_n3k_per_s'] / 1e9:.1f} G N^3k/s ({stats['groups']} groups, largest {stats['max_group']})",
              file=sys.stderr, flush=True)
    return out


def profile_pass(precision, tridiag=0, workload="c2"):
    """The GEMM autotuning pass, one warm-up and one batched step of C2 (or
    C3), then a 1 GiB copy kernel (read 1 GiB, write 1 GiB) to calibrate the
    traffic counters."""
    st = S.config_c3() if workload == "c3" else S.config_c2()
    kpts = S.monkhorst_pack((4, 4, 4) if workload == "c3" else (8, 8, 8))
    w = np.full(len(kpts), 1.0 / len(kpts))
    plan = Plan(st, kpts, w, batch=0, precision=TBK_MIXED if precision == "mixed" else TBK_FP64)
    if tridiag:
        plan.set_tridiag(tridiag)
    plan.autotune(3)                  # the same GEMM tiles as the timed bench
    plan.run(4, 0.005, 10.0)
    r = plan.run(4, 0.005, 10.0)
    plan.close()
    from tbk import _lib
    _lib.check(_lib.lib().tbk_debug_copy(0, 1 << 30), "tbk_debug_copy")   # k_copy16: 1 GiB in, 1 GiB out
    print(json.dumps({"profile_pass": precision, "result": r, "copy_bytes": 1 << 30}))


def bench_c3(args):
    """108-atom strained supercell, N = 972, MP 4x4x4 = 64 k, both precisions."""
    st = S.config_c3()
    kpts = S.monkhorst_pack((4, 4, 4))
    w = np.full(len(kpts), 1.0 / len(kpts))
    out = {"workload": "C3: 108-atom fcc 3x3x3, sigma=0.03 A, <=1% strain, N=972, MP 4^3 = 64 k, mode 4",
           "unit": "k-point solves/s"}
    for name, prec in (("fp64", TBK_FP64), ("mixed", TBK_MIXED)):
        r = run_batched(st, kpts, w, prec, max(1, args.steps // 2), args.warmup, not args.no_autotune, args.tridiag)
        out[name] = {"ksolves_per_s": r["ksolves_per_s"], "ms_per_step": r["ms_per_step"],
                     "stages_ms_per_step": r["stages_ms_per_step"], "result": r["result"],
                     "roofline_solve": solve_roofline(st.nv, len(kpts), r["stages_ms_per_step"], name)}
        print(f"# C3 {name}: {r['ms_per_step']:.1f} ms/step", file=sys.stderr, flush=True)
    return out


def bench_spin(args):
    """The C2 structure (32 atoms, 512 k) with a second parameter channel
    (the first 57 values scaled by 1 + 0.02 N(0,1), seed 7): collinear spin
    polarisation (1024 solves of N = 288 per step) and non-collinear spins
    along alternating directions (512 solves of N = 576), both precisions."""
    from tbk.api import NonCollinearPlan, SpinPlan
    st = S.config_c2()
    kpts = S.monkhorst_pack((8, 8, 8))
    w = np.full(len(kpts), 1.0 / len(kpts))
    rng = np.random.default_rng(7)
    pdn = np.array(st.params.param, float)
    pdn[:57] = pdn[:57] * (1.0 + 0.02 * rng.standard_normal(57))
    t = np.deg2rad(30.0)
    spinat = [[np.sin(t), 0.0, np.cos(t)] if a % 2 else [0.0, 0.0, 1.0] for a in range(st.natoms)]
    out = {"workload": "C2 structure with two parameter channels: jspins=2 (collinear) and lnoncol, mode 4",
           "unit": "k-point solves/s (one solve = one matrix of the path: N=288 per channel, N=576 spinor)"}
    steps = max(1, args.steps // 2)
    for kind in ("collinear", "noncollinear"):
        for name, prec in (("fp64", TBK_FP64), ("mixed", TBK_MIXED)):
            if kind == "collinear":
                plan = SpinPlan(st, pdn, kpts, w, precision=prec)
                nsolve = 2 * len(kpts)
            else:
                plan = NonCollinearPlan(st, pdn, spinat, kpts, w, precision=prec)
                nsolve = len(kpts)
            if not args.no_autotune:
                plan.autotune(3)
            for _ in range(args.warmup):
                plan.run(4, 0.005, 10.0)
            t0 = time.perf_counter()
            for _ in range(steps):
                res = plan.run(4, 0.005, 10.0)
            ms = (time.perf_counter() - t0) * 1e3 / steps
            plan.close()
            out[f"{kind}_{name}"] = {"ms_per_step": ms, "solves_per_s": nsolve / (ms / 1e3), "order": plan.nv,
                                     "result": res}
            print(f"# {kind} {name}: N={plan.nv}, {ms:.1f} ms/step", file=sys.stderr, flush=True)
    print(json.dumps(out))


def bench_c4(args):
    """256 uneven structures (N = 9..576), all solved per step, both precisions."""
    from tbk.api import MultiPlan
    c4 = S.config_c4()
    sts = [st for st, _ in c4]
    ks = [S.monkhorst_pack((q, q, q)) for _, q in c4]
    ws = [np.full(len(k), 1.0 / len(k)) for k in ks]
    nk = sum(len(k) for k in ks)
    work = sum(len(k) * st.nv ** 3 for st, k in zip(sts, ks))
    out = {"workload": "C4: 256 seeded fcc/bcc/hcp structures, 1-64 atoms (N=9-576), MP 2^3-8^3, mode 4",
           "structures": len(sts), "kpoints": nk, "unit": "k-point solves/s"}
    for name, prec in (("fp64", TBK_FP64), ("mixed", TBK_MIXED)):
        m = MultiPlan(sts, ks, ws, precision=prec, graph=True)
        if not args.no_autotune:
            m.autotune(2)
        for _ in range(max(1, args.warmup)):
            m.run(4)
        t0 = time.perf_counter()
        steps = max(1, args.steps // 2)
        for _ in range(steps):
            res = m.run(4)
        dt = (time.perf_counter() - t0) / steps
        bad = sum(1 for r in res if "error" in r)
        out[name] = {"ksolves_per_s": nk / dt, "ms_per_step": dt * 1e3, "groups": m.ngroups,
                     "gib": m.device_bytes() / 2 ** 30, "failed": bad,
                     "n3k_per_s": work / dt,
                     "solve_tflops": SOLVE_FLOP_PER_N3 * work / dt / 1e12}
        m.close()
        print(f"# C4 {name}: {dt * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                    help="precision of the headline value (both are measured)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rates", default="10,20,50,100,200,300,500")
    ap.add_argument("--serve-inflight", type=int, default=8, help="in-flight groups (streams) of the server")
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 108-atom and uneven-batch workloads the default line carries as "
                         "extra keys c3 and c4")
    ap.add_argument("--profile-pass", choices=["fp64", "mixed"], default=None)
    ap.add_argument("--tridiag", type=int, choices=[0, 1, 2], default=0,
                    help="tridiagonalisation of the C2 and C3 plans: 0 the library's choice, 1 one-stage, 2 two-stage")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "spin", "serve"], default="c2",
                    help="c2: the headline line; c3 (N=972, 64 k), c4 (256 uneven "
                         "structures), spin (C2 spin-polarised and non-collinear) and serve (the "
                         "request stream sweep) print their own lines")
    args = ap.parse_args()
    if args.profile_pass:
        profile_pass(args.profile_pass, args.tridiag, args.workload)
        return
    if args.workload == "c3":
        print(json.dumps(bench_c3(args)))
        return
    if args.workload == "spin":
        bench_spin(args)
        return
    if args.workload == "c4":
        print(json.dumps(bench_c4(args)))
        return
    if args.workload == "serve":
        print(json.dumps({"serving": serve_stream([float(x) for x in args.serve_rates.split(",")],
                                                  args.serve_requests, TBK_MIXED, inflight=args.serve_inflight)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    st = S.config_c2()
    kpts = S.monkhorst_pack((8, 8, 8))
    w = np.full(len(kpts), 1.0 / len(kpts))
    n = st.nv
    print(f"# workload C2: natoms={st.natoms} N={n} nk={len(kpts)}", file=sys.stderr, flush=True)

    # ---- batched, both precisions -------------------------------------------
    runs = {}
    for name, prec in (("fp64", TBK_FP64), ("mixed", TBK_MIXED)):
        runs[name] = run_batched(st, kpts, w, prec, args.steps, args.warmup, not args.no_autotune, args.tridiag)
        print(f"# batched {name}: {runs[name]['ms_per_step']:.2f} ms/step, "
              f"{runs[name]['ksolves_per_s']:.1f} k-solves/s", file=sys.stderr, flush=True)
    head = runs[args.precision]

    # ---- plain FP64: one k-point per launch sequence (sample) ----------------
    ns = min(args.plain_sample, len(kpts))
    pplain = Plan(st, kpts[:ns], w[:ns] * len(kpts) / ns, batch=1, precision=TBK_FP64)
    pplain.run(4, 0.005, 10.0)
    t0 = time.perf_counter()
    pplain.run(4, 0.005, 10.0)
    plain_rate = ns / (time.perf_counter() - t0)
    pplain.close()
    print(f"# plain fp64: {plain_rate:.1f} k-solves/s", file=sys.stderr, flush=True)

    # ---- roofline of the headline step's Householder reduction: the
    # resident kernel where the order allows it (one launch, flops), else the
    # one-stage panels (HBM); the FP64 step's panels beside it
    hs = head["stages_ms_per_step"]
    roofline = (resident_roofline(n, len(kpts), hs, args.precision) if "hetrd_resident" in hs
                else panel_roofline(n, len(kpts), hs, args.precision))
    roofline_panel_fp64 = (panel_roofline(n, len(kpts), runs["fp64"]["stages_ms_per_step"], "fp64")
                           if "hetrd_panel" in runs["fp64"]["stages_ms_per_step"] else None)
    # ---- the H/S build (k_phases + k_build_hs, HBM-bound): H and S written
    # once, 2 * 16 * N^2 bytes per k, over the build stage's event time
    build_ms = head["stages_ms_per_step"].get("build_hs", float("nan"))
    build_bytes = 2 * 16 * n * n * len(kpts)
    build_gbs = build_bytes / (build_ms / 1e3) / 1e9
    roofline_build = {"kernel": "k_build_hs", "bound": "hbm", "achieved": build_gbs, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": build_gbs / HBM_PEAK_GBS, "algorithmic_bytes": build_bytes,
                      "stage_ms": build_ms}

    # ---- the eigensolve against the matrix-core peak, both precisions -----
    roofline_solve = {k: solve_roofline(n, len(kpts), r["stages_ms_per_step"], k) for k, r in runs.items()}
    for k, r in roofline_solve.items():
        print(f"# solve {k}: {r['achieved']:.1f} TFLOP/s = {r['frac']:.3f} of {r['peak']}", file=sys.stderr,
              flush=True)

    # ---- the other workloads, as extra keys ---------------------------------
    extra = {}
    if not args.no_extra:
        extra["c3"] = bench_c3(args)
        extra["c4"] = bench_c4(args)

    serving = None
    if not args.no_serve:
        serving = serve_stream([float(x) for x in args.serve_rates.split(",")], args.serve_requests,
                               TBK_MIXED, inflight=args.serve_inflight)

    out = {
        "metric": METRIC,
        "value": head["ksolves_per_s"],
        "unit": "k-point solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded NRL parameters and displacements)",
        "config": {"workload": "C2: 32-atom fcc 2x2x2 supercell, a=3.61 A, sigma=0.05 A, "
                               "sp3d5 N=288 complex128, MP 8x8x8 = 512 k, mode 4 (E_tot + P), "
                               "kT=0.005 Ry, all k batched",
                   "natoms": st.natoms, "N": n, "nk": len(kpts),
                   "precision": args.precision, "parallelism": "single GPU"},
        "batched_fp64_ksolves_per_s": runs["fp64"]["ksolves_per_s"],
        "batched_mixed_ksolves_per_s": runs["mixed"]["ksolves_per_s"],
        "plain_fp64_ksolves_per_s": plain_rate,
        "speedup_vs_plain": head["ksolves_per_s"] / plain_rate,
        "stages_ms_per_step": {k: r["stages_ms_per_step"] for k, r in runs.items()},
        "autotune": {k: r["autotune"] for k, r in runs.items()},
        # host set-up + upload of the structure and k-mesh, once per plan
        "plan_create_ms": {k: r["plan_create_ms"] for k, r in runs.items()},
        "result": {k: r["result"] for k, r in runs.items()},
        "roofline": roofline,
        "roofline_panel_fp64": roofline_panel_fp64,
        "roofline_build": roofline_build,
        "roofline_solve": roofline_solve,
        "serving": serving,
    }
    out.update(extra)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st, kpts, w)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
