"""Benchmark: KKT solves/sec (+ p50/p99 latency) of the plain drop-in path on
one MI355X, on the 10,000-bus / 13,000-branch same-pattern sequence (20 steps,
gamma = 1e5, Ds spread growing from 1e-2..1e2 to 1e-4..1e8).

One "step" is one HykktSolver::execute() of the next system of the sequence
(numeric refactorisation + Schur CG + recovery + residual check; the symbolic
analysis is done once, during warmup).  Inputs are resident in HBM before the
timed region starts (inputs_on_device).  Every timed solve must reach a
relative KKT residual <= 1e-8 over all four block rows.

    python bench.py [--gpus 1] [--steps 20] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Hardware queues.  The served stream runs one solver slot (two streams) per
# resident pattern concurrently, 16 slots; with fewer queues than streams, slots
# share queues and their kernels serialise (at 16 queues the served stream ran
# at 585 solves/s, at 32 at 1,127).  The plain path, one solve at a time, is
# slower with many queues (its short launches wait longer: 191 solves/s at 4
# or 16, 96 at 32), and so is any process while another process on the GPU
# holds many queues.  So the served and batched legs run first, in a child
# process at CONCURRENT_QUEUES, before this process starts the HIP runtime;
# the headline sequence and the plain legs then run here at HIP's default of
# PLAIN_QUEUES: each form in the configuration it is deployed in, alone on the
# GPU.  Set before the HIP runtime starts.
PLAIN_QUEUES = "4"
CONCURRENT_QUEUES = "32"
os.environ["GPU_MAX_HW_QUEUES"] = (CONCURRENT_QUEUES if "--concurrent-legs" in sys.argv
                                   else PLAIN_QUEUES)


# The headline metric this benchmark reports.
METRIC = "KKT solves/sec + p50/p99 latency at rel. resid \u22641e-8, served vs plain, 1 GPU"


def metric_name():
    return METRIC


# MI355X published peaks: FP64 (vector and MFMA) and HBM3E.
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# Kernels whose bound is the FP64 pipes; the rest are priced against HBM.
COMPUTE_BOUND = {"k_dense_nt<128>", "k_dense_nt<32>", "k_potrf_inv"}


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, each scaled against a copy kernel of
    known size), or None if the summary has no line for it."""
    path = os.path.join(ROOT, "profiles", "grid10k_seq20_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(f"hbm_bytes_per_launch {kernel}:"):
                return float(line.split(":", 1)[1])
    return None


def roofline(kt, steps):
    """Roofline entry for the kernel with the largest share of the timed
    region: algorithmic work per launch over its mean event-timed duration."""
    if not kt:
        return None
    name, v = max(kt.items(), key=lambda kv: kv[1]["ms"])
    launches = max(v["launches"], 1)
    sec = 1e-3 * v["ms"] / launches
    if name in COMPUTE_BOUND:
        achieved, peak, unit, bound = v["flops"] / launches / sec / 1e12, PEAK_FP64_TFLOPS, \
            "TFLOP/s", "mfma"
    else:
        achieved, peak, unit, bound = v["bytes"] / launches / sec / 1e9, PEAK_HBM_GBS, "GB/s", \
            "hbm"
    return {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": traffic_per_launch(name),
            "algorithmic_flops_per_launch": v["flops"] / launches,
            "algorithmic_bytes_per_launch": v["bytes"] / launches,
            "avg_launch_us": 1e6 * sec, "launches_per_step": v["launches"] / steps,
            "share_of_step": v["ms"] / steps}


def kernel_entry(name, v, steps):
    """Per-step launches and time of one instrumented kernel, its algorithmic
    work per launch, and the rate that work runs at against its bound."""
    launches = max(v["launches"], 1)
    sec = 1e-3 * v["ms"] / launches
    e = {"launches_per_step": v["launches"] / steps, "ms_per_step": v["ms"] / steps,
         "avg_us": 1e6 * sec, "algorithmic_flops_per_launch": v["flops"] / launches,
         "algorithmic_bytes_per_launch": v["bytes"] / launches}
    if name in COMPUTE_BOUND:
        e["tflops"] = v["flops"] / launches / sec / 1e12 if sec > 0 else None
        e["frac_of_peak"] = e["tflops"] / PEAK_FP64_TFLOPS if sec > 0 else None
    else:
        e["gbs"] = v["bytes"] / launches / sec / 1e9 if sec > 0 else None
        e["frac_of_peak"] = e["gbs"] / PEAK_HBM_GBS if sec > 0 else None
    return e


def cpu_model():
    """The host CPU's model name and logical CPU count (from /proc/cpuinfo)."""
    model, count = None, 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model is None:
                    model = line.split(":", 1)[1].strip()
                if line.startswith("processor"):
                    count += 1
    except OSError:
        pass
    return model, count


def cpu_baseline(g, perm, nsample, gpu_solutions):
    """The CPU oracle (supernodal Cholesky, OpenMP) timed on the first
    `nsample` systems of the same sequence with the same ordering; only its
    numeric phase counts, as the GPU path's analysis is also done once.  Each
    oracle solution is also the reference the GPU's solution of that system
    (gpu_solutions[t]) is measured against."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from hykkt_amd.generator import ds_spread_schedule

    oc = lambda b: O.Csr(b.n, b.m, b.rp, b.ci, b.v)
    t_num, worst, errs = 0.0, 0.0, []
    for t in range(nsample):
        sy = g.system(t, ds_range=ds_spread_schedule(t, 20))
        r = O.execute(oc(sy.H), sy.Ds.v, oc(sy.Jc), oc(sy.Jd), sy.rx, sy.rs, sy.ry, sy.ryd,
                      g.gamma, perm=perm, chol_kind=1)
        t_num += r["stats"]["t_numeric"]
        worst = max(worst, r["stats"]["resid4"])
        xo = np.concatenate([r["dx"], r["ds"], r["dyc"], r["dyd"]])
        if t in gpu_solutions:
            errs.append(float(np.linalg.norm(gpu_solutions[t] - xo) / np.linalg.norm(xo)))
        print(f"bench: cpu oracle solve {t}: {r['stats']['t_numeric']:.2f} s "
              f"resid4={r['stats']['resid4']:.2e} gpu-vs-oracle="
              f"{errs[-1] if errs else float('nan'):.2e}", file=sys.stderr, flush=True)
    model, logical = cpu_model()
    return {"value": nsample / t_num, "unit": "KKT solves/s", "cores": O.threads(),
            "kind": "port", "cpu_model": model, "logical_cpus_visible": logical,
            "sample": f"{nsample} solves (steps 0..{nsample - 1}) of the same 20-step sequence, "
                      "oracle/hykkt_oracle.c execute() with its supernodal left-looking "
                      "Cholesky, same ordering, numeric phase only",
            "max_resid4": worst, "gpu_err_vs_oracle": errs}


def served_vs_plain(rate, seconds, npat=16, budget=200e9, warmup=5.0, legs=("served", "plain")):
    """The served form (continuous batching over 16 resident patterns of 1k,
    5k and 20k buses, 6:3:1) and the plain path (one request at a time through
    the drop-in solver of each pattern) on the same Poisson request stream.
    Both are over capacity at 2,000 requests/s: the figures are saturation
    throughput and queueing latency over `seconds`, p50/p99 over the requests
    that arrived after `warmup` seconds.  `legs`: which of the two run here."""
    from hykkt_amd.served import Server, make_patterns, poisson_requests, summarize
    t0 = time.perf_counter()
    pats = make_patterns(npat, seed=100)
    srv = Server(pats, hbm_budget=budget, sweep_share=8)
    setup = time.perf_counter() - t0
    print(f"bench: served setup {setup:.1f} s, {len(srv.served)} patterns, "
          f"{srv.hbm_used / 1e9:.1f} GB, capacities {srv.capacity}", file=sys.stderr, flush=True)
    reqs = poisson_requests(pats, rate, seconds, seed=7, served=srv.served)
    res = {}
    for mode in legs:
        res[mode] = summarize(*srv.run(reqs, plain=(mode == "plain"), deadline=seconds),
                              patterns=pats, solo_s=srv.solo_s, warmup=warmup)
        print(f"bench: {mode} stream: {res[mode]}", file=sys.stderr, flush=True)
    srv.close()
    for p in pats:
        p.free()
    return {"stream": f"Poisson {rate:.0f} requests/s for {seconds:.0f} s over {npat} resident "
                      "patterns (1k/5k/20k buses, 6:3:1); a pattern's i-th request is its "
                      f"interior-point step i, values from a pool of {len(pats[0].pool)} fresh "
                      "perturbations per pattern with the Ds spread growing with the step, "
                      "gamma 1e5; requests past the window stay pending",
            "capacity_by_pattern": {str(pats[p].nbus) + "/" + str(p): c
                                    for p, c in srv.capacity.items()},
            "patterns_served": len(srv.served), "hbm_used_gb": srv.hbm_used / 1e9,
            "setup_s": setup, "served": res.get("served"), "plain": res.get("plain"),
            "solo_ms": {str(pats[p].nbus) + "/" + str(p): 1e3 * t for p, t in srv.solo_s.items()}}


def ratios(served, batched):
    """Served over plain and batched over plain, once both legs are in."""
    if served and served.get("served") and served.get("plain"):
        sp, pp = served["served"]["solves_per_s"], served["plain"]["solves_per_s"]
        sw, pw = served["served"]["solo_s_per_s"], served["plain"]["solo_s_per_s"]
        served["served_over_plain"] = sp / pp if pp else None
        served["served_over_plain_solo_weighted"] = sw / pw if pw else None
    if batched and batched.get("batched") and batched.get("plain"):
        # systems within the residual bound per second, batched over plain
        batched["batched_over_plain"] = (batched["batched"]["solves_per_s"] /
                                         batched["plain"]["solves_per_s"])


def batched_vs_plain(legs=("batched", "plain")):
    """The batched form (256 systems: 32 patterns of 200..5000 buses x 8 value
    draws, a gamma per system; each pattern's 8 draws in one pass of the batch
    engine, the patterns' passes concurrent) against the same systems one at a
    time through the drop-in path.  `legs`: which of the two run here."""
    from hykkt_amd.batched import make_batch, make_server, run_batch, summarize_batch
    t0 = time.perf_counter()
    pats = make_batch(seed=3)
    srv = make_server(pats, hbm_budget=200e9)
    setup = time.perf_counter() - t0
    if len(srv.served) != len(pats):
        raise RuntimeError(f"batched: slots refused: {srv.refused}")   # over the budget
    res = {}
    for mode in legs:
        r, wall = run_batch(srv, pats, plain=(mode == "plain"))
        res[mode] = summarize_batch(r, wall, pats)
        print(f"bench: {mode} batch: {res[mode]}", file=sys.stderr, flush=True)
    srv.close()
    for p in pats:
        p.free()
    return {"batch": "256 systems: 32 patterns (200..5000 buses, branches 1.3x) x 8 value draws, "
                     "gamma log-uniform 1e4..1e6 per system, Ds log-uniform 1e-4..1e4",
            "hbm_used_gb": srv.hbm_used / 1e9, "setup_s": setup,
            "batched": res.get("batched"), "plain": res.get("plain")}


def concurrent_legs(args):
    """The served and batched legs of the two comparisons, in a child process
    at CONCURRENT_QUEUES hardware queues; their summaries (JSON on the child's
    stdout, its progress lines on stderr).  Called before this process starts
    the HIP runtime, so the child has the GPU to itself."""
    import subprocess
    env = dict(os.environ, GPU_MAX_HW_QUEUES=CONCURRENT_QUEUES)
    cmd = [sys.executable, os.path.abspath(__file__), "--concurrent-legs", "1",
           "--served-seconds", str(args.served_seconds), "--served-rate", str(args.served_rate),
           "--batched", str(args.batched)]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"served and batched legs: child exited with {r.returncode}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def reweight(leg, solo_ms):
    """Solo-weighted throughput of a served-stream leg with the solo times
    `solo_ms` (keys "buses/pattern"): both legs are weighted with the solo
    times measured where one solve runs at a time."""
    solo = {k.split("/")[1]: v / 1e3 for k, v in solo_ms.items()}
    done = leg["within_tol_by_pattern"]
    leg["solo_s_per_s"] = (sum(solo[p] * c for p, c in done.items() if p in solo) / leg["wall_s"]
                           if leg["wall_s"] > 0 else 0.0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="grid10k_seq20")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=3,
                    help="solves of the sequence the CPU oracle is timed on")
    ap.add_argument("--kernel-timing", type=int, default=1)
    ap.add_argument("--batched", type=int, default=1,
                    help="run the 256-system batch, batched vs one at a time")
    ap.add_argument("--served-seconds", type=float, default=60.0,
                    help="length of the served and plain request streams (0: skip)")
    ap.add_argument("--served-rate", type=float, default=2000.0,
                    help="Poisson arrival rate of the served stream (requests/s)")
    ap.add_argument("--concurrent-legs", type=int, default=0,
                    help="(internal) run only the served and batched legs of the two "
                         "comparisons and print their summaries")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run the plain copy kernel over this many MiB (3 times), the "
                         "known-size reference for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the KKT solve of one "
                                   "system does not shard across GPUs"}))
        sys.exit(2)

    if args.concurrent_legs:
        legs = {"served": None, "batched": None}
        if args.served_seconds > 0:
            legs["served"] = served_vs_plain(args.served_rate, args.served_seconds,
                                             legs=("served",))
        if args.batched:
            legs["batched"] = batched_vs_plain(legs=("batched",))
        print(json.dumps(legs), flush=True)
        return
    # the served and batched legs first, while no process here holds the GPU
    conc = (concurrent_legs(args) if args.served_seconds > 0 or args.batched
            else {"served": None, "batched": None})

    import hykkt_amd as H
    from hykkt_amd.generator import ds_spread_schedule, workload

    g = workload(args.workload)
    nseq = 20
    print(f"bench: generating {args.workload} n={g.n} m={g.m} d={g.d}", file=sys.stderr,
          flush=True)
    to_d = lambda a, t=np.float64: H.DeviceArray(np.asarray(a, dtype=t))
    systems = []
    for t in range(nseq):
        rng = ds_spread_schedule(t, nseq) if args.workload == "grid10k_seq20" else None
        sy = g.system(t, ds_range=rng)
        systems.append(dict(H=to_d(sy.H.v), Jc=to_d(sy.Jc.v), Jd=to_d(sy.Jd.v), Ds=to_d(sy.Ds.v),
                            rx=to_d(sy.rx), rs=to_d(sy.rs), ry=to_d(sy.ry), ryd=to_d(sy.ryd)))
    i32 = lambda a: to_d(a, np.int32)
    pat = dict(Hrp=i32(g.H.rp), Hci=i32(g.H.ci), Jcrp=i32(g.Jc.rp), Jcci=i32(g.Jc.ci),
               Jdrp=i32(g.Jd.rp), Jdci=i32(g.Jd.ci), Dsrp=i32(np.arange(g.d + 1)),
               Dsci=i32(np.arange(g.d)))
    H.device_synchronize()

    if args.copy_calibration_mb > 0:
        cnt = args.copy_calibration_mb * (1 << 20) // 8
        src, dst = H.DeviceArray(np.ones(cnt)), H.DeviceArray(np.zeros(cnt))
        for _ in range(3):
            H.device_copy(dst, src)
        H.device_synchronize()
        print(f"bench: copy calibration k_copy {cnt} doubles = {16 * cnt} bytes moved per launch",
              file=sys.stderr, flush=True)
        src.free()
        dst.free()

    s = H.HykktSolver(g.gamma)
    s.options(inputs_on_device=1)

    def bind(k):
        v = systems[k % nseq]
        s.set_H(pat["Hrp"], pat["Hci"], v["H"], g.n, g.n, g.H.nnz)
        s.set_Ds(pat["Dsrp"], pat["Dsci"], v["Ds"], g.d, g.d, g.d)
        s.set_Jc(pat["Jcrp"], pat["Jcci"], v["Jc"], g.m, g.n, g.Jc.nnz)
        s.set_Jd(pat["Jdrp"], pat["Jdci"], v["Jd"], g.d, g.n, g.Jd.nnz)
        s.set_rx(v["rx"], g.n)
        s.set_rs(v["rs"], g.d)
        s.set_ry(v["ry"], g.m)
        s.set_ryd(v["ryd"], g.d)

    for k in range(args.warmup):
        bind(k)
        t0 = time.perf_counter()
        s.execute()
        st = s.stats()
        print(f"bench: warmup {k} {1e3 * (time.perf_counter() - t0):.1f} ms resid4="
              f"{st['resid4']:.2e} cg={st['cg_iters']} factor={st['t_factor_ms']:.1f}ms "
              f"cg={st['t_cg_ms']:.1f}ms sym={st['t_symbolic_ms']:.0f}ms", file=sys.stderr,
              flush=True)

    def timed_pass(profile):
        """K executes over the next systems of the sequence between two device
        synchronizations; with `profile`, HIP events bracket every factor and
        solve launch (a few microseconds each, so this pass runs slower)."""
        r = dict(lat=[], worst=0.0, iters=[], tf=[], tcg=[], tref=[], nref=[], first=0.0, firsts=[])
        if profile:
            s.profile(True)
        H.device_synchronize()
        t_start = time.perf_counter()
        for k in range(args.steps):
            bind(args.warmup + k)
            t0 = time.perf_counter()
            rc = s.execute()
            r["lat"].append(time.perf_counter() - t0)
            st = s.stats()
            r["worst"] = max(r["worst"], st["resid4"])
            r["iters"].append(st["cg_iters"])
            r["tf"].append(st["t_factor_ms"])
            r["tcg"].append(st["t_cg_ms"])
            r["tref"].append(st["t_refine_ms"])
            r["nref"].append(st["refine_steps"])
            r["first"] = max(r["first"], st["resid4_first"])
            r["firsts"].append(st["resid4_first"])
            if rc != 0 or not st["resid4"] <= 1e-8:
                print(f"bench: step {k} rc={rc} resid4={st['resid4']:.3e}", file=sys.stderr,
                      flush=True)
        H.device_synchronize()
        r["total"] = time.perf_counter() - t_start
        r["kt"] = s.kernel_times() if profile else {}
        s.profile(False)
        return r

    # the measured pass runs the plain path with nothing added; the per-kernel
    # figures come from a second pass over the same systems with event timing
    main_pass = timed_pass(False)
    perm0 = s.permutation()
    prof_pass = timed_pass(True) if args.kernel_timing else None
    # the GPU's solutions of the systems the CPU oracle solves (its sample)
    gpu_sol = {}
    if args.cpu_baseline:
        for t in range(args.cpu_sample):
            bind(t)
            s.execute()
            gpu_sol[t] = np.concatenate(s.solution())
    kt = prof_pass["kt"] if prof_pass else {}
    lat, worst, iters, tf, tcg, nref, first, total = (
        main_pass[k] for k in ("lat", "worst", "iters", "tf", "tcg", "nref", "first", "total"))
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": metric_name(),
        "value": args.steps / total,
        "unit": "KKT solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded OPF-like generator, hykkt_amd/generator.py)",
        "config": {"workload": f"{args.workload}: 10,000 buses / 13,000 branches, "
                               f"n={g.n} m={g.m} d={g.d}, gamma=1e5, 20-step same-pattern "
                               "sequence, plain drop-in path",
                   "parallelism": "single GPU"},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "max_resid4": worst,
        "cg_iters_mean": float(np.mean(iters)),
        "factor_ms_mean": float(np.mean(tf)),
        "cg_ms_mean": float(np.mean(tcg)),
        "refine_steps_mean": float(np.mean(nref)),
        "refine_ms_mean": float(np.mean(main_pass["tref"])),
        "max_resid4_before_refinement": first,
        "resid4_before_refinement_quantiles": [float(x) for x in np.percentile(
            main_pass["firsts"], [0, 25, 50, 75, 100])],
        "roofline": roofline(kt, args.steps),
        "kernel_timing_pass": None if prof_pass is None else {
            "ms_per_step": 1e3 * prof_pass["total"] / args.steps,
            "max_resid4": prof_pass["worst"],
            "note": "same K executes again with HIP events around every factor and solve "
                    "launch; roofline and kernels come from this pass"},
        "kernels": {k: kernel_entry(k, v, args.steps) for k, v in kt.items()},
        "served": None,
        "batched": None,
        "cpu_baseline": None,
        "err_vs_oracle": None,
    }
    if args.served_seconds > 0 or args.batched:
        s.close()
    note = (f"served and batched legs first, in a child process at GPU_MAX_HW_QUEUES="
            f"{CONCURRENT_QUEUES}; the headline sequence and the plain legs after it, in this "
            f"process at {PLAIN_QUEUES} (HIP's default); one process on the GPU at a time")
    if args.served_seconds > 0:
        pl = served_vs_plain(args.served_rate, args.served_seconds, legs=("plain",))
        sv = conc["served"]
        sv["plain"] = pl["plain"]
        sv["solo_ms_concurrent_process"] = sv["solo_ms"]
        sv["solo_ms"] = pl["solo_ms"]
        for leg in ("served", "plain"):
            reweight(sv[leg], pl["solo_ms"])
        sv["queues"] = note
        out["served"] = sv
    if args.batched:
        sb = conc["batched"]
        sb["plain"] = batched_vs_plain(legs=("plain",))["plain"]
        sb["queues"] = note
        out["batched"] = sb
    ratios(out["served"], out["batched"])
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, perm0, args.cpu_sample, gpu_sol)
        e = out["cpu_baseline"]["gpu_err_vs_oracle"]
        out["err_vs_oracle"] = {
            "max_rel": max(e) if e else None,
            "what": "||x_gpu - x_oracle|| / ||x_oracle|| over (dx, ds, dyc, dyd), same ordering, "
                    f"on steps 0..{len(e) - 1} of the sequence (the CPU baseline's sample)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
