#!/usr/bin/env python3
"""Benchmark of the MI355X Newton contact step (one GPU).

Workload: 2000 icospheres (subdivision 3: 642 vertices, 1280 triangles, 1920
edges each; radius U(0.05, 0.15)) rejection-packed in a 4 x 4 x 8 box above a
ground plane, uniform random rotations, v0 ~ N(0, 0.5^2), mu = 0.3, seed 2.
Parameters: dt = 0.01, d_hat = 1e-3 (squared barrier threshold), barrier
kappa = 1e-4, orthogonality kappa = 1e9, Newton tol 1e-2, Armijo c1 = 1e-4,
CCD backoff 0.8 (the reference's config.json / settings.h values).

A step is one full adb_step (Newton loop with CCD and line search) on
synthetic input already resident in HBM; by default 1 warm-up step and 100
timed steps, during which the bodies fall and pile up (active contacts grow
from 0 to ~7M, so later steps cost far more than early ones).  Prints one
JSON line (progress lines go to stderr).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBS = 8000.0
MI355X_FP64_TFLOPS = 78.6


def workload(n_bodies=2000):
    from adb_amd import scenes
    sc = scenes.sphere_pack(n_bodies, seed=2, subdiv=3)
    return sc


def params(precision="fp64"):
    from adb_amd.capi import default_params
    return default_params(d_hat=1e-3, kappa=1e-4, kappa_ortho=1e9, mu=0.3, dt=0.01, tol=1e-2,
                          broad_phase_fp32=1 if precision == "mixed" else 0)


def cpu_baseline(sc, p, budget_s, gpu_step_s):
    """The oracle (FP64 C++/OpenMP restatement of the reference path) timed on
    the host: as many whole steps from the initial state as fit in budget_s.
    gpu_step_s are the GPU's per-step times from the same initial state (warm-up
    step first); both sides are compared on steps 1..K-1 of that sample (step 0
    holds the GPU's first-call set-up)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    # the host share of one GPU on the benchmark machines is 16 cores; more
    # threads than that would time other jobs' cores
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = O.lib().oracle_set_threads(min(nproc, 16))
    w = O.OracleWorld(sc, p)
    t0 = time.perf_counter()
    per = []
    pairs = 0
    while True:
        t = time.perf_counter()
        rc, st = w.step()
        per.append(time.perf_counter() - t)
        pairs += st.contacts_processed
        el = time.perf_counter() - t0
        if el >= budget_s or len(per) >= 50:
            break
    k = len(per)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), "")
    except OSError:
        pass
    out = {"value": k / el, "unit": "steps/s", "cores": threads, "kind": "port", "cpu_model": model,
           "nproc": nproc, "threads_note": "OpenMP threads = min(nproc, 16): the host share of one GPU",
           "sample": "first %d steps of the same 2000-icosphere scene from its initial state, %.1f s, "
                     "oracle FP64 restatement with OpenMP (%d threads); contact-pairs/s %.3g"
                     % (k, el, threads, pairs / el)}
    m = min(k, len(gpu_step_s))
    if m > 1:
        cpu_s, gpu_s = sum(per[1:m]), sum(gpu_step_s[1:m])
        out["same_sample"] = {"steps": "1-%d" % (m - 1), "cpu_steps_per_s": round((m - 1) / cpu_s, 4),
                              "gpu_steps_per_s": round((m - 1) / gpu_s, 3), "gpu_over_cpu": round(cpu_s / gpu_s, 1)}
    return out


def run(sc, p, warmup, steps, label="", tune=None):
    """warmup untimed steps, then `steps` timed steps; returns (seconds, totals,
    last stats, world, per-step seconds of the warm-up and timed steps).
    adb_step returns after its stream has drained, so per-step wall times are
    exact; a progress line goes to stderr every 10 steps.  tune: "auto" runs
    the autotuner on the warmed-up state (outside the timed region; the
    launch shapes it picks never change results), a dict applies a tuning."""
    from adb_amd.world import World
    w = World(sc, p)   # World.step raises on any status but ADB_OK, so no failed step is timed
    per_step = []
    for _ in range(warmup):
        t = time.perf_counter()
        w.step()
        per_step.append(time.perf_counter() - t)
    w.tuning_source = "default"
    if tune == "auto":
        from adb_amd import autotune
        _, w.tuning_source = autotune.autotune(w, key=autotune.workload_key(sc, p, "gfx950"), probe_steps=2,
                                               log=lambda m: print(label, m, file=sys.stderr, flush=True))
    elif isinstance(tune, dict):
        w.set_tuning(**tune)
        w.tuning_source = "given"
    w.state()   # synchronizes the world's stream
    tot = {"contacts": 0, "newton": 0, "ls": 0, "pcg": 0, "t_ipc": 0.0, "t_solver": 0.0, "t_ccd": 0.0, "t_ls": 0.0,
           "pcg_not_converged": 0, "pcg_breakdown": 0}
    last = None
    t0 = time.perf_counter()
    for i in range(steps):
        t = time.perf_counter()
        rc, st = w.step()
        per_step.append(time.perf_counter() - t)
        last = st
        tot["contacts"] += st.contacts_processed
        tot["newton"] += st.newton_iters
        tot["ls"] += st.ls_trials
        tot["pcg"] += st.pcg_iters
        tot["t_ipc"] += st.t_ipc
        tot["t_solver"] += st.t_solver
        tot["t_ccd"] += st.t_ccd
        tot["t_ls"] += st.t_line_search
        tot["pcg_not_converged"] += st.pcg_not_converged
        tot["pcg_breakdown"] += st.pcg_breakdown
        if (i + 1) % 10 == 0:
            print("%s step %d/%d: %.1f s, %d contacts" % (label, i + 1, steps, time.perf_counter() - t0, st.n_pt + st.n_ee + st.n_vg),
                  file=sys.stderr, flush=True)
    w.state()
    return time.perf_counter() - t0, tot, last, w, per_step


def max_vertex_deviation(sc, qa, qb):
    """max over vertices of |x_a - x_b| for two states (x = A xbar + p)."""
    import numpy as np
    vo = sc["vert_offset"]
    body = np.repeat(np.arange(len(vo) - 1), np.diff(vo))
    xb = sc["rest_verts"]
    d = (qa - qb).reshape(-1, 12)
    dx = d[body, :3] + np.einsum("nki,nk->ni", d[body, 3:].reshape(-1, 3, 3), xb)
    return float(np.sqrt((dx * dx).sum(1)).max())


# kernels launched once per launch of each profiling scope
KERNEL_OF_SCOPE = {"asm_reduce": ["k_pair_reduce"], "narrow_set": ["k_pair_narrow<0", "k_pair_narrow<2", "k_trial_set"],
                   "narrow_ccd": ["k_pair_narrow<1"],
                   "block_chol": ["k_group_inv"], "contact_group": ["k_contact_c3r", "k_contact_psd9"],
                   "contact": ["k_contact_small<1>", "k_contact_scalars2"], "contact_pk": ["k_contact_pk2"],
                   "spmv": ["k_pcg_spmv"], "pcg_update": ["k_pcg_update"], "contact_classify": ["k_contact_classify"]}
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hbm_traffic_2000_icospheres.json")
SIDE_SCOPES = {"contact_group", "contact_pk"}      # timing scopes on the world's side streams
# FP64 operations per launch measured with rocprofv3 counters (committed
# summary), for the FP64-VALU-bound scopes whose count the host does not know
FP64_OPS_JSON = os.path.join(ROOT, "profiles", "fp64_ops_2000_icospheres.json")
FP64_OPS_OF_SCOPE = {"contact_group", "contact_pk", "narrow_set", "narrow_ccd"}


def measured_fp64_ops(scope, launches):
    """FP64 operations of `launches` launches of the scope from the committed
    counter summary, when it was measured on this workload and launch count."""
    if not os.path.exists(FP64_OPS_JSON):
        return 0.0
    with open(FP64_OPS_JSON) as fh:
        d = json.load(fh)
    e = d.get("scopes", {}).get(scope)
    if not e or e.get("launches") != launches:
        return 0.0
    return e["fp64_ops_per_launch"] * launches


def pmc_traffic(fetch_csv, write_csv, prof_json, scope, copy_mb):
    """HBM bytes per launch of one profiling scope from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch, counters collected for the
    scope's kernels and the calibration kernel only) of
    `bench.py --no-compare --no-cpu-baseline --copy-mb M --profile-json P`.
    Each counter is scaled by what it reads for the plain device copy of M MB
    that the run starts with (the largest elementwise dispatch: M MB read, M
    MB written); for each kernel of the scope the dispatches used are its last
    N, N = the scope's launches in the profiled steps (from P, whose byte
    counts give the algorithmic bytes)."""
    import csv
    with open(prof_json) as fh:
        prof = json.load(fh)
    n = prof[scope]["launches"]
    kernels = KERNEL_OF_SCOPE[scope]
    out = {"kernel_scope": scope, "kernels": kernels, "launches": n, "copy_bytes": copy_mb * 1e6,
           "algorithmic_bytes_per_launch": prof[scope]["bytes"] / n}
    for tag, path in (("fetch", fetch_csv), ("write", write_csv)):
        rows = list(csv.DictReader(open(path)))
        cal = [float(r["Counter_Value"]) for r in rows if "elementwise" in r["Kernel_Name"] or "copyBuffer" in r["Kernel_Name"]]
        scale = copy_mb * 1e6 / (1024.0 * max(cal))
        kib = 0.0
        for k in kernels:
            mine = [float(r["Counter_Value"]) for r in rows if k in r["Kernel_Name"]][-n:]
            kib += sum(mine) / len(mine)
        out[tag + "_kib_raw_avg"] = kib
        out[tag + "_scale"] = scale
        out[tag + "_bytes"] = 1024.0 * scale * kib
    out["traffic_bytes_per_launch"] = out["fetch_bytes"] + out["write_bytes"]
    return out


def pmc_fp64(counter_csv, prof_json, scopes):
    """FP64 operations per launch of FP64-VALU-bound scopes from one rocprofv3
    --pmc run with SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 (wave instructions
    per dispatch) of `bench.py --no-compare --no-cpu-baseline --profile-json
    P`: ops = 64 (ADD + MUL + TRANS) + 128 FMA per dispatch (every lane of
    every issued instruction, so lanes masked off count too: an upper bound on
    the useful operations), averaged over the last N dispatches of each
    kernel, N = the scope's launches in the profiled steps of P."""
    import csv
    with open(prof_json) as fh:
        prof = json.load(fh)
    rows = list(csv.DictReader(open(counter_csv)))
    out = {"source": "rocprofv3 --pmc SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 SQ_INSTS_VALU_FMA_F64 "
                      "SQ_INSTS_VALU_TRANS_F64", "scopes": {},
           "counting": "64 operations per issued ADD / MUL / TRANS wave instruction, 128 per FMA, averaged over the "
                       "scope's launches in the profiled step after the timed ones: every lane of every issued "
                       "instruction, an upper bound on the useful operations"}
    weight = {"SQ_INSTS_VALU_ADD_F64": 64.0, "SQ_INSTS_VALU_MUL_F64": 64.0, "SQ_INSTS_VALU_TRANS_F64": 64.0,
              "SQ_INSTS_VALU_FMA_F64": 128.0}
    for scope in scopes:
        n = prof[scope]["launches"]
        if n <= 0:
            continue
        total = 0.0
        for k in KERNEL_OF_SCOPE[scope]:
            per = {}
            for r in rows:
                if k in r["Kernel_Name"] and r["Counter_Name"] in weight:
                    per.setdefault(r["Dispatch_Id"], 0.0)
                    per[r["Dispatch_Id"]] += weight[r["Counter_Name"]] * float(r["Counter_Value"])
            vals = [per[d] for d in sorted(per, key=int)][-n:]
            total += sum(vals) / max(1, len(vals))
        out["scopes"][scope] = {"kernels": KERNEL_OF_SCOPE[scope], "launches": n, "fp64_ops_per_launch": total}
    return out


def kernel_stats_md(stats_csv, title, cmd, top=25):
    """Markdown table of a rocprofv3 --stats kernel summary (CSV)."""
    import csv
    rows = list(csv.DictReader(open(stats_csv)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    lines = ["# " + title, "", "Command (from the repository root, on one MI355X):", "", "    " + cmd, "",
             "Total kernel time: %.2f ms over all dispatches of the run (warm-up included)." % (tot / 1e6), "",
             "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        name = r["Name"].replace("|", "/")
        if len(name) > 90:
            name = name[:87] + "..."
        lines.append("| `%s` | %s | %.3f | %.1f | %.1f | %.1f | %.2f |" % (
            name, r["Calls"], float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3,
            float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    return "\n".join(lines) + "\n"


def replay_main(args):
    """One step of the workload from a saved state, replayed: every replay
    starts from the same q, qdot, so the timings compare code, not
    trajectories (the pile's Newton counts depend on rounding anywhere)."""
    import numpy as np
    from adb_amd.world import World
    sc = workload(args.bodies)
    p = params(args.precision)
    d = np.load(args.replay)
    w = World(sc, p)
    times, st = [], None
    for i in range(args.warmup + args.steps):
        w.set_state(d["q"], d["qdot"])
        t = time.perf_counter()
        rc, st = w.step()
        if i >= args.warmup:
            times.append(time.perf_counter() - t)
    prof = {}
    if args.profile_steps > 0:
        w.profile_enable(True)
        w.set_state(d["q"], d["qdot"])
        w.step()
        prof = w.profile_read()
        w.profile_enable(False)
        if args.profile_json:
            with open(args.profile_json, "w") as fh:
                json.dump(prof, fh, indent=1)
    ms = 1000.0 * sum(times) / max(1, len(times))
    ranked = sorted(prof.items(), key=lambda kv: -kv[1]["ms"])
    print(json.dumps({"metric": "ms per replayed step", "value": round(ms, 3), "unit": "ms", "higher_is_better": False,
                      "steps": args.steps, "warmup": args.warmup,
                      "config": {"workload": "2000 icospheres (642 v / 1280 t / 1920 e each), one step replayed from %s"
                                             % os.path.basename(args.replay), "precision": args.precision},
                      "newton_iters": st.newton_iters, "pcg_iters": st.pcg_iters, "ls_trials": st.ls_trials,
                      "contacts": st.n_pt + st.n_ee + st.n_vg, "min_ms": round(1000.0 * min(times), 3),
                      "top_scopes_ms": {k: [round(v["ms"], 3), v["launches"]] for k, v in ranked[:10]}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed",
                    help="precision of the reported value; the other one is timed beside it")
    ap.add_argument("--bodies", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the run in the other precision")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--profile-steps", type=int, default=1,
                    help="steps after the timed ones with per-kernel HIP-event timing (roofline figures)")
    ap.add_argument("--profile-json", default="")
    ap.add_argument("--copy-mb", type=float, default=0.0,
                    help="first run one elementwise device kernel that reads and writes this many MB (a "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--pmc-traffic", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "PROFILE_JSON", "SCOPE"),
                    help="no run: summarize two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) into HBM bytes per "
                         "launch of one kernel scope, scaled by the --copy-mb copy; writes --out")
    ap.add_argument("--pmc-fp64", nargs=2, metavar=("COUNTER_CSV", "PROFILE_JSON"),
                    help="no run: summarize a rocprofv3 --pmc run of the FP64 instruction counters into FP64 "
                         "operations per launch of the FP64-VALU-bound scopes; writes --out")
    ap.add_argument("--kernel-stats", nargs=3, metavar=("STATS_CSV", "TITLE", "CMD"),
                    help="no run: write a markdown table of a rocprofv3 --stats kernel summary to --out")
    ap.add_argument("--replay", default="", metavar="NPZ",
                    help="time single steps replayed from a saved state (q, qdot arrays of the same workload, e.g. "
                         "tests/golden/c2_pile_step70.npz): --warmup untimed and --steps timed replays, one profiled")
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    if args.pmc_traffic:
        res = pmc_traffic(*args.pmc_traffic, args.copy_mb)
        if args.out:
            with open(args.out, "w") as fh:
                json.dump(res, fh, indent=1)
        print(json.dumps(res))
        return
    if args.pmc_fp64:
        res = pmc_fp64(args.pmc_fp64[0], args.pmc_fp64[1], sorted(FP64_OPS_OF_SCOPE | {"block_chol"}))
        if args.out:
            with open(args.out, "w") as fh:
                json.dump(res, fh, indent=1)
        print(json.dumps(res))
        return
    if args.kernel_stats:
        md = kernel_stats_md(*args.kernel_stats)
        if args.out:
            with open(args.out, "w") as fh:
                fh.write(md)
        print(md)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the contact step of one scene runs on one device "
              "(multi-GPU sharding is a design note, not implemented); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_mb > 0:
        import torch
        n = int(args.copy_mb * 1e6 / 8)
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.add(src, 1.0, out=dst)      # reads and writes copy_mb MB
        torch.cuda.synchronize()
        del src, dst
    if args.replay:
        replay_main(args)
        return
    sc = workload(args.bodies)
    p = params(args.precision)
    elapsed, tot, last, w, per_step = run(sc, p, args.warmup, args.steps, args.precision,
                                          tune=None if args.no_autotune else "auto")
    q_main, _ = w.state()
    tuning = dict(w.tuning(), source=w.tuning_source)

    # per-kernel timing on extra steps (HIP events on the world's stream)
    prof = {}
    if args.profile_steps > 0:
        w.profile_enable(True)
        for _ in range(args.profile_steps):
            w.step()
        prof = w.profile_read()
        w.profile_enable(False)
    del w
    if args.profile_json:
        with open(args.profile_json, "w") as fh:
            json.dump(prof, fh, indent=1)

    other = None
    if not args.no_compare:
        op = "fp64" if args.precision == "mixed" else "mixed"
        e2, t2, _, w2, _ = run(sc, params(op), args.warmup, args.steps, op,
                               tune={k: v for k, v in tuning.items() if k != "source"})
        q2, _ = w2.state()
        del w2
        other = {"precision": op, "steps_per_s": round(args.steps / e2, 4),
                 "max_vertex_deviation_vs_value_run": max_vertex_deviation(sc, q_main, q2),
                 "newton_iters": t2["newton"], "pcg_iters": t2["pcg"]}

    steps_per_s = args.steps / elapsed
    # dominant kernel scope: the most device time in the profiled steps (side
    # stream scopes included; their event times include sharing the GPU with
    # the main stream's kernels).  FP64-VALU-bound scopes with a counted
    # operation total report FP64 throughput against 78.6 TFLOP/s (MI355X's
    # FP64 peak, the same for vector and matrix); the others report their
    # algorithmic HBM bytes against 8 TB/s.
    ranked = sorted(prof.items(), key=lambda kv: -kv[1]["ms"])
    dom = next(((k, v) for k, v in ranked if v["launches"] > 0), None)
    roofline = None
    if dom:
        k, v = dom
        avg_ms = v["ms"] / v["launches"]
        side = {q: round(prof[q]["ms"], 3) for q in sorted(SIDE_SCOPES) if q in prof}
        flops = v.get("flops", 0.0)
        if flops <= 0.0 and k in FP64_OPS_OF_SCOPE:
            flops = measured_fp64_ops(k, v["launches"])
        if flops > 0.0:
            per = flops / v["launches"]
            tf = per / (avg_ms * 1e-3) / 1e12
            roofline = {"kernel": k, "bound": "mfma", "pipe": "fp64 VALU", "achieved": round(tf, 3),
                        "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / MI355X_FP64_TFLOPS, 4),
                        "traffic": None, "avg_launch_ms": avg_ms, "flops_per_launch": per, "launches": v["launches"],
                        "flops_source": "counted" if v.get("flops", 0.0) > 0.0 else "rocprofv3 FP64 instruction counters",
                        "on_side_stream": k in SIDE_SCOPES, "concurrent_side_stream_ms": side}
        else:
            bytes_per_launch = v["bytes"] / v["launches"]
            gbs = bytes_per_launch / (avg_ms * 1e-3) / 1e9
            roofline = {"kernel": k, "bound": "hbm", "achieved": round(gbs, 1), "peak": MI355X_HBM_GBS,
                        "unit": "GB/s", "frac": round(gbs / MI355X_HBM_GBS, 4), "traffic": None,
                        "avg_launch_ms": avg_ms, "bytes_per_launch": bytes_per_launch, "launches": v["launches"],
                        "on_side_stream": k in SIDE_SCOPES, "concurrent_side_stream_ms": side}
        # HBM bytes per launch measured by rocprofv3 --pmc on this workload (committed summary)
        if os.path.exists(TRAFFIC_JSON):
            with open(TRAFFIC_JSON) as fh:
                tr = json.load(fh)
            # only for the workload it was measured on (same scope, kernels and launches)
            if (tr.get("kernel_scope") == k and tr.get("kernels") == KERNEL_OF_SCOPE.get(k)
                    and tr.get("launches") == v["launches"]):
                roofline["traffic"] = round(tr["traffic_bytes_per_launch"], 0)
                roofline["traffic_source"] = os.path.relpath(TRAFFIC_JSON, ROOT)

    line = {
        "metric": "steps/s at 2k bodies (mixed vs FP64, same traj err); contact-pairs/s",
        "value": round(steps_per_s, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (f32 outward-rounded broad-phase boxes)",
        "data": "synthetic (seeded PCG32 scene, no dataset)",
        "config": {"workload": "2000 icospheres (642 v / 1280 t / 1920 e each), 4x4x8 box, ground, mu=0.3, "
                               "dt=0.01, d_hat=1e-3, seed 2",
                   "bodies": args.bodies, "precision": args.precision},
        "contact_pairs_per_s": round(tot["contacts"] / elapsed, 1),
        "newton_iters": tot["newton"], "ls_trials": tot["ls"], "pcg_iters": tot["pcg"],
        "pcg_unfinished": {"max_iter": tot["pcg_not_converged"], "breakdown": tot["pcg_breakdown"]},
        "phase_s": {"ipc": round(tot["t_ipc"], 4), "solver": round(tot["t_solver"], 4),
                    "ccd": round(tot["t_ccd"], 4), "line_search": round(tot["t_ls"], 4)},
        "contacts_last_step": {"pt": last.n_pt, "ee": last.n_ee, "vg": last.n_vg} if last else None,
        "step_ms": {"first_10": round(1000.0 * sum(per_step[args.warmup:args.warmup + 10]) / max(1, min(10, args.steps)), 2),
                    "median": round(1000.0 * sorted(per_step[args.warmup:])[args.steps // 2], 2) if args.steps else None,
                    "max": round(1000.0 * max(per_step[args.warmup:]), 2) if args.steps else None},
        "other_precision": other,
        "tuning": tuning,
        "top_kernels_ms_profiled_steps": {k: round(v["ms"], 3) for k, v in ranked[:8]},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(sc, p, args.cpu_seconds, per_step)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
