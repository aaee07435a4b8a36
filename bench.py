"""Benchmark of the MI355X ExoPlaSim step: simulated days per second.

Default workload (N=1): the 64-planet mixed ensemble (32 x T21 L5,
24 x T42 L10, 8 x T63 L10; 45-minute steps; seeded rotation, insolation,
pressure and land fraction) advanced together, one batched step per
"step", state resident in HBM. `value` is the aggregate simulated days/s
of the batch. Beside it, the plain FP32 single-run form (one T63 L10
planet, 30-minute step) is timed on the same GPU for the batched-vs-plain
ratio, the same planets are stepped one at a time (each alone) for the
batched-vs-one-at-a-time ratio, the dominant kernel's launches are timed
with HIP events for the roofline figure, and the FP64 CPU oracle is timed
on the host's cores (one planet per thread) on a bounded sample of the
same workload. Prints one JSON line.

The serving loop (continuous batching of Poisson-arriving runs, 4 runs/s,
T21/T42/T63 weights 0.5/0.35/0.15, L5 or L10, 1-360 simulated days each,
ensemble parameters) runs after it over a --serve-seconds arrival window
and reports p50/p99 run-completion latency and aggregate days/s in the
line's "serving" object; --workload serving makes it the measured value.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload batched|plain|serving] [--probe KERNEL]
                    [--serve-seconds S] [--no-serving]
                    [--no-cpu-baseline] [--cpu-cores C] [--no-plain]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simulated days/s at T63 L10, batched vs plain (1 GPU); p50/p99 run latency"


def _package():
    import __graft_entry__ as entry
    return entry._package()


def probe_steps(steps):
    """Steps of the probe window that follows the timed region."""
    return min(steps, 96)


def run_gpu(exo, cfgs, steps, warmup, probes=(), precision=0, graphs=True, tune=False, lanes=0):
    """Steps every config together; returns (wall s, device ms, probed).
    The timed steps replay the captured step graph (graphs=True) or launch
    every kernel (False); each kernel of `probes` then has its single
    launches timed over a further window of steps of its own, after the
    timed region (probed: kernel -> (total ms, launches, missed))."""
    eng = exo.Engine(0, len(cfgs))
    try:
        eng.set_step_graphs(graphs)
        if lanes:
            eng.set_step_lanes(lanes)
        ids = [eng.add_member(c, precision=precision) for c in cfgs]
        eng.step(warmup)
        eng.sync()
        tuned = None
        if tune:
            # transform group sizes for this batch, picked by timing; the
            # state is left as it was
            t_tune = time.perf_counter()
            ms = eng.autotune(trial_steps=8)
            if lanes:
                eng.set_step_lanes(lanes)  # a lane count given on the command line stays
            classes = sorted({(c.truncation, c.levels) for c in cfgs})
            tuned = {"groups": {f"T{t} L{l}": eng.transform_group(t, l) for t, l in classes},
                     "step_graphs": eng.step_graphs(), "lanes": eng.step_lanes(),
                     "best_trial_ms_per_step": ms, "seconds": time.perf_counter() - t_tune}
        t0 = time.perf_counter()
        eng.step(steps)
        eng.sync()
        wall = time.perf_counter() - t0
        dev_ms = eng.last_step_ms()
        probed = {}
        for kernel in probes:
            eng.probe_kernel(kernel, 65536)
            eng.step(probe_steps(steps))
            eng.sync()
            probed[kernel] = eng.probe_result()
            eng.probe_kernel(None)
        for i in ids:
            ts, _ = eng.global_means(i)
            if not (100.0 < ts < 500.0):
                raise RuntimeError(f"member {i} left the physical range: mean Ts {ts}")
        run_gpu.tuned = tuned
        return wall, dev_ms, probed
    finally:
        eng.close()


def cpu_cores():
    """Host cores the CPU baseline may use: OMP_NUM_THREADS where set (the
    box's CPU share), else the cores this process may run on, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(exo, cfgs, seconds_per_case, cores):
    """FP64 oracle steps on the host cores: for each (truncation, levels)
    group of the workload, `cores` planets of the group step at once, one
    per thread (the oracle's C step runs outside the GIL), after their
    start-up steps; the per-step time under that full load, times the
    group's planets over the cores, gives the workload's aggregate
    simulated days/s on `cores` cores."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_harness import OracleRun
    groups = {}
    for c in cfgs:
        groups.setdefault((c.truncation, c.levels, c.timestep_minutes), []).append(c)

    def timed_run(cfg):
        host = exo.Engine(-1, 1)
        try:
            run = OracleRun(host, host.add_member(cfg))
            run.step(run.nstartup, physics=True)  # the start-up phases, untimed
            n = 0
            t0 = time.perf_counter()
            while True:
                run.step(1, physics=True)
                n += 1
                el = time.perf_counter() - t0
                if el >= seconds_per_case or n >= 64:
                    return el / n, n
        finally:
            host.close()

    cpu_s_per_day = 0.0
    days = 0.0
    sample = []
    for (trunc, levels, minutes), members in groups.items():
        picks = [members[i % len(members)] for i in range(cores)]
        with ThreadPoolExecutor(cores) as pool:
            res = list(pool.map(timed_run, picks))
        per_step = sum(r[0] for r in res) / len(res)
        spd = 1440.0 / minutes
        cpu_s_per_day += per_step * spd * len(members) / cores
        days += len(members)
        sample.append(f"T{trunc} L{levels}: {cores} planets at once, {min(r[1] for r in res)}+ main-loop steps "
                      f"each, {per_step * 1e3:.1f} ms/step per planet")
    return {"value": days / cpu_s_per_day, "unit": "simulated days/s", "cores": cores, "kind": "port",
            "sample": f"FP64 C oracle (oracle/liboracle.so) on {cores} host threads, each group of the workload "
                      "sampled with one planet per thread after the start-up steps, scaled by group size: " +
                      "; ".join(sample)}


def run_one_at_a_time(exo, cfgs, precision, steps=32, warmup=4):
    """The same planets stepped one at a time on the GPU (each alone in the
    engine, its step graph replayed): simulated days over the summed time."""
    from web_exoplasim_amd import workloads
    total_s, days = 0.0, 0.0
    with exo.Engine(0, 1) as eng:
        for c in cfgs:
            m = eng.add_member(c, precision=precision)
            eng.step(warmup)
            eng.sync()
            t0 = time.perf_counter()
            eng.step(steps)
            eng.sync()
            total_s += time.perf_counter() - t0
            days += steps / workloads.steps_per_day(c.timestep_minutes)
            eng.remove_member(m)
    return {"value": days / total_s, "unit": "simulated days/s", "steps_per_planet": steps,
            "config": "the same planets, each stepped alone (after its start-up steps), times summed"}


def run_serving(exo, seconds, rate=4.0, seed=5):
    """The serving loop over a `seconds` arrival window."""
    import numpy as np
    from web_exoplasim_amd import workloads
    reqs = workloads.serving_requests(exo, seed=seed, rate=rate, seconds=seconds)
    spd = workloads.steps_per_day(45)
    runs = [(t, cfg, max(1, int(round(days * spd)))) for t, cfg, days in reqs]
    mark = [0]

    def progress(done, total, elapsed):
        # one line per hundred runs ended
        if done // 100 > mark[0] or done == total:
            mark[0] = done // 100
            print(f"serving: {done}/{total} runs ended at {elapsed:.1f} s", file=sys.stderr, flush=True)

    with exo.Engine(0, 64) as eng:
        results, wall = eng.serve(runs, max_chunk_steps=32, prepare_threads=8, progress=progress)
    ok = [r for r in results if r["status"] == 0]
    if len(ok) != len(results):
        raise RuntimeError(f"{len(results) - len(ok)} served runs failed")
    lat = np.array([r["finished_s"] - r["arrival_s"] for r in ok])
    queued = np.array([r["admitted_s"] - r["arrival_s"] for r in ok])
    days = sum(r["steps"] for r in ok) / spd
    return {"value": days / wall, "unit": "simulated days/s", "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "p50_admit_wait_s": float(np.percentile(queued, 50)),
            "runs": len(ok), "arrival_window_s": seconds, "arrival_rate_per_s": rate, "wall_s": wall,
            "simulated_days": days,
            "config": "Poisson arrivals, T21/T42/T63 weights 0.5/0.35/0.15, L5 or L10, length U[1,360] days, "
                      "45-min step, ensemble parameters (seed 5); up to 64 resident runs"}


def copy_calibration():
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=320)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--workload", default="batched", choices=["batched", "plain", "serving"])
    ap.add_argument("--form", default="plain", choices=["split", "plain"],
                    help="operand form of the batched ensemble's Legendre products (the other form is "
                         "timed beside it)")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--probe", default="auto")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default transform group size (4 members) and lanes instead of tuning them")
    ap.add_argument("--lanes", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="streams a replayed step's members run on (0: the engine's default, or the "
                         "autotuner's pick)")
    ap.add_argument("--cpu-seconds", type=float, default=4.0)
    ap.add_argument("--cpu-cores", type=int, default=0, help="host threads of the CPU baseline (default: the "
                    "box's CPU share, OMP_NUM_THREADS)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three device copies of 256 MiB (the known-size kernel that "
                         "scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every member of the workload runs on the one GPU "
              "of this box, and no multi-GPU figure is reported", file=sys.stderr)
        sys.exit(2)
    exo = _package()
    from web_exoplasim_amd import workloads, roofline

    if args.workload == "serving":
        sv = run_serving(exo, args.serve_seconds)
        print(json.dumps({
            "metric": METRIC, "value": sv["value"], "unit": "simulated days/s", "n_gpus": 1,
            "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic seeded planets",
            "config": {"workload": "serving loop: " + sv["config"], "arrival_window_s": args.serve_seconds},
            "serving": sv, "roofline": None, "cpu_baseline": None}))
        return

    precision = 0
    if args.workload == "batched":
        cfgs = workloads.ensemble(exo)
        precision = exo.EXO_PRECISION_SPLIT if args.form == "split" else exo.EXO_PRECISION_FP32
        workload = {"workload": "64-planet mixed ensemble: 32 x T21 L5, 24 x T42 L10, 8 x T63 L10, "
                                "45-min step, seeded parameters (seed 64)",
                    "members": len(cfgs), "timestep_minutes": 45,
                    "form": ("split-operand Legendre (FP16 hi/lo, FP32 accumulate)" if args.form == "split"
                             else "plain FP32")}
    else:
        cfgs = [workloads.plain(exo, 63, 10, 30)]
        workload = {"workload": "plain FP32 single planet T63 L10, 30-min step",
                    "truncation": 63, "levels": 10, "timestep_minutes": 30, "members": 1}
    probe = roofline.DOMINANT[args.workload] if args.probe == "auto" else (args.probe or None)
    secondary = list(roofline.SECONDARY[args.workload]) if args.probe == "auto" else []

    def form_name(k):
        split = precision == exo.EXO_PRECISION_SPLIT and k in ("k_leg_inv", "k_leg_fwd")
        return k + "_split" if split else k

    probe = form_name(probe) if probe else None
    secondary = [form_name(k) for k in secondary]
    probes = ([probe] if probe else []) + secondary
    wall, dev_ms, probed = run_gpu(exo, cfgs, args.steps, args.warmup, probes, precision,
                                   tune=not args.no_autotune, lanes=args.lanes)
    days = sum(args.steps / workloads.steps_per_day(c.timestep_minutes) for c in cfgs)
    value = days / wall
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "simulated days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1000.0 / args.steps,
        "device_ms_per_step": dev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded planets (SH terrain, isothermal 250 K start)",
        "config": workload,
        "autotune": run_gpu.tuned,
    }
    if args.workload == "batched" and not args.no_plain:
        other = exo.EXO_PRECISION_FP32 if precision == exo.EXO_PRECISION_SPLIT else exo.EXO_PRECISION_SPLIT
        ow, _, _ = run_gpu(exo, cfgs, args.steps, args.warmup, (), other, tune=not args.no_autotune)
        name = "batched_plain_fp32" if other == exo.EXO_PRECISION_FP32 else "batched_split_fp16"
        line[name] = {"value": days / ow, "unit": "simulated days/s", "ms_per_step": ow * 1000.0 / args.steps,
                      "form": ("plain FP32" if other == exo.EXO_PRECISION_FP32
                               else "split-operand Legendre (FP16 hi/lo, FP32 accumulate)")}
    if args.workload == "batched" and not args.no_plain:
        plain_cfg = workloads.plain(exo, 63, 10, 30)
        pw, _, _ = run_gpu(exo, [plain_cfg], 480, 48, (), tune=not args.no_autotune)
        plain_tuned = run_gpu.tuned
        plain_value = (480 / 48) / pw
        line["plain_single"] = {"value": plain_value, "unit": "simulated days/s",
                                "ms_per_step": pw * 1000.0 / 480,
                                "config": "plain FP32 single planet T63 L10, 30-min step",
                                "autotune": plain_tuned}
        ew, _, _ = run_gpu(exo, [plain_cfg], 480, 48, (), graphs=False)
        line["plain_single_eager"] = {"value": (480 / 48) / ew, "unit": "simulated days/s",
                                      "ms_per_step": 1e3 * ew / 480,
                                      "config": "as plain_single, every kernel launched (no step graph)"}
        line["batched_vs_plain"] = value / plain_value
        if "batched_split_fp16" in line:
            line["batched_split_vs_plain"] = line["batched_split_fp16"]["value"] / plain_value
    if args.workload == "batched" and not args.no_plain:
        one = run_one_at_a_time(exo, cfgs, precision)
        line["one_at_a_time"] = one
        line["batched_vs_one_at_a_time"] = value / one["value"]
    if args.workload == "batched" and not args.no_serving:
        line["serving"] = run_serving(exo, args.serve_seconds)
    if probe:
        total_ms, launches, missed = probed[probe]
        line["roofline"] = roofline.figure(exo, probe, cfgs, total_ms, launches, missed,
                                           probe_steps(args.steps))
    else:
        line["roofline"] = None
    if secondary:
        line["roofline_secondary"] = [roofline.figure(exo, k, cfgs, *probed[k], probe_steps(args.steps))
                                      for k in secondary]
    line["cpu_baseline"] = (None if args.no_cpu_baseline else
                            cpu_baseline(exo, cfgs, args.cpu_seconds, args.cpu_cores or cpu_cores()))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
