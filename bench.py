"""Benchmark of the MI355X ExoPlaSim step: simulated days per second.

Default workload (N=1): one T63 L10 planet at a 30-minute step in the plain
FP32 form (48 steps = one simulated day), state resident in HBM. A "step"
is one model time step of the whole planet. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain|batched]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480)
    ap.add_argument("--warmup", type=int, default=48)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every member of the workload runs on the "
              "one GPU of this box, and no multi-GPU figure is reported", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__ as entry
    exo = entry._package()

    cfg = exo.default_config(truncation=63, levels=10, timestep_minutes=30)
    steps_per_day = 48
    eng = exo.Engine(0, 1)
    m = eng.add_member(cfg)
    eng.step(args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    dev_ms = eng.last_step_ms()
    ms_per_step = wall * 1000.0 / args.steps
    days = args.steps / steps_per_day
    line = {
        "metric": "simulated days/s at T63 L10, batched vs plain (1 GPU); p50/p99 run latency",
        "value": days / wall,
        "unit": "simulated days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "device_ms_per_step": dev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded planet (SH terrain, isothermal 250 K start)",
        "config": {"workload": "plain FP32 single planet T63 L10, 30-min step",
                   "truncation": 63, "levels": 10, "timestep_minutes": 30, "members": 1},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
