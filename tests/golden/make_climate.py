"""Writes the climate fixtures: global-mean surface temperature and TOA
net flux of the FP64 oracle (and of its FP32 emulation) at every step of
seeded runs, from the engine's own FP64 set-up of the same configuration.

    python tests/golden/make_climate.py CASE [--fp32]

CASE is one of CASES or ENSEMBLE_MEMBERS below. Each run writes tests/golden/climate_<CASE>
[_fp32].json: {"case", "config", "steps", "ts": [...], "toa": [...]}, one
entry per step. The means are area-weighted with the Gaussian weights, as
the engine's exo_global_means computes them (surface level of dt; level 0
of dswfl + dlwfl).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

# name: (config overrides, steps)
CASES = {
    # plain method, 1 day at T21 L10 / 45 min
    "t21_day1": (dict(truncation=21, levels=10, timestep_minutes=45), 32),
    # plain FP32 drop-in: T42 L10, 45 min, 30 days, 30 % land
    "t42_day30": (dict(truncation=42, levels=10, timestep_minutes=45, land_fraction=0.3), 960),
    # precision form: T63 L10, 30 min, 30 days
    "t63_day30": (dict(truncation=63, levels=10, timestep_minutes=30), 1440),
}
# The batched ensemble (64 planets drawn from seed 64: 32 x T21 L5,
# 24 x T42 L10, 8 x T63 L10, 45-minute steps): the first member of each
# class, 3 start-up steps and 349 main-loop steps (just under 11 days).
ENSEMBLE_MEMBERS = {"c4_member0": 0, "c4_member32": 32, "c4_member56": 56}
ENSEMBLE_STEPS = 352


def ensemble_case(exo, case):
    """(config overrides, steps) of an ensemble member case."""
    import importlib
    workloads = importlib.import_module(exo.__name__ + ".workloads")
    cfg = workloads.ensemble(exo, seed=64)[ENSEMBLE_MEMBERS[case]]
    return {name: getattr(cfg, name) for name, _ in cfg._fields_}, ENSEMBLE_STEPS


def global_means(run, weights, nlev, nhor):
    dt = run.field("dt").astype(np.float64)
    sw = run.field("dswfl").astype(np.float64)
    lw = run.field("dlwfl").astype(np.float64)
    ts = dt[nlev * nhor:(nlev + 1) * nhor]
    toa = sw[:nhor] + lw[:nhor]
    return float(np.dot(weights, ts) / weights.sum()), float(np.dot(weights, toa) / weights.sum())


def main():
    import __graft_entry__ as entry
    from oracle_harness import OracleRun
    ap = argparse.ArgumentParser()
    ap.add_argument("case", choices=sorted(CASES) + sorted(ENSEMBLE_MEMBERS))
    ap.add_argument("--fp32", action="store_true")
    args = ap.parse_args()
    exo = entry._package()
    overrides, steps = CASES[args.case] if args.case in CASES else ensemble_case(exo, args.case)
    cfg = exo.default_config(**overrides)
    host = exo.Engine(-1, 1)
    m = host.add_member(cfg)
    _, nlat, nlon, nlev = host.dims(m)
    weights = host.read_table(m, "zgw")
    run = OracleRun(host, m, precision="fp32" if args.fp32 else "fp64")
    ts, toa = [], []
    t0 = time.time()
    for s in range(steps):
        run.step(1, physics=True)
        a, b = global_means(run, weights, nlev, nlat * nlon)
        ts.append(a)
        toa.append(b)
        if (s + 1) % 48 == 0:
            print(f"{args.case} step {s + 1}/{steps} Ts {a:.3f} TOA {b:.3f} ({time.time() - t0:.0f} s)", flush=True)
    out = {"case": args.case, "precision": "fp32" if args.fp32 else "fp64", "config": overrides,
           "steps": steps, "ts": ts, "toa": toa}
    name = f"climate_{args.case}{'_fp32' if args.fp32 else ''}.json"
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(out, f)
    print("wrote", name)


if __name__ == "__main__":
    main()
