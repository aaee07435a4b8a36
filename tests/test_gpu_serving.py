"""The serving loop (exo_serve): runs that arrive while others are on the
GPU are admitted as descriptors and memory free up, each retires at
exactly its step count, and each result equals the same run alone."""
import pytest

pytestmark = pytest.mark.gpu


def _mix(exo):
    spec = [(0.00, 21, 5, 7, 40), (0.00, 42, 10, 8, 23), (0.05, 21, 10, 9, 31),
            (0.30, 63, 10, 10, 12), (0.35, 21, 5, 11, 57), (0.60, 42, 5, 12, 19)]
    out = []
    for arrival, trunc, levels, seed, steps in spec:
        cfg = exo.default_config(truncation=trunc, levels=levels, timestep_minutes=45,
                                 terrain_seed=seed, noise_seed=seed + 100, land_fraction=0.1 * (seed % 5))
        out.append((arrival, cfg, steps))
    return out


def test_served_runs_equal_runs_alone(exo):
    requests = _mix(exo)
    with exo.Engine(0, 3) as eng:  # 3 descriptors: later runs wait for room
        results, wall = eng.serve(requests, max_chunk_steps=8, prepare_threads=2)
    assert len(results) == len(requests)
    for (arrival, cfg, steps), r in zip(requests, results):
        assert r["status"] == 0, r
        assert r["steps"] == steps
        assert r["arrival_s"] <= r["admitted_s"] <= r["finished_s"] <= wall
        with exo.Engine(0, 1) as alone:
            m = alone.add_member(cfg)
            alone.step(steps)
            alone.sync()
            ts, toa = alone.global_means(m)
        print(f"T{cfg.truncation} L{cfg.levels} {steps} steps: served Ts {r['tsurf']:.6f} alone {ts:.6f}")
        assert r["tsurf"] == ts and r["toa_net"] == toa


def test_budget_too_small_fails_the_run_only(exo):
    requests = _mix(exo)[:2]
    with exo.Engine(0, 4) as eng:
        results, _ = eng.serve(requests, hbm_budget_bytes=16 * 1024 * 1024)  # T21 L5 (6.5 MB) fits, T42 L10 (41 MB) never
    assert results[0]["status"] == 0
    assert results[1]["status"] != 0
