"""Climate-scale gate: over the last simulated day of each fixture's run
(day 1 at T21 L10, day 30 at T42 L10 and T63 L10), the mean of the GPU's
area-weighted global-mean surface temperature must be within 0.5 K, and
the mean of its TOA net flux within 1 W/m^2, of the FP64 oracle's on the
same seeded inputs; the surface temperature after the last step must be
within 0.5 K too. The tolerances are stated because threshold physics
makes single-precision trajectories diverge: the FP32 emulation of the
oracle itself differs from the FP64 oracle by up to 0.05 K and 1.6 W/m^2
at single steps (TOA net at one instant follows the clouds of that step),
and by 0.03 W/m^2 in the day-30 daily mean at T42.

Both operand forms of the Legendre products meet the same gates: the plain
FP32 form and the split-operand form (FP16 hi/lo operands, FP32
accumulation), each against the same FP64 fixture.

The fixtures come from tests/golden/make_climate.py (FP64 oracle, the
engine's own FP64 set-up of the same configuration)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["t21_day1", "t42_day30", "t63_day30"]


def _fixture(case):
    path = os.path.join(GOLDEN, f"climate_{case}.json")
    if not os.path.exists(path):
        pytest.skip(f"{path} has not been generated")
    with open(path) as f:
        return json.load(f)


@pytest.mark.parametrize("form", ["plain", "split"])
@pytest.mark.parametrize("case", CASES)
def test_global_means_match_fp64_oracle(exo, case, form):
    fx = _fixture(case)
    cfg = exo.default_config(**fx["config"])
    steps = fx["steps"]
    day = int(round(1440.0 / cfg.timestep_minutes))
    ts, toa = [], []
    with exo.Engine(0, 1) as eng:
        m = eng.add_member(cfg, precision=exo.EXO_PRECISION_SPLIT if form == "split" else exo.EXO_PRECISION_FP32)
        if steps > day:
            eng.step(steps - day)
        for _ in range(min(day, steps)):
            eng.step(1)
            eng.sync()
            a, b = eng.global_means(m)
            ts.append(a)
            toa.append(b)
    n = len(ts)
    ref_ts, ref_toa = np.mean(fx["ts"][-n:]), np.mean(fx["toa"][-n:])
    print(f"{case} {form}: last-day mean GPU Ts {np.mean(ts):.3f} K vs oracle {ref_ts:.3f} K; "
          f"TOA {np.mean(toa):.3f} vs {ref_toa:.3f} W/m^2; after the last step Ts {ts[-1]:.3f} vs "
          f"{fx['ts'][-1]:.3f} K, TOA {toa[-1]:.3f} vs {fx['toa'][-1]:.3f} W/m^2")
    assert np.all(np.isfinite(ts)) and np.all(np.isfinite(toa))
    assert abs(np.mean(ts) - ref_ts) <= 0.5, (np.mean(ts), ref_ts)
    assert abs(np.mean(toa) - ref_toa) <= 1.0, (np.mean(toa), ref_toa)
    assert abs(ts[-1] - fx["ts"][-1]) <= 0.5, (ts[-1], fx["ts"][-1])
