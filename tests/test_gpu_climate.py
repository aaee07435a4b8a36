"""Climate-scale gate: after the run of each fixture (1 day at T21 L10,
30 days at T42 L10 and T63 L10), the GPU's area-weighted global-mean
surface temperature must be within 0.5 K, and its TOA net flux within
1 W/m^2, of the FP64 oracle's on the same seeded inputs. The tolerances
are stated because threshold physics makes single-precision trajectories
diverge: the FP32 emulation of the oracle itself differs from the FP64
oracle by up to 0.03 K and 1.6 W/m^2 at single steps of the T42 run.

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


@pytest.mark.parametrize("case", CASES)
def test_global_means_match_fp64_oracle(exo, case):
    fx = _fixture(case)
    cfg = exo.default_config(**fx["config"])
    steps = fx["steps"]
    ts, toa = [], []
    with exo.Engine(0, 1) as eng:
        m = eng.add_member(cfg)
        day = max(1, steps // 30)
        for at in range(0, steps, day):
            n = min(day, steps - at)
            eng.step(n)
            eng.sync()
            a, b = eng.global_means(m)
            ts.append(a)
            toa.append(b)
    ref_ts, ref_toa = fx["ts"][-1], fx["toa"][-1]
    print(f"{case}: GPU Ts {ts[-1]:.3f} K vs oracle {ref_ts:.3f} K; "
          f"GPU TOA {toa[-1]:.3f} vs oracle {ref_toa:.3f} W/m^2")
    assert np.all(np.isfinite(ts)) and np.all(np.isfinite(toa))
    assert abs(ts[-1] - ref_ts) <= 0.5, (ts[-1], ref_ts)
    assert abs(toa[-1] - ref_toa) <= 1.0, (toa[-1], ref_toa)
