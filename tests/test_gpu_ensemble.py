"""The batched ensemble at full size: 64 planets drawn from seed 64
(32 x T21 L5, 24 x T42 L10, 8 x T63 L10; rotation 0.5-100 days,
insolation 0.6-1.4 x 1367 W/m^2, 0.3-3 bar, land 0-60 %), 45-minute
steps, stepped together for 352 steps (3 start-up steps and just under 11
simulated days) in each operand form of the Legendre products.

Gates, per form:

* the first member of each class, over the last simulated day, is within
  0.5 K (global-mean surface temperature) and 1 W/m^2 (TOA net flux) of
  the FP64 oracle's day means on the same seeded inputs (fixtures
  tests/golden/climate_c4_member{0,32,56}.json from make_climate.py);
* every member's spectral prognostics are finite and its global-mean
  surface temperature lies in 150-400 K;
* a sample of members (one per class) equals, bit for bit, the same member
  stepped alone.
"""
import importlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEPS = 352
DAY = 32
FIXTURES = {0: "c4_member0", 32: "c4_member32", 56: "c4_member56"}
ALONE = (5, 40, 60)


def _fixture(case):
    path = os.path.join(GOLDEN, f"climate_{case}.json")
    if not os.path.exists(path):
        pytest.skip(f"{path} has not been generated")
    with open(path) as f:
        return json.load(f)


@pytest.mark.parametrize("form", ["plain", "split"])
def test_ensemble64_ten_days(exo, form):
    workloads = importlib.import_module(exo.__name__ + ".workloads")
    cfgs = workloads.ensemble(exo, seed=64)
    assert len(cfgs) == 64
    precision = exo.EXO_PRECISION_SPLIT if form == "split" else exo.EXO_PRECISION_FP32
    fixtures = {i: _fixture(c) for i, c in FIXTURES.items()}
    for i, fx in fixtures.items():
        assert fx["steps"] == STEPS
        for name, value in fx["config"].items():
            mine = getattr(cfgs[i], name)
            assert mine == value or (mine != mine and value != value), (i, name)  # NaN: null
    means = {i: ([], []) for i in fixtures}
    with exo.Engine(0, 64) as eng:
        ids = [eng.add_member(c, precision=precision) for c in cfgs]
        eng.step(STEPS - DAY)
        for _ in range(DAY):
            eng.step(1)
            eng.sync()
            for i in fixtures:
                ts, toa = eng.global_means(ids[i])
                means[i][0].append(ts)
                means[i][1].append(toa)
        for i, m in enumerate(ids):
            for name in ("sd", "st", "sz", "sq", "sp"):
                assert np.all(np.isfinite(eng.read_field(m, name))), (i, name)
            ts, toa = eng.global_means(m)
            assert 150.0 < ts < 400.0 and np.isfinite(toa), (i, ts, toa)
        batched = {i: eng.read_arena(ids[i]) for i in ALONE}
    for i, fx in fixtures.items():
        ts, toa = means[i]
        ref_ts, ref_toa = np.mean(fx["ts"][-DAY:]), np.mean(fx["toa"][-DAY:])
        c = cfgs[i]
        print(f"{form} member {i} T{c.truncation} L{c.levels}: last-day Ts {np.mean(ts):.3f} vs oracle "
              f"{ref_ts:.3f} K; TOA {np.mean(toa):.3f} vs {ref_toa:.3f} W/m^2")
        assert abs(np.mean(ts) - ref_ts) <= 0.5, (i, np.mean(ts), ref_ts)
        assert abs(np.mean(toa) - ref_toa) <= 1.0, (i, np.mean(toa), ref_toa)
    for i in ALONE:
        with exo.Engine(0, 1) as alone:
            m = alone.add_member(cfgs[i], precision=precision)
            alone.step(STEPS)
            alone.sync()
            assert np.array_equal(alone.read_arena(m), batched[i]), f"member {i} differs alone"
