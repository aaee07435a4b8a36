"""Device averages against the oracle's averager (GPU).

A T21 L5 planet publishes one snapshot a day for a little over a model
year. Every published snapshot is also read back and fed to the oracle's
restatement of GCM_AVERAGES (oracle/averages.py), so the two see the same
FP32 snapshots in the same order:

* every closed month (its calendar month, sample count and last step, and
  its mean) equals the oracle's bit for bit: both sum in FP64 in publish
  order and round the mean to FP32;
* the annual mean equals the oracle's bit for bit (month / 12 added in
  FP64 and rounded to FP32 month by month);
* the yearly balance (global means of the annual mean) agrees to 1e-12:
  the device sums its points in a tree, the oracle row by row.

A saved averager (the checkpoint's: open month sum, closed months, yearly
balances) loads back unchanged, and averaging continues from it as the
oracle continues from the same state; a loaded history's climatology
equals the oracle's bit for bit.
"""
import numpy as np
import pytest

from oracle import averages as oracle_avg

pytestmark = pytest.mark.gpu


def test_device_averages_match_the_oracle(exo):
    cfg = exo.default_config(truncation=21, levels=5, timestep_minutes=45)
    eng = exo.Engine(0, 1)
    try:
        m = eng.add_member(cfg)
        every = 24 * 60 // 45  # one snapshot a model day
        eng.set_publish_every(every)
        eng.set_averaging(m)
        _, nlat, nlon, _ = eng.dims(m)
        ref = oracle_avg.MonthlyAverager(np.polynomial.legendre.leggauss(nlat)[1])
        # run to just after a publishing step: the snapshot slot then holds
        # it, and the next step's date is the same day as the snapshot's
        while eng.member_step(m) % every != 1:
            eng.step(1)
        days = 0
        while not ref.years and days < 420:
            eng.sync()
            step = eng.member_step(m) - 1
            date = eng.member_date(m)
            ref.add(eng.read_snapshot(m), date[0], date[1], step)
            eng.step(every)
            days += 1
        eng.sync()
        assert ref.years, f"no calendar year closed in {days} days"
        print(f"{days} model days; {len(ref.months)} months; years {[y['year'] for y in ref.years]}")
        assert eng.month_count(m) == len(ref.months)
        for i, (info, mean) in enumerate(ref.months):
            got_info, got = eng.read_month(m, i)
            assert got_info == info, i
            np.testing.assert_array_equal(got, mean.reshape(got.shape))
        annual = eng.read_annual(m)
        np.testing.assert_array_equal(annual, ref.annual().reshape(annual.shape))
        years = eng.yearly_balances(m)
        assert [y["year"] for y in years] == [y["year"] for y in ref.years]
        for got, want in zip(years, ref.years):
            for key in ("toa_net_radiation", "surface_net_heat", "surface_temperature"):
                assert got[key] == pytest.approx(want[key], rel=1e-12, abs=1e-12), key
        assert 150 < years[0]["surface_temperature"] < 400
        assert exo.balance_assess(years)["status"] == "spinning-up"
        # one complete year: its climatology is that year's months
        n_years, climate = oracle_avg.climatology(ref.months)
        for month in (1, 12):
            info, mean, count = eng.read_climatology(m, month)
            assert count == n_years == 1 and info == climate[month - 1][0]
            np.testing.assert_array_equal(mean, climate[month - 1][1].reshape(mean.shape))
        # stopping frees the history
        eng.set_averaging(m, False)
        assert eng.month_count(m) == 0
    finally:
        eng.close()


def _run_days(eng, m, days, every, ref=None):
    """Steps `days` publishing intervals; with ref, each new snapshot also
    goes to the oracle (the one already in the slot was averaged before)."""
    for _ in range(days):
        eng.step(every)
        if ref is not None:
            eng.sync()
            date = eng.member_date(m)
            ref.add(eng.read_snapshot(m), date[0], date[1], eng.member_step(m) - 1)
    eng.sync()


def _same_averager(a, b):
    assert (a["current"] is None) == (b["current"] is None)
    if a["current"] is not None:
        assert {k: v for k, v in a["current"].items() if k != "sum"} == \
            {k: v for k, v in b["current"].items() if k != "sum"}
        np.testing.assert_array_equal(a["current"]["sum"], b["current"]["sum"])
    assert len(a["months"]) == len(b["months"])
    for (ia, va), (ib, vb) in zip(a["months"], b["months"]):
        assert ia == ib
        np.testing.assert_array_equal(va, vb)
    assert a["years"] == b["years"]


def test_saved_averager_loads_and_continues(exo):
    cfg = exo.default_config(truncation=21, levels=5, timestep_minutes=45)
    eng = exo.Engine(0, 1)
    try:
        m = eng.add_member(cfg)
        every = 24 * 60 // 45
        eng.set_publish_every(every)
        eng.set_averaging(m)
        _, nlat, _, _ = eng.dims(m)
        while eng.member_step(m) % every != 1:
            eng.step(1)
        _run_days(eng, m, 45, every)
        saved = eng.averager(m)
        assert saved["current"] is not None and len(saved["months"]) >= 1
        print(f"saved: {len(saved['months'])} months, open month {saved['current']['month']} "
              f"with {saved['current']['samples']} samples")
        eng.set_averaging(m, False)
        eng.load_averager(m, saved)
        _same_averager(eng.averager(m), saved)
        # the oracle continues from the same state
        ref = oracle_avg.MonthlyAverager(np.polynomial.legendre.leggauss(nlat)[1])
        cur = saved["current"]
        ref.current = [cur["year"], cur["month"], cur["sum"].copy(), cur["samples"], cur["step"]]
        ref.months = list(saved["months"])
        _run_days(eng, m, 40, every, ref)
        assert eng.month_count(m) == len(ref.months) > len(saved["months"])
        for i, (info, mean) in enumerate(ref.months):
            got_info, got = eng.read_month(m, i)
            assert got_info == info, i
            np.testing.assert_array_equal(got.reshape(-1), mean.reshape(-1))
        info, total = eng.read_accumulator(m)
        assert info == {"year": ref.current[0], "month": ref.current[1], "samples": ref.current[3],
                        "step": ref.current[4]}
        np.testing.assert_array_equal(total.reshape(-1), ref.current[2].reshape(-1))
    finally:
        eng.close()


def test_loaded_history_climatology(exo):
    cfg = exo.default_config(truncation=21, levels=5)
    eng = exo.Engine(0, 1)
    try:
        m = eng.add_member(cfg)
        _, n = eng.slot(m, "snapshot")
        nhor = eng.dims(m)[1] * eng.dims(m)[2]
        rng = np.random.default_rng(7)
        months = []
        for year in range(2, 15):  # the history keeps years 4..14; the climate the last ten
            for month in range(1, 13):
                months.append(({"year": year, "month": month, "samples": 30, "step": 360 * year + 30 * month},
                               rng.normal(280.0, 30.0, size=(n // nhor, nhor)).astype(np.float32)))
        months = months[-132:]
        years = [{"year": y, "toa_net_radiation": 0.5 * y, "surface_net_heat": -0.25 * y,
                  "surface_temperature": 280.0 + y} for y in range(4, 14)]
        eng.load_averager(m, {"current": None, "months": months, "years": years})
        assert eng.read_accumulator(m) == (None, None)
        assert eng.yearly_balances(m) == years
        n_years, climate = oracle_avg.climatology(months)
        print(f"{len(months)} months held, {n_years} complete years")
        assert n_years == 10 and climate[0][0]["year"] == 14
        for month in range(1, 13):
            info, mean, count = eng.read_climatology(m, month)
            assert count == n_years and info == climate[month - 1][0]
            np.testing.assert_array_equal(mean, climate[month - 1][1])
        with pytest.raises(exo.ExoError):
            eng.load_averager(m, {"current": None, "months": months + months[:1], "years": []})
    finally:
        eng.close()


def test_averaging_members_in_a_mixed_batch(exo):
    """Averaging is per member: in a batch of a T21 L5 planet that averages
    and a T42 L10 planet that does not, the first's months equal the
    oracle's fed with its own snapshots and the second holds none; the
    second starts its months when averaging is switched on; removing an
    averaging member frees it while the other goes on."""
    every = 24 * 60 // 45
    eng = exo.Engine(0, 2)
    try:
        a = eng.add_member(exo.default_config(truncation=21, levels=5, timestep_minutes=45))
        b = eng.add_member(exo.default_config(truncation=42, levels=10, timestep_minutes=45))
        eng.set_publish_every(every)
        eng.set_averaging(a)
        nlat = eng.dims(a)[1]
        ref = oracle_avg.MonthlyAverager(np.polynomial.legendre.leggauss(nlat)[1])
        while eng.member_step(a) % every != 1:
            eng.step(1)
        eng.sync()
        date = eng.member_date(a)
        ref.add(eng.read_snapshot(a), date[0], date[1], eng.member_step(a) - 1)
        _run_days(eng, a, 35, every, ref)
        assert eng.month_count(b) == 0 and eng.month_count(a) == len(ref.months) >= 1
        for i, (info, mean) in enumerate(ref.months):
            got_info, got = eng.read_month(a, i)
            assert got_info == info
            np.testing.assert_array_equal(got.reshape(-1), mean.reshape(-1))
        eng.set_averaging(b)
        _run_days(eng, a, 40, every)
        assert eng.month_count(b) >= 1
        info, _ = eng.read_accumulator(b)
        assert info is not None and info["samples"] >= 1
        eng.remove_member(a)
        eng.step(every)
        eng.sync()
        assert np.all(np.isfinite(eng.read_snapshot(b)))
    finally:
        eng.close()
