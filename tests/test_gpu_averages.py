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
        # stopping frees the history
        eng.set_averaging(m, False)
        assert eng.month_count(m) == 0
    finally:
        eng.close()
