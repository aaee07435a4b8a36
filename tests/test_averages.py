"""Averages and balance on the CPU: the engine's balance test
(exo_balance_assess, host C++) and the oracle's restatement of the
averager (oracle/averages.py), pinned by closed forms.

balance (GCM_BALANCE.assess, balance/index.ts:6-68):
* fewer than 14 years: spinning up, with the years still missing;
* a series a + b k (positive) has 10-year running means a + b (k - 4.5),
  whose 5-point least-squares slopes are exactly b: the drift is |b|, and
  b against 5e-4 W/m^2/yr decides balanced or drifting;
* running means are taken in absolute value (a negative constant is as
  balanced as a positive one);
* only the last 30 slopes count: a ramp that ended long enough ago is
  forgotten.

averager (GCM_AVERAGES): constant snapshots average to the constant; the
month closes when the calendar month changes and holds the samples' mean;
the annual mean of equal months is the month; the diurnal range is the
mean of sum / days; a year's balance appears when December closes.
"""
import math

import numpy as np
import pytest

from oracle import averages as oracle_avg


def test_balance_spinning_up(exo):
    for n in (0, 1, 13):
        out = exo.balance_assess([(1.0, 1.0)] * n)
        assert out["status"] == "spinning-up" and out["years_needed"] == 14 - n
        assert math.isnan(out["toa_drift_per_year"])


@pytest.mark.parametrize("b", [1e-4, 3e-3, -2e-3])
def test_balance_linear_drift_is_the_slope(exo, b):
    years = np.arange(60)
    toa = 2.0 + b * years
    surface = 1.0 + 0.5 * b * years
    out = exo.balance_assess(list(zip(toa, surface)))
    assert out["toa_drift_per_year"] == pytest.approx(abs(b), rel=1e-9)
    assert out["surface_drift_per_year"] == pytest.approx(abs(0.5 * b), rel=1e-9)
    want = "balanced" if abs(b) < 5e-4 and abs(0.5 * b) < 5e-4 else "drifting"
    assert out["status"] == want
    ref = oracle_avg.assess(toa, surface)
    assert ref["toa_drift_per_year"] == pytest.approx(out["toa_drift_per_year"], rel=1e-12)


def test_balance_uses_absolute_running_means(exo):
    out = exo.balance_assess([(-0.7, -0.3)] * 20)
    assert out["status"] == "balanced" and out["toa_drift_per_year"] == 0.0
    # a series crossing zero: |mean| folds the drift back
    years = np.arange(40)
    toa = -0.1 + 0.01 * years
    ours = exo.balance_assess(list(zip(toa, np.zeros(40))))
    ref = oracle_avg.assess(toa, np.zeros(40))
    assert ours["toa_drift_per_year"] == pytest.approx(ref["toa_drift_per_year"], rel=1e-12)
    assert ours["toa_drift_per_year"] < 0.01  # folding lowers it below the raw slope


def test_balance_forgets_old_ramps(exo):
    # a 1 W/m^2/yr ramp for 20 years, then flat for 50: the last 30 slopes
    # see only the flat part
    toa = np.concatenate([np.arange(20) * 1.0, np.full(50, 19.0)])
    out = exo.balance_assess(list(zip(toa, np.ones(70))))
    assert out["status"] == "balanced" and out["toa_drift_per_year"] == 0.0
    short = exo.balance_assess(list(zip(toa[:40], np.ones(40))))
    assert short["status"] == "drifting"


def _weights(nlat):
    return np.polynomial.legendre.leggauss(nlat)[1]


def test_oracle_averager_closed_forms():
    nlat, nlon, fields = 4, 8, 42
    avg = oracle_avg.MonthlyAverager(_weights(nlat))
    rng = np.random.default_rng(0)
    month_values = {}
    step = 0
    for year in (0, 1):
        for month in range(1, 13):
            base = rng.normal(size=(fields, nlat * nlon)).astype(np.float32)
            base[oracle_avg.DAILY_RANGE_DAYS] = 2.0
            base[oracle_avg.DAILY_RANGE_SUM] = 8.0 + month
            month_values[(year, month)] = base
            for _ in range(3):  # three equal snapshots per month
                avg.add(base, year, month, step)
                step += 1
    avg.add(month_values[(0, 1)], 2, 1, step)  # closes December of year 1
    assert len(avg.months) == 24
    for info, mean in avg.months:
        assert info["samples"] == 3
        np.testing.assert_array_equal(mean, month_values[(info["year"], info["month"])])
    assert [y["year"] for y in avg.years] == [0, 1]
    annual = avg.annual()
    months = [month_values[(1, m)] for m in range(1, 13)]
    np.testing.assert_allclose(annual[:35], np.mean(months, axis=0)[:35], rtol=1e-6, atol=1e-6)
    expect_range = np.mean([(8.0 + m) / 2.0 for m in range(1, 13)])
    np.testing.assert_allclose(annual[oracle_avg.DIURNAL_RANGE], expect_range, rtol=1e-6)
    np.testing.assert_array_equal(annual[oracle_avg.DAILY_RANGE_DAYS], 1.0)
    # the balance is the area-weighted mean of the annual field
    w = np.repeat(_weights(nlat), nlon)
    toa = float(np.sum(w * annual[oracle_avg.TOA_NET_RADIATION].astype(np.float64)) / np.sum(w))
    assert avg.years[1]["toa_net_radiation"] == pytest.approx(toa, rel=1e-12)


def test_averaging_needs_a_device_engine(exo):
    host = exo.Engine(-1, 1)
    m = host.add_member(exo.default_config(truncation=21, levels=5))
    with pytest.raises(exo.ExoError):
        host.set_averaging(m)
    assert host.month_count(m) == 0
    assert host.yearly_balances(m) == []
    host.close()


def test_oracle_climatology_closed_forms():
    # constant months worth 100 * year + month: the climatology of month k
    # over the last ten complete years is 100 * (mean of those years) + k
    nh = 6
    months = []
    for year in range(3, 16):
        for month in range(1, 13):
            months.append(({"year": year, "month": month, "samples": 1, "step": 12 * year + month},
                           np.full((2, nh), 100.0 * year + month, dtype=np.float32)))
    # year 16 holds twelve months with one repeated: not complete
    for month in [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 11]:
        months.append(({"year": 16, "month": month, "samples": 1, "step": 0},
                       np.full((2, nh), -1.0, dtype=np.float32)))
    years, climate = oracle_avg.climatology(months[-oracle_avg.HISTORY_MONTHS:])
    # the history keeps 132 months: years 6..15 complete plus year 16's twelve
    assert years == 10
    for k, (info, mean) in enumerate(climate, start=1):
        assert info["year"] == 15 and info["month"] == k
        np.testing.assert_array_equal(mean, np.float32(100.0 * np.mean(np.arange(6, 16)) + k))
    assert oracle_avg.climatology(months[:11]) == (0, [])
