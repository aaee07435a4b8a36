"""CPU restatement of the reference's snapshot averages and balance test:
TEST INFRASTRUCTURE ONLY (the tests compare the engine's device averages
with it; nothing shipped imports it).

* MonthlyAverager follows GCM_AVERAGES (src/model/climate/gcm/averages/
  index.ts): add() (:81-101) sums each snapshot into an FP64 month
  accumulator (open, :71-79) and, when the calendar month changes, closes
  it (close, :50-69: FP32 mean = sum / samples) into a history of
  HISTORY_MONTHS months (:21-23); a close that completes twelve months of
  one calendar year ending in month 12 adds that year's balance (yearly,
  :31-48: global means of the annual mean's TOA net radiation, surface net
  heat and surface temperature). annual() is :103-137 (a Float32Array
  accumulating month / 12; the diurnal temperature range as the months'
  daily ranges, fields/index.ts:492-497).
* climatology is GCM_AVERAGES.view's climatology (:177-222): the complete
  calendar years held (twelve distinct months of one year), the last
  CLIMATE_YEARS of them, each calendar month's mean over those years
  (FP64 sum oldest year first, the mean rounded to FP32).
* global_mean is GCM_GRID.globalMean (grid/index.ts:99-111): latitude
  weights, in FP64, rows in order.
* assess is GCM_BALANCE.assess (balance/index.ts:6-68): ExoPlaSim's
  _isbalanced drift test.
"""
import math

import numpy as np

MONTHS_PER_YEAR = 12
CLIMATE_YEARS = 10
HISTORY_MONTHS = (CLIMATE_YEARS + 1) * MONTHS_PER_YEAR
SURFACE_TEMPERATURE, TOA_NET_RADIATION, SURFACE_NET_HEAT = 0, 16, 17
DIURNAL_RANGE, DAILY_RANGE_SUM, DAILY_RANGE_DAYS = 37, 38, 39


def global_mean(values, latitude_weights):
    """values (NLAT*NLON,) row-major; latitude_weights (NLAT,)."""
    nlat = len(latitude_weights)
    rows = np.asarray(values, dtype=np.float64).reshape(nlat, -1)
    total = 0.0
    weight = 0.0
    for row in range(nlat):
        total += float(np.sum(rows[row] * latitude_weights[row]))
        weight += latitude_weights[row] * rows.shape[1]
    return total / weight


class MonthlyAverager:
    def __init__(self, latitude_weights):
        self.weights = np.asarray(latitude_weights, dtype=np.float64)
        self.current = None  # (year, month, sum FP64, samples, step)
        self.months = []     # (info dict, FP32 mean)
        self.years = []

    def add(self, snapshot, year, month, step):
        """One published snapshot (fields, NHOR) FP32 with its calendar
        month; returns True when a month closed."""
        values = np.asarray(snapshot, dtype=np.float32)
        cur = self.current
        if cur is not None and cur[0] == year and cur[1] == month:
            cur[2] += values.astype(np.float64)
            cur[3] += 1
            cur[4] = step
            return False
        self.current = [year, month, values.astype(np.float64), 1, step]
        if cur is None:
            return False
        mean = (cur[2] / cur[3]).astype(np.float32)
        info = {"year": cur[0], "month": cur[1], "samples": cur[3], "step": cur[4]}
        self.months = (self.months + [(info, mean)])[-HISTORY_MONTHS:]
        year_balance = self.yearly()
        if year_balance is not None:
            self.years.append(year_balance)
        return True

    def yearly(self):
        last12 = self.months[-MONTHS_PER_YEAR:]
        if len(last12) < MONTHS_PER_YEAR:
            return None
        last = last12[-1][0]
        if last["month"] != MONTHS_PER_YEAR or any(m[0]["year"] != last["year"] for m in last12):
            return None
        annual = self.annual()
        return {"year": last["year"],
                "toa_net_radiation": global_mean(annual[TOA_NET_RADIATION], self.weights),
                "surface_net_heat": global_mean(annual[SURFACE_NET_HEAT], self.weights),
                "surface_temperature": global_mean(annual[SURFACE_TEMPERATURE], self.weights)}

    def annual(self):
        if len(self.months) < MONTHS_PER_YEAR:
            return None
        months = [m[1] for m in self.months[-MONTHS_PER_YEAR:]]
        values = np.zeros_like(months[0], dtype=np.float32)
        for month in months:
            values = (values.astype(np.float64) + month.astype(np.float64) / MONTHS_PER_YEAR).astype(np.float32)
        rng = np.zeros_like(values[DAILY_RANGE_SUM])
        for month in months:
            s, d = month[DAILY_RANGE_SUM], month[DAILY_RANGE_DAYS]
            with np.errstate(divide="ignore", invalid="ignore"):
                daily = np.where(d > 0, (s.astype(np.float64) / d.astype(np.float64)).astype(np.float32),
                                 month[DIURNAL_RANGE])
            rng = (rng.astype(np.float64) + daily.astype(np.float64) / MONTHS_PER_YEAR).astype(np.float32)
        values[DAILY_RANGE_SUM] = rng
        values[DAILY_RANGE_DAYS] = 1.0
        values[DIURNAL_RANGE] = rng
        return values


def climatology(months):
    """months: [(info dict, FP32 mean)] oldest first. Returns (years,
    [(info, FP32 mean)] for calendar months 1..12) or (0, [])."""
    grouped = {}
    for info, mean in months:
        grouped.setdefault(info["year"], []).append((info, mean))
    complete = [held for held in grouped.values()
                if len(held) == MONTHS_PER_YEAR and len({i["month"] for i, _ in held}) == MONTHS_PER_YEAR]
    complete.sort(key=lambda held: held[0][0]["year"])
    complete = complete[-CLIMATE_YEARS:]
    if not complete:
        return 0, []
    out = []
    for month in range(1, MONTHS_PER_YEAR + 1):
        samples = [next(m for m in held if m[0]["month"] == month) for held in complete]
        total = np.zeros(samples[-1][1].shape, dtype=np.float64)
        for _, mean in samples:
            total += mean.astype(np.float64)
        out.append((samples[-1][0], (total / len(samples)).astype(np.float32)))
    return len(complete), out


RUNNING_MEAN_YEARS, SLOPE_WINDOW, SLOPES_AVERAGED = 10, 5, 30
THRESHOLD_W_M2_PER_YEAR = 5e-4
MINIMUM_YEARS = RUNNING_MEAN_YEARS + SLOPE_WINDOW - 1


def _slope(values):
    n = len(values)
    x_mean = (n + 1) / 2
    y_mean = sum(values) / n
    num = sum((i + 1 - x_mean) * (v - y_mean) for i, v in enumerate(values))
    den = sum((i + 1 - x_mean) ** 2 for i in range(n))
    return num / den


def _drift(series):
    running = [abs(sum(series[e - RUNNING_MEAN_YEARS:e]) / RUNNING_MEAN_YEARS)
               for e in range(RUNNING_MEAN_YEARS, len(series) + 1)]
    slopes = [_slope(running[e - SLOPE_WINDOW:e]) for e in range(SLOPE_WINDOW, len(running) + 1)]
    last = slopes[-SLOPES_AVERAGED:]
    return abs(sum(last) / len(last))


def assess(toa, surface):
    if len(toa) < MINIMUM_YEARS:
        return {"status": "spinning-up", "years_needed": MINIMUM_YEARS - len(toa),
                "toa_drift_per_year": math.nan, "surface_drift_per_year": math.nan}
    dt, ds = _drift(list(toa)), _drift(list(surface))
    return {"status": "balanced" if dt < THRESHOLD_W_M2_PER_YEAR and ds < THRESHOLD_W_M2_PER_YEAR else "drifting",
            "years_needed": 0, "toa_drift_per_year": dt, "surface_drift_per_year": ds}
