"""Seeded synthetic workloads of the time step.

* `plain(exo, truncation, levels, minutes)`: one planet with the default
  Earth-like parameters (1 bar N2, 360 ppm CO2, 1367 W/m^2, 23.4 deg
  obliquity, 40 % land), the configuration the single-run figures use.
* `ensemble(exo, seed)`: the 64-planet mixed batch, 32 at T21 L5, 24 at
  T42 L10 and 8 at T63 L10, 45-minute steps, with per-planet parameters
  drawn from `seed`: rotation period log-uniform 0.5-100 days, insolation
  uniform 0.6-1.4 x 1367 W/m^2, surface pressure log-uniform 0.3-3 bar and
  land fraction uniform 0-0.6; terrain and noise seeds from the same draw.
* `serving_requests(seed, rate, seconds)`: Poisson arrivals of runs drawing
  truncation T21/T42/T63 with weights 0.5/0.35/0.15, 5 or 10 levels, a
  length uniform in 1-360 simulated days and ensemble parameters.
"""
import math

import numpy as np

ENSEMBLE_MIX = ((21, 5, 32), (42, 10, 24), (63, 10, 8))


def steps_per_day(minutes):
    """Model steps per 86,400 s day (mtspd, setup/pumamod/index.ts:153-158)."""
    return int(round(1440.0 / minutes))


def plain(exo, truncation=63, levels=10, minutes=30, **overrides):
    return exo.default_config(truncation=truncation, levels=levels, timestep_minutes=minutes, **overrides)


def _planet(exo, rng, truncation, levels, minutes):
    rot = math.exp(rng.uniform(math.log(0.5), math.log(100.0)))
    flux = rng.uniform(0.6, 1.4) * 1367.0
    bar = math.exp(rng.uniform(math.log(0.3), math.log(3.0)))
    land = rng.uniform(0.0, 0.6)
    seeds = rng.integers(1, 2**62, size=2)
    return exo.default_config(truncation=truncation, levels=levels, timestep_minutes=minutes,
                              rotation_period_days=rot, flux_wm2=flux, pressure_bar=bar,
                              land_fraction=land, terrain_seed=int(seeds[0]), noise_seed=int(seeds[1]))


def ensemble(exo, seed=64, minutes=45, mix=ENSEMBLE_MIX):
    """Configs of the mixed ensemble, in mix order."""
    rng = np.random.default_rng(seed)
    out = []
    for trunc, levels, count in mix:
        for _ in range(count):
            out.append(_planet(exo, rng, trunc, levels, minutes))
    return out


def serving_requests(exo, seed=5, rate=4.0, seconds=600.0, minutes=45):
    """[(arrival_s, config, days)] of a Poisson stream."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        trunc = int(rng.choice([21, 42, 63], p=[0.5, 0.35, 0.15]))
        levels = int(rng.choice([5, 10]))
        days = float(rng.uniform(1.0, 360.0))
        out.append((t, _planet(exo, rng, trunc, levels, minutes), days))
