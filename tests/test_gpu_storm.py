"""Storm diagnostics on the GPU (k_storm) against the oracle's hurricanemod
(GPU).

A T21 L10 planet with storm diagnostics on (GcmConfig.stormclim) and a warm
tropical band steps a day on the device. The storm kernel runs at the end
of the physics and reads only grid-point state that the rest of the step
leaves as it was (temperature, humidity, winds, vorticity, surface
pressure), so the oracle's storm routine, given the device's state after
the step, computes the same columns. Both work in FP32 (the oracle's FP32
build, in the reference's statement order); the CAPE and potential-
intensity iterations stop on thresholds (0.001 K, 0.5 hPa), so a column
whose iterate lands next to a threshold can take one more or one fewer
step: at least 99 % of the columns agree to 1e-3 relative (1e-3 absolute
for the indices near zero) and every column agrees to 5 %.
A planet with the diagnostics off leaves the storm fields at zero.
"""
import numpy as np
import pytest

from oracle_harness import OracleRun

pytestmark = pytest.mark.gpu

FIELDS = ("capen", "lnb", "mpoti", "chim", "laav", "venti", "vrmpi", "gpi", "shear")


def test_storm_diagnostics_match_the_oracle(exo):
    cfg = exo.default_config(truncation=21, levels=10, sea_surface_k=301.0, initial_temperature=285.0,
                             initial_humidity=0.01)
    cfg.storm_diagnostics = 1
    off = exo.default_config(truncation=21, levels=10)
    gpu = exo.Engine(0, 2)
    host = exo.Engine(-1, 1)
    try:
        g = gpu.add_member(cfg)
        quiet = gpu.add_member(off)
        h = host.add_member(cfg)
        gpu.step(32)
        gpu.sync()
        for name in FIELDS:
            assert np.all(gpu.read_field(quiet, name) == 0.0), name
        oracle = OracleRun(host, h, precision="fp32")
        oracle.arena[:] = gpu.read_arena(g).astype(oracle.dtype)
        for name in FIELDS:
            oracle.field(name)[:] = np.nan
        oracle.routine("storm")
        ran = np.asarray(oracle.field("capen")) >= 0
        print(f"columns with a CAPE computed: {ran.mean():.2f}; "
              f"max potential intensity up to {np.nanmax(oracle.field('mpoti')):.1f} m/s")
        assert ran.mean() > 0.2  # the warm band ran the intensity iteration
        for name in FIELDS:
            got = gpu.read_field(g, name).astype(np.float64)
            ref = oracle.field(name).astype(np.float64)
            assert np.all(np.isfinite(got)), name
            close = np.abs(got - ref) <= 1e-3 * np.abs(ref) + 1e-3
            loose = np.abs(got - ref) <= 5e-2 * np.abs(ref) + 1e-2
            print(f"{name:6s} within 1e-3: {close.mean():.4f}  within 5%: {loose.mean():.4f}")
            assert close.mean() >= 0.99, name
            assert loose.all(), name
    finally:
        gpu.close()
        host.close()
