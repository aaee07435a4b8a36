"""GPU parity of the step against the FP64 oracle, through the C-ABI.

Tolerances: the GPU computes in FP32, the oracle in FP64 on the same
seeded inputs and FP64 set-up. Spectral prognostics must agree per field
to a relative RMS of 2e-5 and a max-abs error of 1e-4 of the field's
max-abs after the start-up steps and early main-loop steps (the reference's
own kernel gate is rel-RMS <= 1e-5 per transform, scripts/exoplasim-webgpu/
verify-kernels.mjs:273; a whole step chains ~30 of them).
"""
import numpy as np
import pytest

from oracle_harness import OracleRun

pytestmark = pytest.mark.gpu

PROGNOSTIC = ("sd", "st", "sz", "sq", "sp", "sdm", "stm", "szm", "sqm", "spm")


def _compare(gpu, g, oracle, names, rel_rms=2e-5, rel_max=1e-4):
    out = {}
    for name in names:
        a = gpu.read_field(g, name).astype(np.float64)
        b = oracle.field(name)
        scale = max(np.sqrt(np.mean(b ** 2)), 1e-30)
        rms = np.sqrt(np.mean((a - b) ** 2)) / scale
        mx = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)
        out[name] = (rms, mx)
        print(f"{name}: rel-RMS {rms:.2e} rel-max {mx:.2e}")
        assert np.all(np.isfinite(a)), name
        assert rms < rel_rms and mx < rel_max, (name, rms, mx)
    return out


@pytest.mark.parametrize("trunc,levels,minutes", [(21, 10, 45), (42, 10, 45), (21, 5, 45), (63, 10, 30)])
def test_dynamics_parity(exo, trunc, levels, minutes):
    cfg = exo.default_config(truncation=trunc, levels=levels, timestep_minutes=minutes)
    if trunc == 42:
        cfg.land_fraction = 0.3
    with exo.Engine(0, 1) as gpu, exo.Engine(-1, 1) as host:
        g = gpu.add_member(cfg)
        h = host.add_member(cfg)
        oracle = OracleRun(host, h)
        steps = 5 if trunc < 63 else 4
        gpu.step(steps)
        gpu.sync()
        oracle.step(steps, physics=True)
        _compare(gpu, g, oracle, PROGNOSTIC)


def test_batch_members_match_alone(exo):
    """A mixed batch computes each member exactly as it would alone."""
    cfgs = [exo.default_config(truncation=21, levels=5),
            exo.default_config(truncation=42, levels=10, land_fraction=0.3),
            exo.default_config(truncation=21, levels=10, terrain_seed=7)]
    with exo.Engine(0, 4) as batch:
        ids = [batch.add_member(c) for c in cfgs]
        batch.step(6)
        batch.sync()
        together = [batch.read_arena(i) for i in ids]
    for c, t in zip(cfgs, together):
        with exo.Engine(0, 1) as alone:
            i = alone.add_member(c)
            alone.step(6)
            alone.sync()
            a = alone.read_arena(i)
        assert np.array_equal(a, t)
