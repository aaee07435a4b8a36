"""GPU parity of the whole step against the FP64 oracle, through the C-ABI.

The GPU computes in FP32, the oracle in FP64 from the same FP64 set-up and
seeded inputs. Two gates:

* Step 1 (north-star level 2): after the first model step every spectral
  prognostic agrees with the FP64 oracle to 1e-5 relative RMS; temperature
  is measured in kelvin against the 250 K state it perturbs (its spectral
  deviation from the reference profile t0 is ~1e-5 of that).
* Later steps (the reference's replay rule, scripts/exoplasim-webgpu/
  verify-dynamics.mjs:56-59, :179-181): per field, rel-RMS(GPU - FP64) <=
  2 x rel-RMS(FP32 emulation - FP64) + 1e-6, where the FP32 emulation is
  the same oracle source compiled with float arithmetic, both RMS-averaged
  over every step of the run. Threshold physics makes single-precision
  trajectories diverge, which this rule measures instead of hiding. Two
  FP32 rounding sequences are equally the FP32 form: plain
  (oracle/liboracle32.so) and with multiply-adds fused, as the GPU
  compiler fuses them (oracle/liboracle32fma.so); the rule takes the
  larger of their errors. Fields that only a few columns carry differ
  between the two by 2x (large-scale rain in 10 columns of T21 L10 at
  steps 7-8: 5.6e-4 plain, 1.2e-3 fused).
"""
import numpy as np
import pytest

from oracle_harness import OracleRun

pytestmark = pytest.mark.gpu

SPECTRAL = ("sd", "st", "sz", "sq", "sp")
STATE = SPECTRAL + ("sdm", "stm", "szm", "sqm", "spm")
PHYSICS = ("dt", "dq", "dswfl", "dlwfl", "dshfl", "dlhfl", "devap", "dtaux", "dsoilt",
           "xts", "ysst", "dcc", "dprc", "dprl")

CASES = [(21, 10, 45, 0.4), (21, 5, 45, 0.4), (42, 10, 45, 0.3), (63, 10, 30, 0.4)]


def _rms(a, b):
    return np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300)


def _runs(exo, trunc, levels, minutes, land, steps):
    cfg = exo.default_config(truncation=trunc, levels=levels, timestep_minutes=minutes,
                             land_fraction=land)
    gpu = exo.Engine(0, 1)
    host = exo.Engine(-1, 1)
    g = gpu.add_member(cfg)
    h = host.add_member(cfg)
    o64 = OracleRun(host, h)
    o32 = OracleRun(host, h, precision="fp32")
    gpu.step(steps)
    gpu.sync()
    o64.step(steps, physics=True)
    o32.step(steps, physics=True)
    return gpu, g, host, o64, o32


@pytest.mark.parametrize("trunc,levels,minutes,land", CASES)
def test_first_step_within_1e5(exo, trunc, levels, minutes, land):
    gpu, g, host, o64, _ = _runs(exo, trunc, levels, minutes, land, 1)
    try:
        k = host.constants(0, 0)
        for name in SPECTRAL:
            a = gpu.read_field(g, name).astype(np.float64)
            b = o64.field(name)
            assert np.all(np.isfinite(a)), name
            if name == "st":
                # kelvin error against the 250 K isothermal state
                err = np.sqrt(np.mean((a - b) ** 2)) * k["ct"] / np.sqrt(2.0)
                rel = err / 250.0
            else:
                rel = _rms(a, b)
            print(f"T{trunc} L{levels} step 1 {name}: rel {rel:.2e}")
            assert rel < 1e-5, (name, rel)
    finally:
        gpu.close()
        host.close()


@pytest.mark.parametrize("trunc,levels,minutes,land", CASES)
def test_every_coefficient_within_1e5_per_level(exo, trunc, levels, minutes, land):
    """Max-norm per level and field, after the first step (a start-up
    step) and after the first main-loop step: |GPU - FP64|_inf <= 1e-5 x
    |FP64|_inf over each level's coefficients. Temperature's spectral field
    is the deviation from the reference profile, so each of its
    coefficients is measured in kelvin against the 250 K state it perturbs
    (|dst|_inf ct / sqrt 2 <= 1e-5 x 250 K).

    Divergence starts from rest and is a small residual of large,
    cancelling terms: sequential FP32 arithmetic itself (the oracle's FP32
    emulation) misses 1e-5 on its top levels, by up to 2.4e-5. Where the
    emulation's own max-norm error exceeds 1e-5, the GPU is held to that
    error instead (no worse than sequential FP32)."""
    cfg = exo.default_config(truncation=trunc, levels=levels, timestep_minutes=minutes, land_fraction=land)
    gpu = exo.Engine(0, 1)
    host = exo.Engine(-1, 1)
    try:
        g = gpu.add_member(cfg)
        h = host.add_member(cfg)
        o64 = OracleRun(host, h)
        o32 = OracleRun(host, h, precision="fp32")
        _, startup = host.phases(h)
        k = host.constants(h, 0)
        done = 0
        worst = {}
        bad = []
        for after in (1, startup + 1):
            gpu.step(after - done)
            gpu.sync()
            o64.step(after - done, physics=True)
            o32.step(after - done, physics=True)
            done = after
            for name in SPECTRAL:
                a = gpu.read_field(g, name).astype(np.float64)
                c = o32.field(name).astype(np.float64)
                b = o64.field(name)
                assert np.all(np.isfinite(a)), name
                planes = a.size // ((trunc + 1) * (trunc + 2))
                a, b, c = (x.reshape(planes, -1) for x in (a, b, c))
                for l in range(planes):
                    if name == "st":
                        scale = np.sqrt(2.0) * 250.0 / k["ct"]
                    else:
                        scale = max(np.max(np.abs(b[l])), 1e-300)
                    rel = np.max(np.abs(a[l] - b[l])) / scale
                    emu = np.max(np.abs(c[l] - b[l])) / scale
                    key = (after, name)
                    worst[key] = max(worst.get(key, (0.0, 0.0)), (rel, emu))
                    if not rel <= max(1e-5, emu):
                        bad.append((after, name, l, rel, emu))
        for (after, name), (rel, emu) in sorted(worst.items()):
            print(f"T{trunc} L{levels} after step {after} {name}: worst level max-norm GPU {rel:.2e} "
                  f"(FP32 emulation {emu:.2e})")
        assert not bad, bad
    finally:
        gpu.close()
        host.close()


@pytest.mark.parametrize("trunc,levels,minutes,land", CASES)
def test_steps_within_fp32_envelope(exo, trunc, levels, minutes, land):
    """Per field, the RMS over steps 1..N of the GPU's relative error must
    stay within twice the same RMS of the FP32 emulation's (single-step
    samples of either fluctuate by 10x as rounding errors cross
    thresholds)."""
    steps = 8 if trunc < 63 else 5
    cfg = exo.default_config(truncation=trunc, levels=levels, timestep_minutes=minutes,
                             land_fraction=land)
    gpu = exo.Engine(0, 1)
    host = exo.Engine(-1, 1)
    try:
        g = gpu.add_member(cfg)
        h = host.add_member(cfg)
        o64 = OracleRun(host, h)
        o32 = OracleRun(host, h, precision="fp32")
        o32f = OracleRun(host, h, precision="fp32fma")
        names = STATE + PHYSICS
        eg = {n: [] for n in names}
        ec = {n: [] for n in names}
        ef = {n: [] for n in names}
        for _ in range(steps):
            gpu.step(1)
            gpu.sync()
            for o in (o64, o32, o32f):
                o.step(1, physics=True)
            for n in names:
                a = gpu.read_field(g, n).astype(np.float64)
                b = o64.field(n)
                assert np.all(np.isfinite(a)), n
                eg[n].append(_rms(a, b))
                ec[n].append(_rms(o32.field(n).astype(np.float64), b))
                ef[n].append(_rms(o32f.field(n).astype(np.float64), b))
        bad = []
        for n in names:
            ga = np.sqrt(np.mean(np.square(eg[n])))
            ca = max(np.sqrt(np.mean(np.square(ec[n]))), np.sqrt(np.mean(np.square(ef[n]))))
            ok = ga <= 2 * ca + 1e-6
            print(f"T{trunc} L{levels} steps 1..{steps} {n}: GPU {ga:.2e} FP32-emulation {ca:.2e} {'ok' if ok else 'FAIL'}")
            if not ok:
                bad.append((n, ga, ca))
                print(f"  per step: GPU {' '.join(f'{x:.1e}' for x in eg[n])}; "
                      f"FP32 emulation {' '.join(f'{x:.1e}' for x in ec[n])}")
        assert not bad, bad
    finally:
        gpu.close()
        host.close()


def test_batch_members_match_alone(exo):
    """A mixed batch computes each member bit for bit as it would alone."""
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


def test_ensemble_planets_first_step(exo):
    """Planets drawn as the benchmark ensemble draws them (rotation period
    0.5-100 days, insolation 0.6-1.4 x 1367 W/m^2, 0.3-3 bar, land 0-60 %),
    two at each of T21 L5, T42 L10 and T63 L10 plus a T42 L5 and a T63 L5,
    stepped together in one batch: every member's first step is within
    1e-5 of the FP64 oracle started from its own set-up."""
    import importlib
    wl = importlib.import_module(exo.__name__ + ".workloads")
    rng = np.random.default_rng(2026)
    shapes = [(21, 5), (21, 5), (42, 10), (42, 10), (63, 10), (63, 10), (42, 5), (63, 5)]
    cfgs = [wl._planet(exo, rng, t, l, 45) for t, l in shapes]
    gpu = exo.Engine(0, len(cfgs))
    host = exo.Engine(-1, len(cfgs))
    try:
        gids = [gpu.add_member(c) for c in cfgs]
        hids = [host.add_member(c) for c in cfgs]
        gpu.step(1)
        gpu.sync()
        for (t, l), c, g, h in zip(shapes, cfgs, gids, hids):
            oracle = OracleRun(host, h)
            oracle.step(1, physics=True)
            k = host.constants(h, 0)
            for name in SPECTRAL:
                a = gpu.read_field(g, name).astype(np.float64)
                b = oracle.field(name)
                assert np.all(np.isfinite(a)), name
                if name == "st":
                    rel = np.sqrt(np.mean((a - b) ** 2)) * k["ct"] / np.sqrt(2.0) / 250.0
                else:
                    rel = _rms(a, b)
                assert rel < 1e-5, (t, l, c.rotation_period_days, c.pressure_bar, name, rel)
            print(f"T{t} L{l} rot {c.rotation_period_days:6.2f} d, {c.pressure_bar:4.2f} bar, "
                  f"land {c.land_fraction:.2f}: step 1 within 1e-5")
    finally:
        gpu.close()
        host.close()


@pytest.mark.parametrize("trunc,levels", [(21, 10), (42, 5)])
def test_reference_fresh_start(exo, trunc, levels):
    """The reference's fresh start (initfd: the restoration profile's mean
    temperature, planetary vorticity, no humidity, ln(ps) noise from
    gfortran's random numbers with the base namelist's SEED) under the
    reference's default atmosphere (1.011 bar of air, 300 ppm CO2): the
    first step within 1e-5 of the FP64 oracle, and 6 steps within the FP32
    envelope."""
    cfg = exo.default_config(truncation=trunc, levels=levels, initial_state=1, pressure_bar=1.011,
                             gas_constant=8314.46261815324 / 28.970253, co2_ppm=300,
                             eccentricity=0.016715, obliquity_deg=23.5)
    gpu = exo.Engine(0, 1)
    host = exo.Engine(-1, 1)
    try:
        g = gpu.add_member(cfg)
        h = host.add_member(cfg)
        o64 = OracleRun(host, h)
        o32 = OracleRun(host, h, precision="fp32")
        o32f = OracleRun(host, h, precision="fp32fma")
        k = host.constants(h, 0)
        names = STATE + PHYSICS
        eg = {n: [] for n in names}
        ec = {n: [] for n in names}
        for step in range(1, 7):
            gpu.step(1)
            gpu.sync()
            for o in (o64, o32, o32f):
                o.step(1, physics=True)
            for n in names:
                a = gpu.read_field(g, n).astype(np.float64)
                b = o64.field(n)
                assert np.all(np.isfinite(a)), (step, n)
                if step == 1 and n in SPECTRAL:
                    if n == "st":
                        rel = np.sqrt(np.mean((a - b) ** 2)) * k["ct"] / np.sqrt(2.0) / 250.0
                    else:
                        rel = _rms(a, b)
                    print(f"T{trunc} L{levels} fresh start, step 1 {n}: rel {rel:.2e}")
                    assert rel < 1e-5, (n, rel)
                eg[n].append(_rms(a, b))
                ec[n].append(max(_rms(o32.field(n).astype(np.float64), b),
                                 _rms(o32f.field(n).astype(np.float64), b)))
        bad = []
        for n in names:
            ga = np.sqrt(np.mean(np.square(eg[n])))
            ca = np.sqrt(np.mean(np.square(ec[n])))
            print(f"T{trunc} L{levels} fresh start, steps 1..6 {n}: GPU {ga:.2e} FP32-emulation {ca:.2e}")
            if not ga <= 2 * ca + 1e-6:
                bad.append((n, ga, ca))
        assert not bad, bad
    finally:
        gpu.close()
        host.close()
