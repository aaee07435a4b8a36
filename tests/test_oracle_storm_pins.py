"""Pins of the oracle's storm diagnostics (hurricanemod, physics/storm/
wgsl/storm.wgsl) by what the meteorology fixes, on crafted columns (CPU):

* level choice and kinematics: the 850, 600 and 200 hPa levels are the
  sigma levels nearest 0.85, 0.6 and 0.2; the shear is the difference of
  the wind speeds there and the absolute-vorticity field is the 850 hPa
  relative vorticity times the rotation rate;
* a saturated 600 hPa level has no entropy deficit (chi = 0); a dry
  column has no CAPE and its neutral-buoyancy level is unset (-1);
* over cool water (SST <= 5 C) nothing is computed (-1) and the
  ventilation index is 1, whose reduced value is 0;
* a tropical sounding (SST 300 K, 6.5 K/km, 80 % humidity) has CAPE of
  the observed order (10^2..10^4 J/kg), a level of neutral buoyancy in
  the upper troposphere and a maximum potential intensity of tens of m/s
  (Emanuel 1986 gives about 70 m/s at 300 K); a warmer or moister surface
  parcel is more buoyant (more CAPE), a colder one less;
* the reduced ventilation index x is the real root of x^3 - x + 2 v /
  (3 sqrt(3) v_t) = 0 (its trigonometric form) for 0 < v < v_t.
FP64 oracle: the checks hold to rounding.
"""
import math

import numpy as np
import pytest

from oracle_harness import OracleRun

OUT = ("capen", "lnb", "mpoti", "chim", "laav", "venti", "vrmpi", "gpi", "shear")


@pytest.fixture(scope="module")
def member(exo):
    host = exo.Engine(-1, 1)
    cfg = exo.default_config(truncation=21, levels=10)
    cfg.storm_diagnostics = 1
    m = host.add_member(cfg)
    yield host, m
    host.close()


def _planes(run, name, nh):
    return run.field(name).reshape(-1, nh)


def _bolton(tc):
    return 6.112 * np.exp(17.67 * tc / (243.5 + tc))


def _tropical(run, nh, sst=300.0, rh=0.8, lapse=0.0065):
    """dt/dq planes (levels + surface) of a standard-lapse sounding."""
    L = run.dims[3]
    sigma = run.table("sigma")
    rd = run.constant("gascon")
    dp = run.field("dp")
    dp[:] = 101325.0
    dt = _planes(run, "dt", nh)
    dq = _planes(run, "dq", nh)
    t = sst * sigma ** (rd * lapse / 9.81)
    t = np.maximum(t, 200.0)
    p = 1013.25 * sigma
    e = rh * _bolton(t - 273.15)
    q = 0.622 * e / (p - e)
    dt[:L] = t[:, None]
    dq[:L] = q[:, None]
    dt[L] = sst
    dq[L] = q[-1]
    for name in ("du", "dv", "gz"):
        _planes(run, name, nh)[:] = 0.0
    return t, q


def _nearest(sigma, target):
    best = 0
    for l in range(len(sigma)):
        if abs(sigma[l] - target) < abs(sigma[best] - target):
            best = l
    return best


def test_storm_flag_and_constants(member):
    host, m = member
    run = OracleRun(host, m)
    assert run.constant("nstormdiag") == 1.0
    assert run.constant("ckcd") == pytest.approx(0.9)
    assert run.constant("vithresh") == pytest.approx(0.145)


def test_levels_shear_and_vorticity(member):
    host, m = member
    run = OracleRun(host, m)
    nh, L = run.dims[1] * run.dims[2], run.dims[3]
    _tropical(run, nh)
    rng = np.random.default_rng(11)
    du, dv, gz = (_planes(run, n, nh) for n in ("du", "dv", "gz"))
    du[:L] = rng.normal(0, 15, (L, nh))
    dv[:L] = rng.normal(0, 15, (L, nh))
    gz[:L] = rng.normal(0, 1e-5, (L, nh))
    run.routine("storm")
    sigma = run.table("sigma")
    i850, i200 = _nearest(sigma, 0.85), _nearest(sigma, 0.2)
    assert abs(sigma[i850] - 0.85) < 0.1 and abs(sigma[i200] - 0.2) < 0.1
    want = np.abs(np.hypot(du[i200], dv[i200]) - np.hypot(du[i850], dv[i850]))
    np.testing.assert_allclose(run.field("shear"), want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(run.field("laav"), gz[i850] * run.constant("ww"), rtol=1e-12)


def test_saturated_600_has_no_deficit_and_dry_has_no_cape(member):
    host, m = member
    run = OracleRun(host, m)
    nh, L = run.dims[1] * run.dims[2], run.dims[3]
    _tropical(run, nh)
    i600 = _nearest(run.table("sigma"), 0.6)
    _planes(run, "dq", nh)[i600] = 0.5  # far above saturation at 600 hPa
    run.routine("storm")
    assert np.all(run.field("chim") == 0.0)
    # a dry column
    run = OracleRun(host, m)
    _tropical(run, nh)
    _planes(run, "dq", nh)[:] = 0.0
    run.routine("storm")
    assert np.all(run.field("capen") == 0.0) and np.all(run.field("lnb") == -1.0)
    assert np.all(run.field("chim") == 1.0)


def test_cool_water_is_skipped(member):
    host, m = member
    run = OracleRun(host, m)
    nh = run.dims[1] * run.dims[2]
    _tropical(run, nh, sst=278.0)
    run.routine("storm")
    for name in ("capen", "lnb", "mpoti"):
        assert np.all(run.field(name) == -1.0), name
    assert np.all(run.field("venti") == 1.0) and np.all(run.field("vrmpi") == 0.0)


def test_tropical_cape_and_potential_intensity(member):
    host, m = member
    nh = None
    runs = {}
    for label, sst, rh in (("base", 300.0, 0.8), ("moist", 300.0, 0.9), ("warm", 302.0, 0.8), ("cool", 296.0, 0.8)):
        run = OracleRun(host, m)
        nh = run.dims[1] * run.dims[2]
        _tropical(run, nh, sst=sst, rh=rh)
        run.routine("storm")
        runs[label] = run
    base = runs["base"]
    cape, lnb, vmax = (float(base.field(n)[0]) for n in ("capen", "lnb", "mpoti"))
    assert 100.0 < cape < 1e4
    assert 100.0 < lnb < 500.0  # hPa
    assert 30.0 < vmax < 150.0
    assert float(runs["moist"].field("capen")[0]) > cape
    assert float(runs["warm"].field("capen")[0]) > cape
    assert float(runs["cool"].field("capen")[0]) < cape
    assert float(runs["warm"].field("mpoti")[0]) > vmax
    # every column of a horizontally uniform input agrees
    for n in OUT:
        assert np.all(base.field(n) == base.field(n)[0]), n


def test_reduced_ventilation_is_the_cubic_root(member):
    host, m = member
    run = OracleRun(host, m)
    nh, L = run.dims[1] * run.dims[2], run.dims[3]
    _tropical(run, nh)
    sigma = run.table("sigma")
    i850, i200 = _nearest(sigma, 0.85), _nearest(sigma, 0.2)
    du = _planes(run, "du", nh)
    du[i850] = 5.0
    du[i200] = 5.0 + np.linspace(0.0, 80.0, nh)  # shear 0 .. 80 m/s
    run.routine("storm")
    vt = run.constant("vithresh")
    v = run.field("venti")
    x = run.field("vrmpi")
    inside = (v > 0) & (v < vt)
    assert inside.sum() > nh // 10 and np.any(v >= vt)
    resid = x[inside] ** 3 - x[inside] + 2.0 * v[inside] / (3.0 * math.sqrt(3.0) * vt)
    assert np.max(np.abs(resid)) < 1e-12
    assert np.all(x[~inside] == 0.0)


@pytest.fixture(scope="module")
def member5(exo):
    host = exo.Engine(-1, 1)
    cfg = exo.default_config(truncation=21, levels=5)
    cfg.storm_diagnostics = 1
    m = host.add_member(cfg)
    yield host, m
    host.close()


def test_dry_adiabat_cape_closed_form(member5):
    """A hot, nearly dry surface parcel (r = 1.2e-6) has its lifting
    condensation level above the top of an L5 column, so it rises dry-
    adiabatically, T = T_s (p / p_s)^(R / c_pd), to every level. An
    environment 1 K colder than that (same mixing ratio) gives it the
    density excess b = 1 K (1 + r / eps) / (1 + r) at every level above the
    surface one, and CAPE is the trapezoid sum R sum (b_j + b_j+1)
    (p_j+1 - p_j) / (p_j + p_j+1)."""
    host, m = member5
    run = OracleRun(host, m)
    nh, L = run.dims[1] * run.dims[2], run.dims[3]
    sigma = run.table("sigma")
    rd = run.constant("gascon")
    assert 1013.25 * sigma[0] > 80.0  # the condensation level (about 75 hPa) is above the top
    run.field("dp")[:] = 101325.0
    p = 1013.25 * sigma
    ts, r = 320.0, 1.2e-6
    adiabat = ts * (p / p[-1]) ** (rd / 1005.7)
    dt, dq = _planes(run, "dt", nh), _planes(run, "dq", nh)
    dt[:L] = (adiabat - np.where(np.arange(L) < L - 1, 1.0, 0.0))[:, None]
    dq[:L] = r
    dt[L] = 300.0
    dq[L] = r
    for name in ("du", "dv", "gz"):
        _planes(run, name, nh)[:] = 0.0
    run.routine("storm")
    eps = rd / 461.5
    b = np.where(np.arange(L) < L - 1, 1.0, 0.0) * (1 + r / eps) / (1 + r)
    want = rd * np.sum((b[:-1] + b[1:]) * (p[1:] - p[:-1]) / (p[:-1] + p[1:]))
    np.testing.assert_allclose(run.field("capen"), want, rtol=1e-9)
    # buoyant up to the top: the neutral-buoyancy level is the top level
    np.testing.assert_allclose(run.field("lnb"), p[0], rtol=1e-12)


def test_saturation_decides_the_entropy_deficit(member):
    """Bolton's (1980) saturation vapour pressure e_s = 6.112 exp(17.67 T_c
    / (T_c + 243.5)) hPa sets the saturation mixing ratio r_s = eps e_s /
    (p - e_s) at 600 hPa: just above it the level counts as saturated (chi
    = 0), just below it does not (chi > 0)."""
    host, m = member
    for factor, saturated in ((1.002, True), (0.998, False)):
        run = OracleRun(host, m)
        nh = run.dims[1] * run.dims[2]
        t, q = _tropical(run, nh)
        sigma = run.table("sigma")
        i600 = _nearest(sigma, 0.6)
        p600 = 1013.25 * sigma[i600]
        eps = run.constant("gascon") / 461.5
        es = _bolton(t[i600] - 273.15)
        _planes(run, "dq", nh)[i600] = factor * eps * es / (p600 - es)
        run.routine("storm")
        chi = run.field("chim")
        assert np.all(chi == 0.0) if saturated else np.all(chi > 0.0), factor
