"""Pins of the oracle's remaining parameterization details (CPU), each by a
property that a plausible slip (a mistyped coefficient or exponent, a
wrong branch point, a swapped argument, a dropped term) breaks:

* lwr's band absorptivities (lwr.wgsl:36-179): H2O, CO2 and O3 absorb
  nothing along a zero path and the H2O transmission in the CO2 band is 1
  there; each fit is continuous across its branch point (H2O and O3 at
  0.01, CO2 at 1, the H2O transmission at 2), the condition the reference
  derives its offsets from (physics/radiation/index.ts:108-125), and
  continuous everywhere else on a dense path grid; absorptivities rise
  with the path and stay within [0, 1], the transmission falls.
* swr's ozone and water-vapour absorptions (Lacis and Hansen's fits,
  swr.wgsl): zero along a zero path, rising and saturating (concave) with
  the path, at most the whole band over the paths a column can hold.
* the sea-ice surface temperature (ice.wgsl): the implicit step closes
  the energy balance of the top ice layer, rho_i c_i h_min (T - T_old) /
  dt = net surface heat + conduction from the ocean through the slab
  (k_i / h_i without snow; the thickness-weighted conductivity over the
  ice-plus-snow depth with snow; none past the thickness cap); above the
  melting point the surplus melts instead.
* the glacier albedo (surface.wgsl): the albedo runs linearly from its
  maximum at 263.16 K down to its minimum at the melting point, per band.
* the wet-soil albedo (surface.wgsl, NWETSOIL): a dry bucket has the dry
  land albedo, a full one the ocean's; in between the albedo a inverts the
  retention curve, bucket fraction = (1 + (5.2 a)^4)^(-4/3); frozen soil
  has the land albedo.
* the cloud liquid water (mkclouds.wgsl): where there is cloud, the liquid
  water density rho_l = q_l p / (R T) is Kiehl et al.'s (1998, CCM3)
  profile 0.21 g/m^3 exp(-z / h_l) with the scale height h_l = 700 m x
  ln(1 + W), W the column water vapour in kg/m^2, at the levels'
  hypsometric heights.
* the weathering supply limit (carbonstep.wgsl, NSUPPLY): the limited
  weathering follows the unlimited one where it is small against the
  supply cap, approaches the cap where it is large, rises with it, and
  W = -cap ln(1 - W_lim / cap) recovers the unlimited rate.
* the storm's potential intensity (storm.wgsl:140-208): the outflow ratio
  is SST over the environment's temperature at the saturated parcel's
  level of neutral buoyancy, and the intensity is Bister and Emanuel's
  sqrt(Ck/Cd (T_s/T_o) (CAPE* - CAPE)).

* the two coefficients no property fixes, the CO2 absorptivity's slope per
  decade above its branch point and the ozone absorption's ultraviolet
  exponent: the oracle's rise per decade and its large-path log-log slope
  recover the literals read out of the reference's WGSL text
  (tests/golden/make_fit_constants.py, lwr.wgsl:122, swr.wgsl:36).

FP64 oracle: the identities hold to rounding, continuity to the size of a
step of the dense grid.
"""
import math

import numpy as np
import pytest

from oracle_harness import OracleRun


@pytest.fixture(scope="module")
def member(exo):
    host = exo.Engine(-1, 1)
    cfg = exo.default_config(truncation=21, levels=10, star_temperature_k=0.0, vegetation=0, glaciers=0)
    cfg.storm_diagnostics = 1
    m = host.add_member(cfg)
    yield host, m
    host.close()


def _run(member):
    host, m = member
    run = OracleRun(host, m)
    return run, run.dims[1] * run.dims[2], run.dims[3]


def _planes(run, name, nh):
    return run.field(name).reshape(-1, nh)


def _continuous(a):
    """No jump on a dense grid: every increment is within 3x of the larger
    of its neighbours (a branch taken at the wrong point, or a branch whose
    offset is wrong, jumps by orders of magnitude more)."""
    d = np.abs(np.diff(a))
    nb = np.maximum(np.concatenate([[d[1]], d[:-1]]), np.concatenate([d[1:], [d[-2]]]))
    return np.all(d <= 3 * nb + 1e-15)


LWR = {"lwr_h2o": 0.01, "lwr_co2": 1.0, "lwr_o3": 0.01, "lwr_h2o_transmission": 2.0}


@pytest.mark.parametrize("fit", sorted(LWR))
def test_lwr_fit_zero_path_and_branch_continuity(member, fit):
    run, _, _ = _run(member)
    at0 = run.radiation_fit(fit, np.array([0.0]))[0]
    if fit == "lwr_h2o_transmission":
        assert abs(at0 - 1.0) < 1e-12, at0
    else:
        assert abs(at0) < 1e-12, at0
    b = LWR[fit]
    below, above = run.radiation_fit(fit, np.array([b, b * (1 + 1e-12)]))
    assert abs(above - below) < 1e-9, (fit, below, above)
    u = np.logspace(-9, 1, 40001)
    a = run.radiation_fit(fit, u)
    assert _continuous(a), fit
    if fit == "lwr_h2o_transmission":
        assert np.all(np.diff(a) <= 0) and np.all(a >= 0) and a[0] <= 1
    else:
        assert np.all(np.diff(a) >= 0) and np.all(a >= 0) and np.all(a <= 1), fit


@pytest.mark.parametrize("fit,top", [("swr_o3", 12.0), ("swr_h2o", 1000.0)])
def test_swr_absorption_zero_rising_saturating(member, fit, top):
    run, _, _ = _run(member)
    assert run.radiation_fit(fit, np.array([0.0]))[0] == 0.0
    u = np.linspace(0.0, top, 20001)
    a = run.radiation_fit(fit, u)
    assert np.all(np.diff(a) > 0), fit
    assert np.all(np.diff(a, 2) <= 1e-15), fit  # concave: the band saturates
    assert np.all(a <= 1.0)
    # the absorbed fraction of the band grows far slower than the path
    r = run.radiation_fit(fit, np.array([1e-4, 1e-3, 1e-2, 1e-1, 1.0]))
    assert np.all(np.diff(np.log10(r)) < 1.0)


def _ice_columns(run, nh, rng, snow):
    L = run.dims[3]
    run.field("dls")[:] = 0.0
    run.field("xls")[:] = 0.0
    hi = rng.uniform(0.2, 8.0, nh)
    hi[::7] = 9.5  # past the thickness cap (xmaxd = 9 m): no conduction
    run.field("xiced")[:] = hi
    run.field("xicec")[:] = 1.0
    run.field("xicecc")[:] = 1.0
    run.field("xsnow")[:] = rng.uniform(0.0, 0.3, nh) if snow else 0.0
    run.field("xsst")[:] = rng.uniform(270.0, 272.0, nh)
    run.field("xts")[:] = rng.uniform(230.0, 275.0, nh)
    run.field("dprs")[:] = 0.0
    run.field("sea_cprsa")[:] = 0.0
    run.field("sea_cheata")[:] = 0.0
    run.field("sea_naccua")[:] = 0.0
    heat = rng.uniform(-300.0, 300.0, nh)
    for name in ("dshfl", "dlhfl"):
        run.field(name)[:] = 0.0
    _planes(run, "dswfl", nh)[L] = heat
    _planes(run, "dlwfl", nh)[L] = 0.0
    return heat


@pytest.mark.parametrize("snow", [False, True])
def test_ice_surface_temperature_closes_the_energy_balance(member, snow):
    run, nh, L = _run(member)
    rng = np.random.default_rng(11 + snow)
    heat = _ice_columns(run, nh, rng, snow)
    hi = run.field("xiced").copy()
    hs = 1e3 / run.constant("crhosn") * run.field("xsnow")
    sst = run.field("xsst").copy()
    t_old = np.minimum(run.field("xts"), run.constant("tmelt"))
    run.routine("sea_ice")
    ts = run.field("xts")
    melt = run.field("xqmelt")
    ki, ks = run.constant("ckapi"), run.constant("ckapsn")
    kbar = (ks * hs + ki * hi) / (hs + hi)
    if not snow:
        np.testing.assert_allclose(kbar, ki, rtol=1e-14)
    conductance = np.where(hi >= run.constant("xmaxd"), 0.0, kbar / (hi + hs))
    store = run.constant("xmind") * run.constant("crhoi") * run.constant("cpi") / run.constant("xdt")
    tm = run.constant("tmelt")
    frozen = ts < tm
    assert frozen.any() and (~frozen).any()
    # below the melting point: storage = surface heat + conduction, no melt
    lhs = store * (ts - t_old)
    rhs = heat + conductance * (sst - ts)
    np.testing.assert_allclose(lhs[frozen], rhs[frozen], rtol=1e-10, atol=1e-9)
    assert np.all(melt[frozen] == 0)
    # at the melting point the surplus melts
    np.testing.assert_allclose(ts[~frozen], tm)
    np.testing.assert_allclose(melt[~frozen], (heat + conductance * (sst - tm) - store * (tm - t_old))[~frozen],
                               rtol=1e-10, atol=1e-9)
    assert np.all(melt[~frozen] >= -1e-9)


def _land_columns(run, nh, ts):
    run.field("dls")[:] = 1.0
    run.field("land_dts")[:] = ts
    run.field("dsnowz")[:] = 0.0
    run.field("dforest")[:] = 0.0
    run.field("dp")[:] = 1e5


def test_glacier_albedo_ramp(member):
    run, nh, _ = _run(member)
    t0, tm = 263.16, run.constant("tmelt")
    ts = np.linspace(240.0, 290.0, nh)
    ts[:5] = [t0, tm, 0.5 * (t0 + tm), t0 - 1e-3, tm + 1e-3]
    _land_columns(run, nh, ts)
    run.field("dglac")[:] = 1.0
    # distinct band albedos, so that a band reading another's is seen
    for suffix, (mx, mn) in (("1", (0.85, 0.55)), ("2", (0.7, 0.45))):
        run.set_constant("albgmax" + suffix, mx)
        run.set_constant("albgmin" + suffix, mn)
    run.routine("land")
    for field, band, suffix in (("dalb", 0, ""), ("dsalb", 0, "1"), ("dsalb", 1, "2")):
        a = run.field(field).reshape(-1, nh)[band]
        mx, mn = run.constant("albgmax" + suffix), run.constant("albgmin" + suffix)
        assert mx > mn
        assert abs(a[0] - mx) < 1e-12 and abs(a[1] - mn) < 1e-12, (field, band)
        assert abs(a[2] - 0.5 * (mx + mn)) < 1e-12
        assert a[3] == mx and a[4] == mn
        order = np.argsort(ts)
        assert np.all(np.diff(a[order]) <= 1e-15)
        inside = (ts > t0) & (ts < tm)
        np.testing.assert_allclose(a[inside], mx + (mn - mx) * (ts[inside] - t0) / (tm - t0), rtol=1e-12)


def test_wet_soil_albedo_curve(member):
    run, nh, _ = _run(member)
    run.set_constant("nwetsoil", 1.0)
    wsmax = run.constant("wsmax")
    ts = np.full(nh, 285.0)
    ts[::4] = 260.0  # frozen soil
    ts[2::8] = 268.0
    ts[3::16] = 273.2  # just above freezing
    _land_columns(run, nh, ts)
    run.field("dglac")[:] = 0.0
    run.set_constant("doceanalb1", 0.06)
    run.set_constant("doceanalb2", 0.09)
    bf = np.linspace(0.0, 1.0, nh)
    bf[1], bf[5] = 0.0, 1.0
    run.field("dwater")[:] = bf * wsmax
    run.routine("land")
    a = run.field("dalbclim")
    warm = ts > 273.16
    assert warm[1] and warm[5]
    assert abs(a[1] - run.constant("alblandmax")) < 1e-12
    assert abs(a[5] - run.constant("doceanalb1")) < 1e-12
    # the second band: the same curve, ending at the ocean's band-2 albedo
    assert abs(run.field("dalbclim2")[5] - run.constant("doceanalb2")) < 1e-12
    np.testing.assert_allclose(run.field("dalbclim1")[warm], a[warm], rtol=1e-14)
    mid = warm & (bf > 0.15) & (bf < 0.85)  # both the dry and wet weights vanish
    np.testing.assert_allclose((1 + (5.2 * a[mid]) ** 4) ** (-4.0 / 3.0), bf[mid], rtol=1e-10)
    np.testing.assert_allclose(a[~warm], run.constant("albland"), rtol=1e-14)
    assert np.all(np.diff(a[mid]) < 0)  # wetter soil is darker


def test_cloud_liquid_water_is_the_ccm3_profile(member):
    run, nh, L = _run(member)
    rng = np.random.default_rng(5)
    sigma, sigmah = run.table("sigma"), run.table("sigmah")
    rd, ga = run.constant("gascon"), run.constant("ga")
    ps = rng.uniform(7e4, 1.04e5, nh)
    run.field("dp")[:] = ps
    t = 288.0 * sigma[:, None] ** 0.19 + rng.uniform(-5, 5, (L, nh))
    q = 0.015 * sigma[:, None] ** 3 * rng.uniform(0.2, 1.0, nh)
    dt = _planes(run, "dt", nh)
    dq = _planes(run, "dq", nh)
    dt[:L] = t
    dq[:L] = q
    for name in ("dtdt", "dqdt", "dw", "icclev", "icctot", "dprc"):
        run.field(name)[:] = 0.0
    _planes(run, "dq", nh)[:L] *= 1.0
    # near-saturated layers make stratiform cloud in most columns
    run.routine("mkclouds")
    cover = _planes(run, "dcc", nh)[:L]
    ql = _planes(run, "dql", nh)[:L]
    w = (q * np.diff(np.concatenate([[0.0], sigmah]))[:, None]).sum(0) * ps / ga
    np.testing.assert_allclose(run.field("dqvi"), w, rtol=1e-12)
    hl = 700.0 * np.log(1.0 + w)
    # hypsometric heights of the full levels: mid-layer (in height) below
    # the top layer, and half the log-distance above the top half level
    thick = rd / ga * t[1:] * np.log(sigmah[1:] / sigmah[:-1])[:, None]
    zhalf = np.vstack([np.cumsum(thick[::-1], axis=0)[::-1], np.zeros((1, nh))])  # half levels 0..L-1
    z = np.empty((L, nh))
    z[1:] = zhalf[1:] + 0.5 * thick
    z[0] = zhalf[0] + 0.5 * rd / ga * t[0] * np.log(sigmah[0] / sigma[0])
    rho = ql * sigma[:, None] * ps / (rd * t)
    cloudy = cover > 0
    assert cloudy.sum() > nh
    expect = np.maximum(0.21e-3 * np.exp(-z / hl), 1e-9 * sigma[:, None] * ps / (rd * t))
    np.testing.assert_allclose(rho[cloudy], expect[cloudy], rtol=1e-10)
    assert np.all(ql[~cloudy] == 0)


def test_weathering_supply_limit(member):
    run, nh, L = _run(member)
    rng = np.random.default_rng(9)
    run.field("dls")[:] = 1.0
    run.field("carbon_step")[0] = 1.0
    run.field("dp")[:] = 1e5
    _planes(run, "dt", nh)[L] = rng.uniform(274.0, 320.0, nh)
    run.field("dprc")[:] = 10.0 ** rng.uniform(-16, -6, nh)  # precipitation spans ten decades
    run.field("dprl")[:] = 0.0
    scale = run.constant("plarad") ** 2 / (run.constant("rad_earthsq") * 0.29)
    run.set_constant("nsupply", 0.0)
    run.routine("carbon")
    w = run.field("localweathering") / scale
    cap = float(np.median(w))
    run.set_constant("wmax", cap * run.constant("vearth"))
    run.set_constant("nsupply", 1.0)
    run.routine("carbon")
    wl = run.field("localweathering") / scale
    assert w.min() < 0.05 * cap and w.max() > 20 * cap
    assert np.all(wl <= w * (1 + 1e-12)) and np.all(wl <= cap)
    small = w < 1e-3 * cap
    np.testing.assert_allclose(wl[small], w[small], rtol=1e-3)
    large = w > 20 * cap
    np.testing.assert_allclose(wl[large], cap, rtol=1e-8)
    order = np.argsort(w)
    assert np.all(np.diff(wl[order]) >= 0)
    ok = wl < cap * (1 - 1e-6)
    np.testing.assert_allclose(-cap * np.log(1 - wl[ok] / cap), w[ok], rtol=1e-6)


def test_storm_outflow_ratio_and_intensity(member):
    """A tropical sounding topped by an isothermal 205 K layer: the
    saturated parcel's level of neutral buoyancy lies in it, so the outflow
    temperature is 205 K exactly and the ratio is SST / 205 K."""
    run, nh, L = _run(member)
    sigma = run.table("sigma")
    rd = run.constant("gascon")
    run.field("dp")[:] = 101325.0
    sst = np.linspace(299.0, 303.0, nh)
    dt = _planes(run, "dt", nh)
    dq = _planes(run, "dq", nh)
    t = sst[None, :] * sigma[:, None] ** (rd * 0.0065 / 9.81)
    t = np.maximum(t, 205.0)
    dt[:L] = t
    dt[L] = sst
    p = 1013.25 * sigma[:, None]
    es = 6.112 * np.exp(17.67 * (t - 273.15) / (243.5 + t - 273.15))
    eps = rd / 461.5
    dq[:L] = 0.8 * eps * es / (p - es)
    dq[L] = dq[L - 1]
    cols = np.arange(0, nh, 97)
    checked = 0
    for h in cols:
        parts = run.storm_potential(h)
        if parts is None:
            continue
        ratio, tout, pout, capes, cape, vmax = parts
        # the outflow temperature is the environment's at that level, linear
        # in pressure between the model levels
        assert tout == pytest.approx(np.interp(pout, p[:, 0], t[:, h]), rel=1e-12)
        assert ratio == pytest.approx(sst[h] / tout, rel=1e-12)
        pl = p[:, 0]
        below = int(np.searchsorted(pl, pout))  # first level at or beneath the outflow
        bracket = t[max(below - 1, 0):below + 1, h]
        if not np.all(bracket == 205.0):
            continue  # neutral buoyancy below the isothermal layer
        assert tout == pytest.approx(205.0, rel=1e-12)
        assert ratio == pytest.approx(sst[h] / 205.0, rel=1e-12)
        assert capes > cape
        assert vmax == pytest.approx(math.sqrt(run.constant("ckcd") * ratio * (capes - cape)), rel=1e-12)
        checked += 1
    assert checked >= 5, checked


def _fit_constants():
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "radiation_fit_constants.json")
    with open(path) as f:
        return json.load(f)


def test_co2_log_branch_slope_is_the_reference_literal(member):
    """Above the branch point the CO2 absorptivity is C log10(u) + offset:
    its rise over each decade of path is C, which must equal the literal
    read from lwr.wgsl:122 (tests/golden/make_fit_constants.py)."""
    run, _, _ = _run(member)
    c = _fit_constants()["lwr_co2_log_slope"]["value"]
    a = run.radiation_fit("lwr_co2", np.array([2.0, 20.0, 200.0, 2000.0]))
    assert np.allclose(np.diff(a), c, rtol=0, atol=1e-12), (np.diff(a), c)


def test_ozone_uv_exponent_is_the_reference_literal(member):
    """For large ozone paths the ultraviolet term 1.082 x / (1 + 138.6 x)^E
    dominates the ozone absorption (the other two terms fall as 1/x and
    1/x^2), so the absorption grows as x^(1 - E): the log-log slope between
    two large paths recovers E, which must equal the literal read from
    swr.wgsl:36 (tests/golden/make_fit_constants.py)."""
    run, _, _ = _run(member)
    e = _fit_constants()["swr_o3_uv_exponent"]["value"]
    a = run.radiation_fit("swr_o3", np.array([1e12, 1e13]))
    measured = 1.0 - math.log10(a[1] / a[0])
    assert abs(measured - e) < 1e-7, (measured, e)
