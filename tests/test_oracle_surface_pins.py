"""Pins of the set-up and surface pieces that the rest of GcmConfig switches
on (CPU), each against what the physics or the definitions fix:

* the star's spectrum (SPECTRUM.split, physics/radiation/spectrum/index.ts:
  58-168): the energy shares of the bands below and above 0.75 um are
  Planck integrals, the Rayleigh coefficient is the star's lambda^-4-
  weighted mean over the Sun's, normalized to the Sun's default band
  share, and each surface's band albedo is its spectral albedo averaged
  over the star in each band (numerical integrals here, by adaptive
  quadrature, against the reference's 1024-point trapezoids);
* GCM_ATMOSPHERE.resolve (atmosphere/index.ts:84-160): R / mu of a mix,
  the native atmosphere, CO2 as a mixing ratio, and the reference's errors;
* the Keplerian orbit (calendar/solar/index.ts:13-66): the declination and
  distance factor follow Kepler's equation, solved here by bisection; the
  synchronous sun stands still over the substellar longitude;
* the surface set-up's reads (setup/surface/index.ts:76-107): a 12-month
  climatology gains December before and January after, other counts fail
  with the reference's message; without a surface-temperature climatology
  the ocean starts at the freezing point under sea ice;
* vegstep (physics/vegetation/wgsl/vegstep.wgsl): litter moves carbon from
  the vegetation to the soil pool without loss, soil respiration doubles
  per 10 K (Q10 = 2), frozen ground grows nothing, forest cover rises with
  vegetation carbon, nveg = 1 feeds nothing back;
* glacierstep (physics/glacier/wgsl/glacierstep.wgsl): land that kept
  glacelim of snow through a year becomes glacier at the year's end.
"""
import math
import os
import re

import numpy as np
import pytest
from scipy import integrate, optimize

from oracle_harness import OracleRun

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HC_OVER_K = 0.0143877735383
LMIN = 316.036116751e-9
HINGE = 7.5e-7


def _planck(lam, t):
    return lam ** -5 / math.expm1(HC_OVER_K / (lam * t))


def _integral(f, a, b):
    # adaptive quadrature over log-spaced pieces
    edges = np.geomspace(a, b, 41)
    return sum(integrate.quad(f, x0, x1, epsrel=1e-12, limit=200)[0] for x0, x1 in zip(edges[:-1], edges[1:]))


def _specblock():
    text = open(os.path.join(ROOT, "web-exoplasim_amd", "csrc", "specblock.inc")).read()
    out = {}
    for name, body in re.findall(r"kSpec_(\w+)\[kSpecPoints\] = \{([^}]*)\}", text):
        out[name] = np.array([float(x) for x in body.replace("\n", " ").split(",") if x.strip()])
    return out


@pytest.mark.parametrize("star", [3500.0, 5780.0, 9000.0])
def test_star_spectrum_split_is_planck(exo, star):
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config(star_temperature_k=star))
        k = e.constants(m, 0)
    z1 = _integral(lambda x: _planck(x, star), LMIN, HINGE)
    z2 = _integral(lambda x: _planck(x, star), HINGE, 1e-4)
    assert k["nstartemp"] == 1
    assert k["zsolar1"] == pytest.approx(z1 / (z1 + z2), abs=2e-4)
    assert k["zsolar1"] + k["zsolar2"] == pytest.approx(1.0, abs=1e-15)
    assert k["zsolars1"] == k["zsolar1"] and k["zsolars2"] == k["zsolar2"]
    # Rayleigh: the star's lambda^-4 moment over its first band, against
    # the Sun's over its whole spectrum
    cross = _integral(lambda x: _planck(x, star) * (x * 1e6) ** -4, LMIN, 1e-4)
    sun = 5772.0
    sun_cross = _integral(lambda x: _planck(x, sun) * (x * 1e6) ** -4, LMIN, 1e-4)
    sun_all = _integral(lambda x: _planck(x, sun), LMIN, 1e-4)
    rcoeff = cross * 0.517 / z1 / (sun_cross / sun_all)
    assert k["rcoeff"] == pytest.approx(rcoeff, rel=5e-4)
    # band albedos: the tabulated spectral albedo (per cent) averaged over
    # the star in each band, then one broadband value for both bands
    spec = _specblock()
    lam_um = spec["wavelengths"]
    b = np.array([_planck(x * 1e-6, star) for x in lam_um])
    one = lam_um <= 0.75 + 1e-12
    for table, const in (("groundblend", "dgroundalb1"), ("oceanblend", "doceanalb1"),
                         ("seaicemin", "dicealbmn1"), ("seaicemax", "dicealbmx1"),
                         ("glacalbmin", "albgmin1"), ("iceblendmax", "albsmax1"), ("iceblendmin", "albsmin1")):
        alb = spec[table]
        a1 = 0.01 * np.trapezoid(b[one] * alb[one], lam_um[one] * 1e-6) / z1
        two = ~one
        two[np.argmax(two) - 1] = True  # the band-two integral starts at the last band-one point
        a2 = 0.01 * np.trapezoid(b[two] * alb[two], lam_um[two] * 1e-6) / z2
        broad = k["zsolar1"] * a1 + k["zsolar2"] * a2
        assert k[const] == pytest.approx(broad, rel=3e-4), table
        assert k[const.replace("1", "2")] == k[const], table  # simple albedo: one value


def test_grey_split_without_a_star(exo):
    with exo.Engine(-1, 1) as e:
        k = e.constants(e.add_member(exo.default_config(star_temperature_k=0.0)), 0)
    assert k["nstartemp"] == 0 and k["zsolar1"] == 0.517 and k["zsolar2"] == 0.483 and k["rcoeff"] == 1
    assert k["dgroundalb1"] == 0.2 and k["doceanalb1"] == 0.069 and k["albgmin1"] == 0.6


def test_atmosphere_resolve_closed_forms(exo):
    r = 8314.46261815324
    bar, gas, co2 = exo.atmosphere_resolve()
    assert (bar, co2) == (1.011, 300.0) and gas == pytest.approx(r / 28.970253, rel=1e-15)
    bar, gas, co2 = exo.atmosphere_resolve({"N2": 0.8, "O2": 0.2})
    assert bar == pytest.approx(1.0) and gas == pytest.approx(r / (0.8 * 28.0134 + 0.2 * 31.9988), rel=1e-14)
    bar, gas, co2 = exo.atmosphere_resolve({"N2": 0.999, "CO2": 0.001})
    assert co2 == pytest.approx(1000.0, rel=1e-12)
    bar, gas, co2 = exo.atmosphere_resolve({"CO2": 2.0})
    assert bar == 2.0 and co2 == 1e6 and gas == pytest.approx(r / 44.01, rel=1e-15)
    bar, gas, co2 = exo.atmosphere_resolve({"N2": 0.5}, mode="pressureAndGasConstant", pressure_bar=3.0)
    assert bar == 3.0 and gas == pytest.approx(r / 28.970253, rel=1e-15)  # gas constant still Earth's
    for kwargs, message in (
            (dict(partial_pressures={"Ar": -0.1}), "Argon (Ar) partial pressure must be finite and nonnegative."),
            (dict(partial_pressures={"N2": 2.0}, mode="pressureAndGasConstant", pressure_bar=1.0),
             "The specified partial pressures exceed total surface pressure."),
            (dict(mode="pressureAndGasConstant", pressure_bar=0.0),
             "Total pressure and gas constant must be finite and positive.")):
        with pytest.raises(exo.ExoError) as err:
            exo.atmosphere_resolve(**kwargs)
        assert str(err.value) == message


def test_keplerian_orbit_follows_keplers_equation(exo):
    e_, obl, periapse, m0 = 0.3, 30.0, 75.0, 40.0
    cfg = exo.default_config(keplerian=1, eccentricity=e_, obliquity_deg=obl, periapse_longitude_deg=periapse,
                             initial_mean_anomaly_deg=m0)
    steps_per_year = round(cfg.year_days * 86400 / (cfg.timestep_minutes * 60))
    with exo.Engine(-1, 1) as e:
        m = e.add_member(cfg)
        for step in (0, 1000, 4321, 9000, 11000):
            v = e.step_values(m, step)
            mean = (step % steps_per_year) / steps_per_year * 2 * math.pi + math.radians(m0)
            mean = mean % (2 * math.pi)
            ea = optimize.brentq(lambda x: x - e_ * math.sin(x) - mean, -1.0, 2 * math.pi + 1, xtol=1e-15)
            nu = 2 * math.atan(math.sqrt((1 + e_) / (1 - e_)) * math.tan(ea / 2))
            lon = math.radians(periapse + 180) + nu
            assert v[0] == pytest.approx(math.sin(math.radians(obl)) * math.sin(lon), abs=1e-9), step
            assert v[3] == pytest.approx((1 / (1 - e_ * math.cos(ea))) ** 2, rel=1e-9), step


def test_synchronous_sun_stands_over_the_substellar_point(exo):
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config(synchronous=1, substellar_longitude_deg=90.0))
        phases = {e.step_values(m, s)[2] for s in (0, 7, 31, 500)}
        k = e.constants(m, 0)
    pi = 3.14159265359  # pumamod's
    assert phases == {(1 - 90.0 / 360) * 2 * pi}
    assert k["rad_synchronous"] == 1


def _initial(e, m, name):
    layout = {n: (o, c) for n, o, c in e.layout(m)}
    off, n = layout[name]
    return e.read_table(m, "initial")[off:off + n]


def test_surface_reads_expand_and_fail_like_the_reference(exo):
    cfg = exo.default_config()
    nh = 64 * 32
    months = np.repeat(np.arange(1.0, 13.0), nh).reshape(12, nh) + 270.0
    with exo.Engine(-1, 2) as e:
        m = e.add_member_surface(cfg, {169: months})
        sst = _initial(e, m, "xclsst").reshape(14, nh)
        assert np.array_equal(sst[:, 0], [282.0] + [270.0 + i for i in range(1, 13)] + [271.0])
        with pytest.raises(exo.ExoError) as err:
            e.add_member_surface(cfg, {169: months[:3]})
        assert str(err.value) == "N032_surf_0169.sra has 3 fields, expected 14"


def test_ocean_without_sst_starts_frozen_under_ice(exo):
    cfg = exo.default_config(sea_surface_k=float("nan"))
    with exo.Engine(-1, 1) as e:
        m = e.add_member(cfg)
        k = e.constants(m, 0)
        sea = _initial(e, m, "dls") < 0.5
        assert sea.any()
        np.testing.assert_array_equal(_initial(e, m, "yclsst2")[sea], k["tfreeze"])
        np.testing.assert_array_equal(_initial(e, m, "dicec")[sea], 1.0)
        assert np.all(_initial(e, m, "diced")[sea] > 0)


@pytest.fixture(scope="module")
def land(exo):
    host = exo.Engine(-1, 1)
    m = host.add_member(exo.default_config(truncation=21, levels=10, star_temperature_k=0.0))
    yield host, m
    host.close()


def _land_run(land, nveg=2):
    host, m = land
    run = OracleRun(host, m)
    run.set_constant("nveg", float(nveg))
    nh = run.dims[1] * run.dims[2]
    L = run.dims[3]
    run.field("dls")[:] = 1.0
    run.field("dglac")[:] = 0.0
    run.field("dsnowz")[:] = 0.0
    run.field("dwmax")[:] = 0.5
    run.field("dwatc")[:] = 0.2
    run.field("dp")[:] = 1e5
    run.field("co2")[0] = 360.0
    run.field("dfd").reshape(-1, nh)[L] = 300.0
    run.field("devap")[:] = -2e-5
    run.field("dagg")[:] = 0.5
    run.field("dgrow")[:] = 1.0
    run.field("dmr")[:] = 2.0
    run.field("dsc")[:] = 1.0
    return run, nh, L


def test_vegetation_carbon_budget_and_respiration(land):
    run, nh, L = _land_run(land)
    rng = np.random.default_rng(3)
    ts = rng.uniform(245.0, 300.0, nh)
    ts[::2] = np.where(ts[::2] > 262.0, ts[::2] - 20.0, ts[::2])  # half the columns frozen
    run.field("land_dts")[:] = ts
    cveg = rng.uniform(0.0, 8.0, nh)
    csoil = rng.uniform(0.5, 20.0, nh)
    run.field("dcveg")[:] = cveg
    run.field("dcsoil")[:] = csoil
    run.routine("land")
    dt = run.constant("deltsec")
    npp = run.field("dnpp")
    nv, ns = run.field("dcveg"), run.field("dcsoil")
    tm = run.constant("tmelt")
    frozen = ts <= tm
    assert frozen.any() and (~frozen).any()
    assert np.all(npp[frozen] == 0) and np.all(npp[~frozen] >= 0)
    # litter leaves the vegetation pool at 1/tau_veg and enters the soil
    litter = cveg / run.constant("tau_veg")
    np.testing.assert_allclose(nv - cveg, (npp - litter) * dt, rtol=1e-9, atol=1e-18)
    resp = (litter * dt - (ns - csoil)) / dt
    # Q10 = 2: respiration per unit soil carbon doubles every 10 K
    rate = resp / csoil * run.constant("tau_soil")
    np.testing.assert_allclose(rate, 2.0 ** ((ts - tm - 10.0) / 10.0), rtol=1e-9)
    # forest cover rises with vegetation carbon, from none at no carbon
    forest = run.field("dforest")
    order = np.argsort(nv)
    assert np.all(np.diff(forest[order]) >= -1e-15) and forest.min() >= 0 and forest.max() <= 1
    assert forest[np.argmin(nv)] == 0.0


def test_vegetation_feeds_back_only_with_nveg_2(land):
    outs = {}
    for nveg in (1, 2):
        run, nh, L = _land_run(land, nveg)
        run.field("land_dts")[:] = 290.0
        run.field("dcveg")[:] = np.linspace(0.0, 10.0, nh)
        run.field("dcsoil")[:] = 5.0
        run.field("dz0")[:] = 0.123
        run.routine("land")
        outs[nveg] = {n: run.field(n).copy() for n in ("dz0", "dwmax", "dalb", "dforest", "dcveg", "dnpp")}
    assert np.all(outs[1]["dz0"] == 0.123) and np.all(outs[1]["dwmax"] == 0.5)
    assert not np.all(outs[2]["dz0"] == 0.123) and not np.all(outs[2]["dwmax"] == 0.5)
    np.testing.assert_array_equal(outs[1]["dcveg"], outs[2]["dcveg"])
    np.testing.assert_array_equal(outs[1]["dnpp"], outs[2]["dnpp"])


def test_glacier_promotion_at_the_year_end(land):
    host, m = land
    for year_end in (0.0, 1.0):
        run = OracleRun(host, m)
        nh = run.dims[1] * run.dims[2]
        run.set_constant("nglacier", 1.0)
        lim = run.constant("glacelim")
        run.field("dls")[:] = 1.0
        snow = np.linspace(0.0, 2 * lim, nh)
        run.field("dsnowz")[:] = snow
        run.field("dpersist")[:] = 1.0
        run.field("dglac")[:] = 0.0
        run.field("land_dts")[:] = 260.0
        run.field("surface_step")[6] = year_end
        run.routine("land")
        glac, persist = run.field("dglac"), run.field("dpersist")
        kept = snow >= lim
        if year_end:
            np.testing.assert_array_equal(glac, np.where(kept, 1.0, 0.0))
            np.testing.assert_array_equal(persist, 1.0)  # persistence restarts on land
        else:
            np.testing.assert_array_equal(glac, 0.0)
            np.testing.assert_array_equal(persist, np.where(kept, 1.0, 0.0))
