"""Pins of the oracle's parameterizations by what the physics fixes, on
crafted columns (CPU). Each check is chosen so that a plausible slip in a
routine (a dropped term, a wrong sign or index, a wrong weight, a swapped
coefficient) fails it; planting such slips one at a time in a copy of the
oracle fails at least one test each.

* lwr: over an isothermal column on a black surface the upward flux is
  sigma T^4 at every level whatever the absorbers (Kirchhoff); a dry,
  clear, CO2- and ozone-free column is transparent; opaque clouds of
  covers c1, c2 overlap at random (downward flux (1 - (1-c1)(1-c2))
  sigma T^4 below them) and an ocean of emissivity 0.98 reflects the rest;
  the Stefan-Boltzmann constant is the physical one.
* swr: no sunlight below the horizon; the downward flux at the top is the
  insolation S0 d^-2 mu0; a column without absorbers passes the same net
  flux through every level; absorbers and cloud layers only take flux
  away; clouds raise the planetary albedo.
* solang: mu0 is the spherical-trigonometry zenith cosine on the
  Gauss-Legendre latitudes.
* radstep: the column's heating is the net flux convergence.
* mkdca: dry adjustment conserves column enthalpy and moisture, leaves a
  statically stable column, and is idempotent.
* mklsp: the member's gas constants follow the ideal-gas relations;
  condensation releases the textbook latent heat into moist air and
  brings the level to saturation by the Tetens (Murray) formula.
* mkrain: falling water is conserved between evaporation and the ground;
  evaporation cools by L_v, melting and freezing exchange L_f.
* vdiff: implicit diffusion with closed ends conserves u, v and q, is a
  backward-Euler step, mixes moisture with the heat mixing length in
  neutral air and damps heat more than momentum in stable air (Louis).
* fixer: global moisture is conserved and no humidity stays negative.
* kuo: the cloud layers' moisture supply moistens or rains out (column
  water conserved) and the rain's latent heat heats the column; the
  shallow branch only redistributes.
* soil: water (bucket, snow pack, runoff) is conserved, over snow the
  pack sublimates, and on bare land the surface and soil layers store the
  net surface heat flux through series conductances.
* mko3: Green's (1964) ozone profile at hypsometric heights.
* seastep: mixed-layer energy (with the flux handed to the ice), a
  saturated surface air and Charnock's roughness.
* surflx: the log law in neutral air, Louis's stability dependence, and
  the lowest layer's momentum, enthalpy and water budgets against the
  stress, sensible and latent heat fluxes and evaporation.

* mkradv: river routing moves water between cells without making any.
* sea_ice: thickness change = melt energy / (rho_i L_f), the snow-pack
  budget, snow melt flux = melt x L_f of snow, ice surface <= melting point.
* mkclouds: Slingo's convective cover spread by random overlap, the
  stratiform cover from relative humidity, the column total, the vapour
  path.
* land: the river store gains the runoff, the snow-albedo weight
  d / (d + 1 cm), saturated surface air.
* carbon: the Walker-Hays-Kasting weathering law against the present Earth.

The glacier and wet-soil albedo curves, the ice surface temperature's
implicit solution, the weathering supply limit, the cloud liquid-water
profile, the radiation's absorber fits and the storm's outflow ratio are
pinned in tests/test_oracle_fit_pins.py.
Each test states the tolerance it needs: FP64 arithmetic, so conservation
holds to rounding.
"""
import numpy as np
import pytest

from oracle_harness import OracleRun

STEFAN_BOLTZMANN = 5.670374419e-8  # W m^-2 K^-4, CODATA 2018


@pytest.fixture(scope="module")
def column(exo):
    host = exo.Engine(-1, 1)
    # grey radiation and no vegetation or glacier model: each pin below
    # isolates one parameterization (tests/test_oracle_surface_pins.py pins
    # the star spectrum, vegetation and glaciers)
    m = host.add_member(exo.default_config(truncation=21, levels=10, star_temperature_k=0.0, vegetation=0,
                                           glaciers=0))
    _, nlat, nlon, nlev = host.dims(m)
    yield host, m, nlat, nlon, nlev
    host.close()


def _run(column):
    host, m, nlat, nlon, nlev = column
    return OracleRun(host, m), nlat * nlon, nlev


def _planes(run, name, nh):
    return run.field(name).reshape(-1, nh)


def test_lwr_isothermal_upward_flux_is_sigma_t4(column):
    run, nh, L = _run(column)
    rng = np.random.default_rng(1)
    T0 = 271.3
    _planes(run, "dt", nh)[:] = T0
    run.field("dls")[:] = 1.0  # land: emissivity 1
    run.field("dp")[:] = rng.uniform(6e4, 1.05e5, nh)
    _planes(run, "dq", nh)[:] = rng.uniform(0, 0.02, (L + 1, nh))
    _planes(run, "dqo3", nh)[:] = rng.uniform(0, 1e-5, (L, nh))
    _planes(run, "dqco2", nh)[:] = rng.uniform(0, 1e-3, (L, nh))
    _planes(run, "dcc", nh)[:] = rng.uniform(0, 1, (L + 1, nh))
    _planes(run, "dql", nh)[:] = rng.uniform(0, 1e-4, (L + 1, nh))
    run.routine("lwr")
    sbk = run.constant("sbk")
    assert abs(sbk / STEFAN_BOLTZMANN - 1) < 1e-3
    up = _planes(run, "dftu", nh)
    np.testing.assert_allclose(up, -sbk * T0 ** 4, rtol=1e-10)
    # downward flux is the emission of the layers above: from 0 at the top
    # up to at most sigma T^4
    down = _planes(run, "dftd", nh)
    assert np.all(np.abs(down[0]) < 1e-9)
    assert np.all(down >= -1e-9) and np.all(down <= sbk * T0 ** 4 * (1 + 1e-9))


def test_lwr_clear_dry_column_is_transparent(column):
    run, nh, L = _run(column)
    T0 = 250.0
    _planes(run, "dt", nh)[:] = T0
    run.field("dls")[:] = 0.0  # ocean: emissivity 0.98
    for name in ("dq", "dcc", "dql"):
        run.field(name)[:] = 0.0
    for name in ("dqo3", "dqco2"):
        run.field(name)[:] = 0.0
    run.routine("lwr")
    sbk = run.constant("sbk")
    net = _planes(run, "dlwfl", nh)
    np.testing.assert_allclose(net, -0.98 * sbk * T0 ** 4, rtol=2e-5)


def test_lwr_black_clouds_random_overlap_over_ocean(column):
    """Opaque cloud layers of cover c1, c2 in a transparent isothermal
    column over an ocean of emissivity 0.98: below them the downward flux
    is (1 - (1 - c1)(1 - c2)) sigma T^4 (random overlap), above them zero;
    at the top the upward flux is the cloud emission plus the surface's
    emission and reflection seen through the gaps."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(7)
    T0 = 280.0
    _planes(run, "dt", nh)[:] = T0
    run.field("dls")[:] = 0.0
    run.field("dp")[:] = 1e5
    for name in ("dq", "dqo3", "dqco2", "dcc"):
        run.field(name)[:] = 0.0
    _planes(run, "dql", nh)[:] = 1e-3  # optical depth > 50 in every layer
    cols = np.arange(nh)
    k1 = rng.integers(1, L // 2, nh)
    k2 = rng.integers(L // 2, L - 1, nh)
    c1 = rng.uniform(0, 1, nh)
    c2 = rng.uniform(0, 1, nh)
    cc = _planes(run, "dcc", nh)
    cc[k1, cols] = c1
    cc[k2, cols] = c2
    run.routine("lwr")
    sbk = run.constant("sbk")
    B = sbk * T0 ** 4
    down = _planes(run, "dftd", nh)
    up = -_planes(run, "dftu", nh)
    level = np.arange(L + 1)[:, None]
    cover = np.where(level > k1, c1, 0.0)
    cover = 1 - (1 - cover) * (1 - np.where(level > k2, c2, 0.0))
    # clear layers are not exactly transparent (band transmissivities are
    # capped at 1 - 1e-6), hence 1e-5
    np.testing.assert_allclose(down, cover * B, rtol=1e-5, atol=1e-5 * B)
    eps = 0.98
    gap = (1 - c1) * (1 - c2)
    total = 1 - gap
    expect_top = (total + gap * (eps + (1 - eps) * total)) * B
    np.testing.assert_allclose(up[0], expect_top, rtol=1e-5)


def _swr_column(run, nh, L, rng, mu0):
    rs = run.field("radiation_step")
    rs[3] = 1.03  # gdist2
    run.field("gmu0")[:] = mu0
    run.field("dp")[:] = 1e5
    _planes(run, "dt", nh)[:] = 288.0 * np.append(run.table("sigma"), 1.0)[:, None] ** 0.19
    # land without ice: the direct and diffuse surface albedos agree
    run.field("dls")[:] = 1.0
    run.field("dicec")[:] = 0.0
    alb = rng.uniform(0.05, 0.6, nh)
    run.field("dalb")[:] = alb
    run.field("dsalb")[:] = np.concatenate([alb, alb])
    return run.constant("gsol0") * rs[3] * mu0


def _swr_net(run, nh):
    run.routine("swr")
    return _planes(run, "dswfl", nh).copy(), _planes(run, "dfd", nh).copy()


def test_swr_conserves_in_a_transparent_column(column):
    run, nh, L = _run(column)
    rng = np.random.default_rng(2)
    assert run.constant("nstartemp") == 0
    mu0 = rng.uniform(0.05, 1.0, nh)
    mu0[::5] = 0.0  # night columns
    top = _swr_column(run, nh, L, rng, mu0)
    for name in ("dq", "dqo3", "dcc", "dql"):
        run.field(name)[:] = 0.0
    net, down = _swr_net(run, nh)
    night = mu0 == 0
    day = ~night
    assert np.all(net[:, night] == 0) and np.all(down[:, night] == 0)
    np.testing.assert_allclose(down[0, day], top[day], rtol=1e-12)
    # no absorber anywhere: what enters at the top reaches the surface
    np.testing.assert_allclose(net[:, day], np.broadcast_to(net[0, day], net[:, day].shape), rtol=1e-12)
    alb = run.field("dalb")[day]
    assert np.all(net[0, day] < top[day] * (1 - alb) * (1 + 1e-12))  # Rayleigh adds reflection
    assert np.all(net[0, day] > 0)


def test_swr_clear_absorbing_column_decreases_downwards(column):
    run, nh, L = _run(column)
    rng = np.random.default_rng(3)
    mu0 = rng.uniform(0.05, 1.0, nh)
    top = _swr_column(run, nh, L, rng, mu0)
    run.field("dcc")[:] = 0.0
    run.field("dql")[:] = 0.0
    _planes(run, "dq", nh)[:] = rng.uniform(0, 0.015, (L + 1, nh))
    _planes(run, "dqo3", nh)[:] = rng.uniform(0, 1e-5, (L, nh))
    net, down = _swr_net(run, nh)
    np.testing.assert_allclose(down[0], top, rtol=1e-12)
    assert np.all(np.diff(net, axis=0) <= 1e-12 * top)  # every layer absorbs
    assert np.all(net[-1] > 0) and np.all(net[0] <= top)
    # more vapour absorbs more
    _planes(run, "dq", nh)[:] *= 2.0
    net2, _ = _swr_net(run, nh)
    assert np.all(net2[-1] < net[-1])


def test_swr_cloudy_column_stays_within_insolation(column):
    run, nh, L = _run(column)
    rng = np.random.default_rng(4)
    mu0 = rng.uniform(0.05, 1.0, nh)
    top = _swr_column(run, nh, L, rng, mu0)
    _planes(run, "dq", nh)[:] = rng.uniform(0, 0.015, (L + 1, nh))
    _planes(run, "dcc", nh)[:] = rng.uniform(0, 1, (L + 1, nh)) * (rng.uniform(0, 1, (L + 1, nh)) > 0.5)
    _planes(run, "dql", nh)[:] = rng.uniform(0, 2e-4, (L + 1, nh))
    net, down = _swr_net(run, nh)
    np.testing.assert_allclose(down[0], top, rtol=1e-12)
    assert np.all(net >= 0) and np.all(net <= top * (1 + 1e-12))
    # the cloud-top reflectance raises the planetary albedo: less enters
    # the column than without clouds. The cloudy fraction of a layer trades
    # vapour absorption for cloud absorption, so a thin cloud may let in up
    # to 1e-4 of the insolation more (seen: 6e-6).
    run.field("dcc")[:] = 0.0
    clear, _ = _swr_net(run, nh)
    assert np.all(net[0] <= clear[0] + 1e-4 * top)
    assert np.mean(net[0] < clear[0] - 1e-2 * top) > 0.5


def test_swr_cloud_layer_absorbs(column):
    """One cloud layer per column in a dry, ozone-free column: every layer
    of the free atmosphere (above the lowest, whose Rayleigh layer has its
    own direct and diffuse reflectances) passes on no more net flux than it
    receives."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(8)
    mu0 = rng.uniform(0.05, 1.0, nh)
    top = _swr_column(run, nh, L, rng, mu0)
    for name in ("dq", "dqo3", "dcc"):
        run.field(name)[:] = 0.0
    _planes(run, "dcc", nh)[rng.integers(0, L - 1, nh), np.arange(nh)] = rng.uniform(0.2, 1, nh)
    _planes(run, "dql", nh)[:] = rng.uniform(1e-6, 5e-4, (L + 1, nh))
    net, _ = _swr_net(run, nh)
    assert np.all(np.diff(net[:L], axis=0) <= 1e-12 * top)
    assert np.all(net[:L] >= 0)


def test_solang_is_the_zenith_cosine(column):
    run, nh, L = _run(column)
    host, m, nlat, nlon, _ = column
    delta = np.deg2rad(17.0)
    rs = run.field("radiation_step")
    rs[0], rs[1], rs[2] = np.sin(delta), np.cos(delta), 0.7
    run.routine("solang")
    phi = np.arcsin(np.polynomial.legendre.leggauss(nlat)[0][::-1])  # north to south
    lam = 2 * np.pi * np.arange(nlon) / nlon
    hour = rs[2] + lam - np.pi
    mu = (np.sin(phi)[:, None] * np.sin(delta) + np.cos(phi)[:, None] * np.cos(delta) * np.cos(hour)[None, :])
    got = run.field("gmu0").reshape(nlat, nlon)
    dawn = run.constant("zdawn")
    expect = np.where(mu > dawn, mu, 0.0)
    np.testing.assert_allclose(got, expect, atol=1e-11)


def _unstable_column(run, nh, L, rng):
    sigma = run.table("sigma")
    # a lapse rate of 12-14 K per scale height fraction, beyond the dry
    # adiabat near the ground, with random kinks
    t = 300.0 * sigma[:, None] ** 0.33 + rng.normal(0, 3.0, (L, nh))
    t[-3:] += rng.uniform(2, 8, (3, nh))
    return t


def test_mkdca_conserves_and_stabilises(column):
    run, nh, L = _run(column)
    rng = np.random.default_rng(3)
    dt = _planes(run, "dt", nh)
    dt[:L] = _unstable_column(run, nh, L, rng)
    _planes(run, "dtdt", nh)[:] = 0.0
    _planes(run, "dqdt", nh)[:] = 0.0
    dq = _planes(run, "dq", nh)
    dq[:L] = rng.uniform(1e-4, 1e-2, (L, nh))
    run.routine("mkdca")
    ds = run.table("dsigma")[:, None]
    skap = run.table("vdiff_skap")[:, None]
    dT = _planes(run, "dca_dtdt", nh)
    dQ = _planes(run, "dca_dqdt", nh)
    scale_t = np.sum(ds * np.abs(dT), axis=0) + 1e-30
    scale_q = np.sum(ds * np.abs(dQ), axis=0) + 1e-30
    assert np.all(np.abs(np.sum(ds * dT, axis=0)) <= 1e-12 * scale_t)
    assert np.all(np.abs(np.sum(ds * dQ, axis=0)) <= 1e-12 * scale_q)
    assert np.mean(np.any(dT != 0, axis=0)) > 0.5  # the columns were unstable
    limited = run.field("dca_limited") != 0
    dts = run.constant("deltsec2")
    t_new = dt[:L] + dT * dts
    theta = t_new / skap
    stable = np.all(theta[:-1] >= theta[1:] - 1e-9, axis=0)
    assert np.all(stable | limited)
    # idempotent on the adjusted column
    dt[:L] = t_new
    dq[:L] = dq[:L] + dQ * dts
    run.routine("mkdca")
    again = _planes(run, "dca_dtdt", nh)
    assert np.max(np.abs(again[:, ~limited])) < 1e-9


def test_mklsp_latent_heat_and_saturation(column):
    run, nh, L = _run(column)
    rng = np.random.default_rng(4)
    sigma = run.table("sigma")
    ps = rng.uniform(8e4, 1.03e5, nh)
    run.field("dp")[:] = ps
    t = 295.0 * sigma[:, None] ** 0.19 + rng.normal(0, 1, (L, nh))
    p = sigma[:, None] * ps[None, :]
    # the member's gas: dry-air constants from its gas constant by the
    # ideal-gas relations, vapour and latent heat from the textbook
    R, kappa = run.constant("gascon"), run.constant("akap")
    assert abs(R - 8314.46 / 28.0134) < 1e-3  # N2 (the default atmosphere)
    cp = run.constant("acpd")
    assert abs(cp - R / kappa) < 1e-9 * cp
    eps = run.constant("rdbrv")
    assert abs(eps / (R / 461.5) - 1) < 2e-3  # R_v = 461.5 J/(kg K)
    cpv = cp * (1 + run.constant("adv"))
    assert abs(cpv / 1860.0 - 1) < 1e-2  # vapour cp
    Lv = 2.501e6
    assert abs(run.constant("alv") / Lv - 1) < 1e-3
    es = 610.78 * np.exp(17.27 * (t - 273.16) / (t - 35.86))  # Tetens (Murray 1967)
    qs = eps * es / (p - (1 - eps) * es)
    ratio = rng.uniform(0.6, 1.5, (L, nh))
    q = qs * ratio
    _planes(run, "dt", nh)[:L] = t
    _planes(run, "dq", nh)[:L] = q
    _planes(run, "dtdt", nh)[:] = 0.0
    _planes(run, "dqdt", nh)[:] = 0.0
    run.routine("mklsp")
    dts = run.constant("deltsec2")
    dT = _planes(run, "dtdt", nh)[:L] * dts
    dQ = _planes(run, "dqdt", nh)[:L] * dts
    sub = ratio < 0.99
    assert np.all(dT[sub] == 0) and np.all(dQ[sub] == 0)
    sup = ratio > 1.01
    assert np.all(dQ[sup] < 0)
    # latent heat: dT = -L / (cp_moist) dq with cp_moist = cp (1 - q) + cpv q
    # to first order in q
    lcp = -dT[sup] / dQ[sup]
    expect = Lv / (cp + (cpv - cp) * q[sup])
    np.testing.assert_allclose(lcp, expect, rtol=2e-3)
    # saturated at the new temperature (two Newton steps)
    tn, qn = t[sup] + dT[sup], q[sup] + dQ[sup]
    esn = 610.78 * np.exp(17.27 * (tn - 273.16) / (tn - 35.86))
    qsn = eps * esn / (p[sup] - (1 - eps) * esn)
    np.testing.assert_allclose(qn, qsn, rtol=2e-3)
    # the condensed water falls: dprll = -dq/dt dsigma
    prl = _planes(run, "dprll", nh)
    np.testing.assert_allclose(prl[sup], -(dQ[sup] / dts) * np.broadcast_to(run.table("dsigma")[:, None],
                                                                               (L, nh))[sup], rtol=1e-12)


def test_vdiff_conserves_column_momentum_and_moisture(column):
    run, nh, L = _run(column)
    rng = np.random.default_rng(5)
    sigma = run.table("sigma")
    _planes(run, "dt", nh)[:L] = 290.0 * sigma[:, None] ** 0.2 + rng.normal(0, 2, (L, nh))
    _planes(run, "du", nh)[:L] = rng.normal(5, 10, (L, nh))
    _planes(run, "dv", nh)[:L] = rng.normal(0, 8, (L, nh))
    _planes(run, "dq", nh)[:L] = rng.uniform(1e-5, 1e-2, (L, nh))
    for name in ("dudt", "dvdt", "dqdt", "dtdt"):
        run.field(name)[:] = 0.0
    run.routine("vdiff")
    ds = run.table("dsigma")[:, None]
    moved = 0
    for name in ("dudt", "dvdt", "dqdt"):
        x = _planes(run, name, nh)[:L]
        scale = np.sum(ds * np.abs(x), axis=0) + 1e-30
        assert np.all(np.abs(np.sum(ds * x, axis=0)) <= 1e-10 * scale), name
        moved += np.count_nonzero(x)
    assert moved > 0


def _vdiff_coefficients(x, x_new, ds):
    """Exchange coefficients K_l (interface l between layers l and l+1) of
    a backward-Euler diffusion step ds_l (x'_l - x_l) = K_l (x'_{l+1} - x'_l)
    - K_{l-1} (x'_l - x'_{l-1}) with closed ends, recovered from its input
    and output: the flux through interface l is the cumulative change above
    it. Also returns the flux through the bottom (zero for a closed end)."""
    flux = np.cumsum(ds * (x_new - x), axis=0)
    return flux[:-1] / (x_new[1:] - x_new[:-1]), flux[-1]


def _vdiff_run(run, nh, L, rng, theta_v):
    skap = run.table("vdiff_skap")[:, None]
    f = np.sort(rng.uniform(-1, 1, (L, nh)), axis=0)  # strictly monotone: shear everywhere
    u = 8.0 + 6.0 * f
    q = 2e-3 * (1 + 0.5 * f)
    moist = 1 / run.constant("rdbrv") - 1
    _planes(run, "dt", nh)[:L] = theta_v * skap / (1 + moist * q)
    _planes(run, "du", nh)[:L] = u
    _planes(run, "dv", nh)[:L] = 0.0
    _planes(run, "dq", nh)[:L] = q
    for name in ("dudt", "dvdt", "dqdt", "dtdt"):
        run.field(name)[:] = 0.0
    run.routine("vdiff")
    dts = run.constant("deltsec2")
    ds = run.table("dsigma")[:, None]
    km, bottom_u = _vdiff_coefficients(u, u + _planes(run, "dudt", nh)[:L] * dts, ds)
    kh, bottom_q = _vdiff_coefficients(q, q + _planes(run, "dqdt", nh)[:L] * dts, ds)
    assert np.all(np.abs(bottom_u) < 1e-9 * np.sum(ds * np.abs(u), axis=0))
    assert np.all(np.abs(bottom_q) < 1e-9 * np.sum(ds * np.abs(q), axis=0))
    assert np.all(km > 0) and np.all(kh > 0)
    ratio = run.table("vdiff_mixh2")[:L - 1, None] / run.table("vdiff_mixm2")[:L - 1, None]
    return kh / km, ratio


def test_vdiff_neutral_column_mixes_with_the_mixing_lengths(column):
    """Uniform virtual potential temperature (Richardson number 0): both
    stability functions are 1, so moisture's exchange coefficient is
    momentum's times the ratio of the squared heat and momentum mixing
    lengths, interface by interface."""
    run, nh, L = _run(column)
    got, ratio = _vdiff_run(run, nh, L, np.random.default_rng(9), 290.0)
    np.testing.assert_allclose(got, np.broadcast_to(ratio, got.shape), rtol=1e-8)


def test_vdiff_stable_air_damps_heat_more_than_momentum(column):
    """Louis (1979): in stable air (Richardson number > 0) the heat
    stability function falls faster than the momentum one, so moisture's
    coefficient drops below its neutral share of momentum's. (In unstable
    air the two enhancements depend on the mixing lengths, which differ
    here, so no order is fixed there.)"""
    run, nh, L = _run(column)
    height = np.arange(L - 1, -1, -1)[:, None]  # layer 0 is the top
    got, ratio = _vdiff_run(run, nh, L, np.random.default_rng(10), 290.0 + 2.0 * height)
    assert np.all(got < ratio * (1 - 1e-6))


def test_fixer_conserves_global_moisture(column):
    run, nh, L = _run(column)
    host, m, nlat, nlon, _ = column
    rng = np.random.default_rng(6)
    q = rng.uniform(0, 5e-3, (L, nh))
    neg = rng.uniform(0, 1, (L, nh)) < 0.08
    q[neg] = -rng.uniform(0, 2e-4, neg.sum())
    _planes(run, "dq", nh)[:L] = q
    run.field("gqdt")[:] = 0.0
    ps = rng.uniform(7e4, 1.03e5, nh)
    run.field("dp")[:] = ps
    ds = run.table("dsigma")[:, None]
    w = run.table("zgw")
    before = np.sum(w * ps * np.sum(ds * q, axis=0))
    run.routine("fixer")
    after_q = _planes(run, "dq", nh)[:L]
    after = np.sum(w * ps * np.sum(ds * after_q, axis=0))
    assert np.all(after_q >= 0)
    assert abs(after / before - 1) < 1e-12


def test_radstep_heating_is_the_flux_convergence(column):
    """Radiative heating of the column's moist air (cp of dry air times
    1 + (cpv/cp - 1) q) equals the net flux entering at the top minus the
    net flux leaving at the surface, and the short- and longwave parts
    add up to it."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(12)
    ps = rng.uniform(6e4, 1.05e5, nh)
    run.field("dp")[:] = ps
    _planes(run, "dt", nh)[:] = rng.uniform(200, 300, (L + 1, nh))
    q = rng.uniform(0, 0.02, (L + 1, nh))
    _planes(run, "dq", nh)[:] = q
    _planes(run, "dswfl", nh)[:] = rng.uniform(0, 800, (L + 1, nh))
    _planes(run, "dlwfl", nh)[:] = rng.uniform(-400, 0, (L + 1, nh))
    run.field("dtdt")[:] = 0.0
    run.routine("radstep")
    flux = _planes(run, "dswfl", nh) + _planes(run, "dlwfl", nh)
    dtdt = _planes(run, "dtdt", nh)[:L]
    ga, cp = run.constant("ga"), run.constant("acpd")
    cpv = cp * (1 + run.constant("adv"))
    mass = run.table("dsigma")[:, None] * ps / ga  # kg/m^2 per layer
    heat = np.sum(mass * (cp * (1 - q[:L]) + cpv * q[:L]) * dtdt, axis=0)
    # cp (1 + adv q) is the first-order form of cp (1 - q) + cpv q
    scale = np.abs(flux).max(axis=0)
    assert np.all(np.abs(heat - (flux[0] - flux[L])) <= 0.03 * scale)
    exact = np.sum(mass * cp * (1 + run.constant("adv") * q[:L]) * dtdt, axis=0)
    assert np.all(np.abs(exact - (flux[0] - flux[L])) <= 1e-12 * scale)
    parts = _planes(run, "dtdtswr", nh)[:L] + _planes(run, "dtdtlwr", nh)[:L]
    np.testing.assert_allclose(parts, dtdt, rtol=1e-12, atol=1e-18)


def _saturation(run, t, p):
    eps = run.constant("rdbrv")
    es = 610.78 * np.exp(17.2693882 * (t - 273.16) / (t - 35.86))  # Tetens (Murray 1967)
    return eps * es / (p - (1 - eps) * es)


def _rain_column(run, nh, L, rng, t, humidity):
    ps = rng.uniform(8e4, 1.03e5, nh)
    run.field("dp")[:] = ps
    p = run.table("sigma")[:, None] * ps
    _planes(run, "dt", nh)[:] = t
    q = humidity * _saturation(run, t[:L], p)
    _planes(run, "dq", nh)[:L] = q
    for name in ("dtdt", "dqdt", "dprcl", "dprll", "dprscl", "dprsll"):
        run.field(name)[:] = 0.0
    return ps, q


def _rain_budget(run, nh, L, ps):
    """Column water in (fluxes made per level) against water out (surface
    precipitation and evaporation), in the flux units sigma s^-1."""
    w = run.constant("ga") * 1000.0 / ps
    made = np.sum(_planes(run, "dprcl", nh)[:L] + _planes(run, "dprll", nh)[:L], axis=0)
    evap = np.sum(run.table("dsigma")[:, None] * _planes(run, "dqdt", nh)[:L], axis=0)
    surface = (run.field("dprc") + run.field("dprl")) * w
    return made, evap, surface, run.field("dprs") * w


def test_mkrain_evaporation_conserves_water_and_cools(column):
    """Rain falling through warm subsaturated air: what evaporates plus
    what reaches the ground is what fell, the air is cooled by L_v per kg
    evaporated, and no level is pushed past saturation."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(13)
    t = 280.0 + 20.0 * np.append(run.table("sigma"), 1.0)[:, None]  # above freezing throughout
    ps, q = _rain_column(run, nh, L, rng, np.broadcast_to(t, (L + 1, nh)).copy(), rng.uniform(0.3, 0.95, (L, nh)))
    ds = run.table("dsigma")[:, None]
    _planes(run, "dprll", nh)[:L] = rng.uniform(0, 1e-7, (L, nh)) * ds
    _planes(run, "dprcl", nh)[:L] = rng.uniform(0, 1e-7, (L, nh)) * ds
    run.routine("mkrain")
    made, evap, surface, snow = _rain_budget(run, nh, L, ps)
    np.testing.assert_allclose(evap + surface, made, rtol=1e-12)
    assert np.all(evap > 0) and np.all(surface >= 0) and np.all(snow == 0)
    dq = _planes(run, "dqdt", nh)[:L]
    dT = _planes(run, "dtdt", nh)[:L]
    cp, adv = run.constant("acpd"), run.constant("adv")
    assert abs(run.constant("alv") / 2.501e6 - 1) < 1e-3
    wet = dq > 0
    np.testing.assert_allclose(-dT[wet] / dq[wet], (run.constant("alv") / (cp * (1 + adv * q)))[wet], rtol=1e-12)
    dts = run.constant("deltsec2")
    p = run.table("sigma")[:, None] * ps
    after = _saturation(run, t[:L] + dT * dts, p)
    assert np.all(q + dq * dts <= after * (1 + 2e-3))


def test_mkrain_melting_and_freezing_exchange_latent_heat_of_fusion(column):
    """Saturated air (nothing evaporates): snow made at the top falls into
    air above freezing and melts there, rain made at the top falls into
    air below freezing and freezes; either way the column's enthalpy
    changes by L_f per kg, L_f = L_s - L_v the textbook 3.34e5 J/kg, and
    all the water reaches the ground in the new phase."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(14)
    ds = run.table("dsigma")[:, None]
    cp, adv = run.constant("acpd"), run.constant("adv")
    lf = run.constant("als") - run.constant("alv")
    assert abs(lf / 3.34e5 - 1) < 5e-3
    ga = run.constant("ga")
    for warm in (True, False):
        t = np.full((L + 1, nh), 283.0 if warm else 255.0)
        ps, q = _rain_column(run, nh, L, rng, t, 1.05)
        falling = rng.uniform(1e-9, 1e-7, nh)
        _planes(run, "dprll", nh)[0] = falling
        if warm:
            _planes(run, "dprsll", nh)[0] = falling  # all of it snow
        run.routine("mkrain")
        made, evap, surface, snow = _rain_budget(run, nh, L, ps)
        assert np.all(evap == 0)
        np.testing.assert_allclose(surface, made, rtol=1e-12)
        np.testing.assert_allclose(snow, 0.0 if warm else made, rtol=1e-12, atol=0)
        dT = _planes(run, "dtdt", nh)[:L]
        heat = np.sum(ds * cp * (1 + adv * q) * dT, axis=0)
        np.testing.assert_allclose(heat, (-lf if warm else lf) * falling, rtol=1e-12)


def test_kuo_consumes_the_moisture_supply(column):
    """Kuo convection: in a convecting column the moisture supply of the
    cloud layers (convergence plus earlier tendencies) either moistens the
    cloud or rains out, so the column's total moisture tendency plus the
    convective rain is unchanged, and the rain's latent heat L_v is the
    heating of the column's moist air. Where the supply is not positive the
    shallow branch only redistributes heat and moisture (both column sums
    stay). Crafted tropical soundings with convergence in the lowest four
    layers; both branches occur."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(15)
    sig = run.table("sigma")
    ps = rng.uniform(9.5e4, 1.03e5, nh)
    run.field("dp")[:] = ps
    tsurf = rng.uniform(290, 305, nh)
    t = tsurf * sig[:, None] ** rng.uniform(0.12, 0.2, nh)
    _planes(run, "dt", nh)[:L] = t
    _planes(run, "dt", nh)[L] = tsurf + rng.uniform(0, 3, nh)
    q = _saturation(run, t, sig[:, None] * ps) * rng.uniform(0.5, 0.98, (L, nh))
    _planes(run, "dq", nh)[:L] = q
    _planes(run, "dq", nh)[L] = q[-1] * 1.05
    run.field("dqt")[:] = 0.0
    to_model = ps / (run.constant("psurf") * run.constant("ww"))
    _planes(run, "dqt", nh)[L - 4:L] = rng.uniform(-0.2, 1, (4, nh)) * 1e-7 * to_model
    dqdt0 = rng.uniform(-2e-9, 2e-9, (L, nh))
    dtdt0 = rng.uniform(-1e-5, 1e-5, (L, nh))
    _planes(run, "dqdt", nh)[:L] = dqdt0
    _planes(run, "dtdt", nh)[:L] = dtdt0
    run.routine("kuo")
    ds = run.table("dsigma")[:, None]
    dqt = _planes(run, "dqt", nh)[:L]  # in physical units now
    dqdt = _planes(run, "dqdt", nh)[:L]
    dtdt = _planes(run, "dtdt", nh)[:L]
    rain = np.sum(_planes(run, "dprcl", nh)[:L], axis=0)
    assert np.all(_planes(run, "dprscl", nh)[:L] == 0) and np.all(rain >= 0)
    before = np.sum(ds * (dqt + dqdt0), axis=0)
    after = np.sum(ds * (dqt + dqdt), axis=0) + rain
    scale = np.sum(ds * (np.abs(dqt) + np.abs(dqdt0)), axis=0)
    assert np.all(np.abs(after - before) <= 1e-12 * scale)
    deep = rain > 0
    shallow = ~deep & np.any(dtdt != dtdt0, axis=0)
    assert deep.sum() > 200 and shallow.sum() > 0
    cp, adv, lv = run.constant("acpd"), run.constant("adv"), run.constant("alv")
    heat = np.sum(ds * cp * (1 + adv * q) * (dtdt - dtdt0), axis=0)
    np.testing.assert_allclose(heat[deep], lv * rain[deep], rtol=1e-10)
    dT = np.sum(ds * (dtdt - dtdt0), axis=0)
    dQ = np.sum(ds * (dqdt - dqdt0), axis=0)
    # the diffusion steps T and q themselves: rounding of T (300 K) and q
    # over one time step
    dts = run.constant("deltsec2")
    assert np.all(np.abs(dT[shallow]) <= 1e-14 * 300.0 / dts)
    assert np.all(np.abs(dQ[shallow]) <= 1e-14 * 0.03 / dts)


def _land_column(run, nh, rng):
    run.field("dls")[:] = 1.0
    run.field("dglac")[:] = 0.0
    run.field("dwmax")[:] = rng.uniform(0.1, 0.5, nh)
    run.field("dwatc")[:] = run.field("dwmax") * rng.uniform(0.2, 1.0, nh)
    run.field("land_dtsm")[:] = rng.uniform(265, 300, nh)
    _planes(run, "dsoilt", nh)[:5] = rng.uniform(265, 300, (5, nh))
    run.field("dshfl")[:] = rng.uniform(-80, 40, nh)
    run.field("dlhfl")[:] = rng.uniform(-150, 0, nh)
    run.field("dflux")[:] = 0.0
    _planes(run, "dflux", nh)[-1] = rng.uniform(-50, 400, nh)


def test_soil_conserves_water_and_heat(column):
    """The land column: rain, snowfall and evaporation go into the bucket,
    the snow pack and runoff with nothing lost; on snow-free land the
    surface layer and the five soil layers (backward-Euler heat equation,
    insulated bottom) store exactly the net surface heat flux."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(16)
    _land_column(run, nh, rng)
    snow = rng.uniform(0, 0.3, nh) * (rng.uniform(0, 1, nh) < 0.5)
    run.field("dsnowz")[:] = snow
    dts_s = run.constant("deltsec")
    rain = rng.uniform(0, 3e-7, nh)
    snowfall = rng.uniform(0, 1e-7, nh) * (rng.uniform(0, 1, nh) < 0.5)
    run.field("dprl")[:] = rain + snowfall  # totals include their snow
    run.field("dprc")[:] = rng.uniform(0, 2e-7, nh)
    run.field("dprs")[:] = snowfall
    run.field("devap")[:] = -rng.uniform(0, 5e-8, nh)
    before_soil = _planes(run, "dsoilt", nh)[:5].copy()
    before_t = run.field("land_dtsm").copy()
    before_w = run.field("dwatc").copy()
    water_in = (run.field("devap") + run.field("dprl") + run.field("dprc")) * dts_s
    flux = (run.field("dshfl") + run.field("dlhfl") + _planes(run, "dflux", nh)[-1]).copy()
    assert np.all(before_w + water_in >= 0)  # the bucket does not run dry
    run.routine("soil")
    stored = (run.field("dwatc") - before_w) + run.field("drunoff") * dts_s + (run.field("dsnowz") - snow)
    np.testing.assert_allclose(stored, water_in, rtol=0, atol=1e-15)
    assert np.all(run.field("dwatc") <= run.field("dwmax"))
    assert np.any(run.field("drunoff") > 0)
    # over snow, evaporation is sublimation from the pack: where nothing
    # melts the pack gains snowfall and loses sublimation, and the bucket
    # only rain
    new_snow = run.field("dsnowz")
    kept = (snow > 0) & (run.field("dsmelt") == 0) & (new_snow > 0)
    assert kept.sum() > 50
    np.testing.assert_allclose((new_snow - snow)[kept],
                               ((snowfall + run.field("devap")) * dts_s)[kept], rtol=1e-9, atol=1e-16)
    bare = (snow == 0) & (snowfall == 0)
    assert bare.sum() > 100
    cap = run.constant("soilcap")
    soilz = run.table("dsoilz")[:, None]
    heat = cap * (run.constant("dztop") * (run.field("land_dts") - before_t)
                  + np.sum(soilz * (_planes(run, "dsoilt", nh)[:5] - before_soil), axis=0))
    # rounding of the temperatures (300 K) times the column's heat capacity
    atol = 1e-13 * cap * (run.constant("dztop") + soilz.sum()) * 300.0
    np.testing.assert_allclose(heat[bare], flux[bare] * dts_s, rtol=1e-9, atol=atol)
    # the heat flux between layers l and l+1 (the change stored below the
    # interface) is the conductivity over the distance between the layer
    # centres times the new temperature difference (backward Euler)
    new = _planes(run, "dsoilt", nh)[:5]
    below = np.cumsum((cap * soilz * (new - before_soil) / dts_s)[::-1], axis=0)[::-1][1:]
    gap = (new[:-1] - new[1:])[:, bare]
    conductance = below[:, bare] / gap
    expect = np.broadcast_to(2 * run.constant("soildiff") / (soilz[:-1] + soilz[1:]), gap.shape)
    clear = np.abs(gap) > 0.5  # well-conditioned quotients (K)
    np.testing.assert_allclose(conductance[clear], expect[clear], rtol=1e-8)


def test_mko3_is_the_green_profile(column):
    """Ozone from Green's (1964) profile: the amount above height z is
    a (1 + e^(-b/c)) / (1 + e^((z - b)/c)) with the column total a, at
    hypsometric heights z from the layers' temperatures; mixing ratios
    convert with the density of ozone at STP (2.14 kg/m^3, amounts in cm)."""
    run, nh, L = _run(column)
    host, m, nlat, nlon, _ = column
    rng = np.random.default_rng(17)
    ps = rng.uniform(6e4, 1.05e5, nh)
    run.field("dp")[:] = ps
    t = rng.uniform(190, 300, (L, nh))
    _planes(run, "dt", nh)[:L] = t
    season = 0.6
    run.field("radiation_step")[4] = season
    run.routine("mko3")
    ga, R = run.constant("ga"), run.constant("gascon")
    b, c = run.constant("bo3"), run.constant("co3")
    mu = np.polynomial.legendre.leggauss(nlat)[0][::-1]  # sin(latitude), north to south
    sinlat = np.repeat(mu, nlon)
    total = run.constant("a0o3") + run.constant("a1o3") * np.abs(sinlat) + run.constant("aco3") * sinlat * season
    sh = run.table("sigmah")
    assert sh[L - 1] == 1.0
    # height of each layer's top interface above the surface
    thickness = R * t / ga * np.log(sh[:, None] / np.append(np.nan, sh[:-1])[:, None])
    ztop = np.cumsum(thickness[::-1], axis=0)[::-1]
    above = lambda z: total * (1 + np.exp(-b / c)) / (1 + np.exp((z - b) / c))
    amount = _planes(run, "dqo3", nh)[:L] * (100.0 / 2.14) * run.table("dsigma")[:, None] * ps / ga
    zbottom = np.vstack([ztop[1:], np.zeros((1, nh))])
    np.testing.assert_allclose(amount[1:], above(zbottom[1:]) - above(ztop[1:]), rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(amount[0], above(zbottom[0]), rtol=1e-10)
    np.testing.assert_allclose(amount.sum(axis=0), total, rtol=1e-12)


def test_seastep_mixed_layer_energy_and_surface_air(column):
    """Sea points: the mixed layer (sea water, rho c_p of 1030 kg/m^3 times
    4180 J/(kg K)) warms by the heat it gets, and what freezing takes from
    it is handed to the ice as flux, so rho c_p h dSST/dt plus the flux to
    the ice is the net heat in. The surface air is saturated at the sea's
    temperature (Tetens), and the open-water roughness is Charnock's
    z0 = alpha u*^2 / g with u*^2 = |tau| / rho_air."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(18)
    assert run.constant("ice_coupled") == 1
    assert abs(run.constant("crhos") * run.constant("cps") / (1030 * 4180) - 1) < 1e-12
    run.field("dls")[:] = 0.0
    run.field("xls")[:] = 0.0
    run.field("yls")[:] = 0.0
    run.field("ice_sums")[:] = 0.0
    run.field("xsnow")[:] = 0.0
    heat = rng.uniform(-300, 300, nh)
    run.field("xcflux")[:] = heat
    fldo = rng.uniform(-30, 30, nh)
    run.field("yfldo")[:] = fldo
    tf = run.constant("tfreeze")
    sst = rng.uniform(tf - 1.5, tf + 20, nh)
    run.field("ysst")[:] = sst
    ice = rng.uniform(0, 1, nh) < 0.3
    run.field("xiced")[:] = np.where(ice, rng.uniform(0.1, 2, nh), 0.0)
    run.field("xicec")[:] = np.where(ice, 1.0, 0.0)
    mld = rng.uniform(20, 100, nh)
    run.field("ymld")[:] = mld
    ts = rng.uniform(265, 305, nh)
    run.field("xts")[:] = ts
    ps = rng.uniform(9e4, 1.04e5, nh)
    run.field("dp")[:] = ps
    taux, tauy = rng.normal(0, 0.3, nh), rng.normal(0, 0.3, nh)
    run.field("dtaux")[:] = taux
    run.field("dtauy")[:] = tauy
    run.routine("seastep")
    dtmix = run.constant("dtmix")
    stored = 1030 * 4180 * mld * (run.field("ysst") - sst) / dtmix
    to_ice = run.field("xoheat")
    # rounding of the SST (300 K) times rho c_p h / dt (up to 1.6e5 W/m^2/K)
    np.testing.assert_allclose(stored + to_ice, heat + fldo, rtol=1e-9, atol=1e-7)
    assert np.all(run.field("ysst")[ice] == tf)
    assert np.all(to_ice[~ice & (sst > tf + 1)] == 0)
    eps = run.constant("rdbrv")
    np.testing.assert_allclose(_planes(run, "dq", nh)[L], _saturation(run, ts, ps), rtol=2e-3)
    np.testing.assert_allclose(_planes(run, "dt", nh)[L], ts, rtol=0)
    rho = ps / (run.constant("gascon") * ts)
    z0 = np.maximum(0.018 * np.hypot(taux, tauy) / rho / run.constant("ga"), run.constant("dz0sea"))
    np.testing.assert_allclose(run.field("dz0")[~ice], z0[~ice], rtol=1e-12)


def _surface_layer(run, nh, L, rng, ts):
    ps = rng.uniform(8e4, 1.04e5, nh)
    run.field("dp")[:] = ps
    u, v = rng.normal(0, 8, nh), rng.normal(0, 8, nh)
    t = rng.uniform(260, 300, nh)
    q = rng.uniform(1e-3, 1.5e-2, nh)
    for name, x in (("du", u), ("dv", v), ("dt", t), ("dq", q)):
        _planes(run, name, nh)[L - 1] = x
    for name in ("dudt", "dvdt", "dtdt", "dqdt"):
        run.field(name)[:] = 0.0
    _planes(run, "dt", nh)[L] = ts(t)
    _planes(run, "dq", nh)[L] = rng.uniform(1e-3, 2e-2, nh)
    run.field("dz0")[:] = 10 ** rng.uniform(-4, 0, nh)
    run.field("drhs")[:] = 1.0
    dry = rng.uniform(0, 1, nh) < 0.3  # nearly empty buckets
    run.field("dwatc")[:] = np.where(dry, rng.uniform(0, 2e-4, nh), rng.uniform(0, 0.2, nh))
    return ps, u, v, t, q


def test_surflx_log_law_and_layer_budgets(column):
    """Bulk surface fluxes into the lowest layer:
    - in neutral air (equal virtual potential temperature of air and
      surface) the transfer coefficient is the log law (kappa /
      ln(1 + z/z0))^2 |u|, kappa = 0.4, z the hypsometric height of the
      lowest level; stable air transfers less, unstable air more;
    - the stress is rho_s C |u| u on the new wind (backward Euler), and it
      is exactly the momentum the layer loses; u*^3 = (|tau| / rho_s)^1.5;
    - the sensible heat flux is what the layer's moist air loses, beside
      the kinetic energy the stress dissipates into heat;
    - the evaporation is the water the layer gains, its latent heat the
      textbook L_v (L_s over cold land), and land never evaporates more
      than its bucket holds."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(19)
    sig = run.table("sigma")[L - 1]
    kappa_air = run.constant("akap")
    assert run.table("flux_lnsig")[0] == pytest.approx(np.log(sig), rel=1e-12)
    assert run.table("flux_sigfac")[0] == pytest.approx(sig ** -kappa_air, rel=1e-12)
    moist = 1 / run.constant("rdbrv") - 1
    R, ga = run.constant("gascon"), run.constant("ga")
    for land in (1.0, 0.0):
        run.field("dls")[:] = land
        # neutral: the surface's virtual temperature equals the air's
        # potential virtual temperature at the surface
        ps, u, v, t, q = _surface_layer(run, nh, L, rng, lambda t: t)  # surface set below
        qs = _planes(run, "dq", nh)[L].copy()
        theta_v = t * sig ** -kappa_air * (1 + moist * q)
        ts = theta_v / (1 + moist * qs)
        _planes(run, "dt", nh)[L] = ts
        run.routine("surflx")
        z = R * 0.5 * (t + theta_v) * np.log(1 / sig) / ga
        speed = np.sqrt(np.maximum(run.constant("zumin"), u * u + v * v))
        log_law = speed * (0.4 / np.log(1 + z / run.field("dz0"))) ** 2
        # the Richardson number is zero up to rounding (~1e-16), which the
        # stability functions' square and cube roots lift to ~1e-7
        np.testing.assert_allclose(run.field("dtransm"), log_law, rtol=1e-6)
        np.testing.assert_allclose(run.field("dtransh"), log_law, rtol=1e-6)
        # stable (air 3 K warmer than neutral) and unstable (3 K colder)
        for shift, cmp in ((3.0, np.less), (-3.0, np.greater)):
            _planes(run, "du", nh)[L - 1], _planes(run, "dv", nh)[L - 1] = u, v
            _planes(run, "dt", nh)[L - 1] = t
            _planes(run, "dq", nh)[L - 1] = q
            _planes(run, "dt", nh)[L] = ts - shift
            for name in ("dudt", "dvdt", "dtdt", "dqdt"):
                run.field(name)[:] = 0.0
            run.routine("surflx")
            assert np.all(cmp(run.field("dtransm"), log_law)), (land, shift)
            assert np.all(cmp(run.field("dtransh"), log_law)), (land, shift)
        # very stable air (bulk Richardson number above 1): Louis's heat
        # transfer falls to well below half of the momentum transfer
        _planes(run, "dt", nh)[L] = ts - 15.0
        run.routine("surflx")
        theta_s = (ts - 15.0) * (1 + moist * qs)
        ri = ga * z * (theta_v - theta_s) / (speed ** 2 * theta_v)
        strong = ri > 1
        assert strong.sum() > 100
        assert np.all(run.field("dtransh")[strong] < 0.5 * run.field("dtransm")[strong])
    # budgets on random (non-neutral) layers over land and sea
    run.field("dls")[:] = (rng.uniform(0, 1, nh) < 0.5).astype(float)
    ps, u, v, t, q = _surface_layer(run, nh, L, rng, lambda t: t + rng.uniform(-8, 8, t.shape))
    ts = _planes(run, "dt", nh)[L].copy()
    run.routine("surflx")
    dts = run.constant("deltsec2")
    mass = run.table("dsigma")[L - 1] * ps / ga
    un = u + _planes(run, "dudt", nh)[L - 1] * dts
    vn = v + _planes(run, "dvdt", nh)[L - 1] * dts
    rho_s = ps / (R * ts)
    trans_m = run.field("dtransm")
    np.testing.assert_allclose(run.field("dtaux"), rho_s * trans_m * un, rtol=1e-10)
    np.testing.assert_allclose(run.field("dtauy"), rho_s * trans_m * vn, rtol=1e-10)
    np.testing.assert_allclose(mass * (un - u) / dts, -run.field("dtaux"), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(mass * (vn - v) / dts, -run.field("dtauy"), rtol=1e-9, atol=1e-12)
    tau = np.hypot(run.field("dtaux"), run.field("dtauy"))
    np.testing.assert_allclose(run.field("dust3"), (tau / rho_s) ** 1.5, rtol=1e-10)
    cp, adv = run.constant("acpd"), run.constant("adv")
    heating = mass * cp * (1 + adv * q) * _planes(run, "dtdt", nh)[L - 1]
    dissipated = -mass * 0.5 * (un ** 2 - u ** 2 + vn ** 2 - v ** 2) / dts
    assert run.constant("ndheat") != 0
    np.testing.assert_allclose(heating, -run.field("dshfl") + dissipated, rtol=1e-8, atol=1e-6)
    gained = mass * _planes(run, "dqdt", nh)[L - 1]
    np.testing.assert_allclose(gained, -run.field("devap") * 1000.0, rtol=1e-10, atol=1e-16)
    latent = np.where((ts > run.constant("tmelt")) | (run.field("dls") < 0.5), 2.501e6, 2.834e6)
    np.testing.assert_allclose(run.field("dlhfl"), run.field("devap") * 1000.0 * latent, rtol=2e-3)
    on_land = run.field("dls") > 0
    bucket = run.field("dwatc") / run.constant("deltsec")
    assert np.all(-run.field("devap")[on_land] <= bucket[on_land] * (1 + 1e-12))
    assert np.any(-run.field("devap")[on_land] >= bucket[on_land] * (1 - 1e-12))  # the limit binds


def test_mkradv_moves_river_water_without_making_any(column):
    """River routing (mkradv): water crosses each cell face out of one cell
    and into its neighbour, weighted by the cell areas, so the area-weighted
    sum of the runoff is unchanged except for what leaves through the
    southern edge of the last latitude row."""
    run, nh, L = _run(column)
    host, m, nlat, nlon, _ = column
    rng = np.random.default_rng(20)
    run.field("dls")[:] = (rng.uniform(0, 1, nh) < 0.6).astype(float)
    run.field("driver")[:] = rng.uniform(0, 50, nh)
    area = rng.uniform(1e9, 5e10, nh)
    run.field("darea")[:] = area
    run.field("duroff")[:] = rng.uniform(-1e-6, 1e-6, nh)
    run.field("dvroff")[:] = rng.uniform(-1e-6, 1e-6, nh)
    before = rng.uniform(0, 1e-6, nh)
    run.field("drunoff")[:] = before
    run.routine("mkradv")
    after = run.field("drunoff")
    river = run.field("driver") * area * run.field("dls")
    last = slice((nlat - 1) * nlon, nh)
    v_last = run.field("dvroff")[last]
    lost = np.sum(np.where(v_last < 0, -v_last * river[last], 0.0))
    change = np.sum(area * (after - before))
    scale = np.sum(area * np.abs(after - before))
    assert abs(change + lost) <= 1e-12 * scale
    assert np.count_nonzero(after != before) > nh // 2


def test_sea_ice_mass_and_latent_heat_budgets(column):
    """Sea ice: the thickness change is the melt energy over rho_i L_f (the
    textbook 920 kg/m^3 and 3.28e5 J/kg of sea ice); the snow pack changes
    by snowfall minus melt; the snow melt flux is the melt times the
    latent heat of fusion of snow (3.337e5 J/kg); an ice surface never
    exceeds the melting point, open water is at the sea temperature."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(21)
    rho_i, lf_i, lf_s = run.constant("crhoi"), run.constant("clfi"), run.constant("clfsn")
    assert abs(rho_i / 920 - 1) < 0.02 and abs(lf_i / 3.34e5 - 1) < 0.03 and abs(lf_s / 3.337e5 - 1) < 1e-3
    xdt = run.constant("xdt")
    run.field("dls")[:] = 0.0
    run.field("xls")[:] = 0.0
    run.field("sea_naccua")[:] = 0.0
    run.field("sea_cheata")[:] = 0.0
    run.field("sea_cprsa")[:] = 0.0
    run.field("dshfl")[:] = rng.uniform(-60, 20, nh)
    run.field("dlhfl")[:] = rng.uniform(-80, 0, nh)
    _planes(run, "dswfl", nh)[L] = rng.uniform(0, 300, nh)
    _planes(run, "dlwfl", nh)[L] = rng.uniform(-120, -20, nh)
    snowfall = rng.uniform(0, 5e-8, nh) * (rng.uniform(0, 1, nh) < 0.6)
    run.field("dprs")[:] = snowfall
    tf = run.constant("tfreeze")
    run.field("xsst")[:] = rng.uniform(tf - 0.5, tf + 3, nh)
    ice = rng.uniform(0, 4, nh) * (rng.uniform(0, 1, nh) < 0.7)
    run.field("xiced")[:] = ice
    run.field("xicec")[:] = np.where(ice > 0, rng.uniform(0, 1, nh), 0.0)
    run.field("xicecc")[:] = run.field("xicec")
    snow = rng.uniform(0, 0.3, nh) * (ice > 0)
    run.field("xsnow")[:] = snow
    run.field("xts")[:] = rng.uniform(240, 275, nh)
    run.field("xcflux")[:] = 0.0
    run.field("xcfluxf")[:] = 0.0
    run.field("xoflux")[:] = rng.uniform(-20, 20, nh)
    run.routine("sea_ice")
    new_ice = run.field("xiced")
    melt = run.field("ximelt")
    np.testing.assert_allclose(new_ice - ice, -melt * xdt / (rho_i * lf_i), rtol=0, atol=1e-8)
    assert np.any(melt > 0) and np.any(melt < 0)  # both melting and freezing
    smelt = run.field("xsmelt")
    np.testing.assert_allclose(run.field("xsnow") - snow, (snowfall - smelt) * xdt, rtol=0, atol=1e-12)
    np.testing.assert_allclose(run.field("xsmflx"), smelt * 1000.0 * lf_s, rtol=1e-12, atol=1e-9)
    covered = ice >= run.constant("xmind")
    assert np.all(run.field("xts")[covered] <= run.constant("tmelt"))
    np.testing.assert_array_equal(run.field("xts")[~covered], run.field("xsst")[~covered])  # open water
    assert np.all(new_ice >= 0) and np.all(new_ice <= run.constant("xmaxd"))


def test_mkclouds_slingo_cover_and_random_overlap(column):
    """Clouds (mkclouds): convective cloud of total cover Slingo's
    0.245 + 0.125 ln P (P in mm/day, clamped to [0.05, 0.8]) spread over
    the convective levels so that their random overlap gives that total;
    stratiform cover ((RH - RHc) / (1 - RHc))^2 above the critical
    relative humidity (Slingo 1987), with Tetens saturation; the column
    total is the random overlap of the layers; the column's vapour path is
    sum q dsigma p_s / g."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(22)
    sig = run.table("sigma")
    ps = rng.uniform(9e4, 1.03e5, nh)
    run.field("dp")[:] = ps
    t = 290.0 * sig[:, None] ** 0.19 + rng.normal(0, 1, (L, nh))
    _planes(run, "dt", nh)[:L] = t
    _planes(run, "dtdt", nh)[:] = 0.0
    _planes(run, "dqdt", nh)[:] = 0.0
    rh = rng.uniform(0.3, 1.0, (L, nh))
    qs = _saturation(run, t, sig[:, None] * ps)
    _planes(run, "dq", nh)[:L] = rh * qs
    conv = (rng.uniform(0, 1, (L, nh)) < 0.3) & (sig[:, None] > 0.4)
    run.field("icclev")[:] = 0.0
    _planes(run, "icclev", nh)[:L] = conv.astype(float)
    n_conv = conv.sum(axis=0)
    run.field("icctot")[:] = n_conv
    rain = rng.uniform(0, 3e-7, nh) * (n_conv > 0)  # m/s
    run.field("dprc")[:] = rain
    run.field("dw")[:] = 0.0
    assert run.constant("clwfac") <= 0  # no low-level vertical-velocity factor
    run.routine("mkclouds")
    cc = _planes(run, "dcc", nh)
    total = 0.245 + 0.125 * np.log(np.maximum(rain, 1e-300) * 86400e3)  # mm/day
    total = np.clip(total, 0.05, 0.8)
    rhc = run.table("rcrit")[:, None]
    strat = np.where(rh > rhc, ((rh - rhc) / (1 - rhc)) ** 2, 0.0)
    pure_conv = conv & (rh < rhc)  # a convective level without stratiform cloud
    cols = np.where((n_conv > 0) & np.all(~conv | pure_conv, axis=0))[0]
    assert cols.size > 50
    conv_cover = np.where(conv[:, cols], cc[:L, cols], 0.0)
    np.testing.assert_allclose(1 - np.prod(1 - conv_cover, axis=0), total[cols], rtol=1e-6)
    plain = ~conv & (rh > rhc + 0.05)
    np.testing.assert_allclose(cc[:L][plain], np.minimum(strat[plain], 1.0), rtol=1e-5)
    np.testing.assert_allclose(cc[L], 1 - np.prod(1 - cc[:L], axis=0), rtol=1e-12, atol=1e-15)
    vapour = np.sum(rh * qs * run.table("dsigma")[:, None], axis=0) * ps / run.constant("ga")
    np.testing.assert_allclose(run.field("dqvi"), vapour, rtol=1e-12)


def test_land_surface_river_store_albedo_and_surface_air(column):
    """Land surface: the river store gains the cell's runoff over the time
    step; a snow pack blends the snow albedo into the ground's as
    d / (d + 1 cm) (a convex weight: the ground's without snow, the snow's
    under deep snow); the surface air takes the land temperature and its
    Tetens saturation humidity."""
    run, nh, L = _run(column)
    rng = np.random.default_rng(23)
    land = rng.uniform(0, 1, nh) < 0.6
    run.field("dls")[:] = land.astype(float)
    run.field("dglac")[:] = 0.0
    river = rng.uniform(0, 10, nh)
    run.field("driver")[:] = river
    runoff = rng.uniform(0, 1e-6, nh)
    run.field("drunoff")[:] = runoff
    # pairs of columns (2i, 2i + 1) alike but for their snow depth: the
    # snow albedo is then the same for both and cancels from the ratio
    ts = np.repeat(rng.uniform(240, 300, nh // 2), 2)
    run.field("land_dts")[:] = ts
    ps = rng.uniform(8e4, 1.04e5, nh)
    run.field("dp")[:] = ps
    snow = 10 ** rng.uniform(-5, 1, nh)
    snow[::4] = 0.0
    run.field("dsnowz")[:] = snow
    run.field("dforest")[:] = np.repeat(rng.uniform(0, 1, nh // 2), 2)
    land[1::2] = land[::2]
    run.field("dls")[:] = land.astype(float)
    run.routine("land")
    np.testing.assert_allclose(run.field("driver"), np.where(land, river + runoff * run.constant("deltsec"), river),
                               rtol=1e-12)
    ground = run.field("dalbclim")
    alb = run.field("dalb")
    bare = land & (snow == 0)
    np.testing.assert_array_equal(alb[bare], ground[bare])
    w = snow / (snow + 0.01)
    even = np.arange(0, nh, 2)
    both = land[even] & (snow[even] > 0) & (snow[even + 1] > 0)
    a, b = even[both], even[both] + 1
    assert a.size > 100 and np.all(ground[a] == ground[b])
    np.testing.assert_allclose((alb[a] - ground[a]) * w[b], (alb[b] - ground[b]) * w[a], rtol=1e-9, atol=1e-15)
    assert np.all((alb[land] > 0) & (alb[land] < 1))
    np.testing.assert_array_equal(_planes(run, "dt", nh)[L][land], ts[land])
    np.testing.assert_allclose(_planes(run, "dq", nh)[L][land], _saturation(run, ts, ps)[land], rtol=1e-12)


def test_carbon_weathering_law(column):
    """Carbon: silicate weathering follows Walker, Hays and Kasting (1981),
    W = (pCO2 / pCO2_E)^beta exp(k (T - 288 K)) (P / P_E)^0.65, relative to
    the present Earth (W = 1 there), scaled by the planet's land area
    against Earth's (radius^2 over Earth's, over Earth's land fraction
    0.29); none over the sea or below freezing; precipitation is converted
    from m/s to cm/yr; the running means gain the time-weighted values."""
    run, nh, L = _run(column)
    if run.constant("ncarbon") == 0:
        pytest.skip("carbon cycle off in this configuration")
    rng = np.random.default_rng(24)
    assert run.constant("nsupply") == 0
    beta, k, P_E, co2_E = (run.constant(n) for n in ("cbeta", "kact", "pearth", "co2earth"))
    run.field("carbon_step")[:] = 1.0
    land = rng.uniform(0, 1, nh) < 0.6
    run.field("dls")[:] = land.astype(float)
    ts = rng.uniform(260, 310, nh)
    _planes(run, "dt", nh)[L] = ts
    ps = rng.uniform(8e4, 1.1e5, nh)
    run.field("dp")[:] = ps
    co2 = 400.0
    run.field("co2")[:] = co2
    seconds_per_year = 3.154e7  # 365.0 days
    precip_cm = rng.uniform(10, 300, nh)
    run.field("dprc")[:] = precip_cm / 100.0 / seconds_per_year * 0.3
    run.field("dprl")[:] = precip_cm / 100.0 / seconds_per_year * 0.7
    for name in ("localavgprecip", "localavgtemps", "localavgweather"):
        run.field(name)[:] = 0.0
    run.routine("carbon")
    np.testing.assert_allclose(run.field("localprecip"), precip_cm, rtol=1e-12)
    pco2 = co2 * ps * 1e-5  # ppm of the column, in units of Earth's 1 bar
    W = (pco2 / co2_E) ** beta * np.exp(k * (ts - 288.0)) * (precip_cm / P_E) ** 0.65
    area = run.constant("plarad") ** 2 / (6.371e6 ** 2 * 0.29)
    warm = land & (ts >= 273.15)
    np.testing.assert_allclose(run.field("localweathering")[warm], W[warm] * area, rtol=2e-3)
    assert np.all(run.field("localweathering")[~warm] == 0)
    tw = run.constant("timeweight")
    np.testing.assert_allclose(run.field("localavgtemps"), ts * tw, rtol=1e-12)
    np.testing.assert_allclose(run.field("localavgweather"), run.field("localweathering") * tw, rtol=1e-12)
