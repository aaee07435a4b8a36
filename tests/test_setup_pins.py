"""CPU pins of the FP64 host set-up against closed forms and library
routines (no GPU): Gauss-Legendre grid, associated Legendre functions and
their derivative form, physics filters, Fourier tables, sigma levels,
semi-implicit inverse, restoration profile, calendar and the seeded planet.
"""
import numpy as np
import pytest
from scipy.special import gammaln, lpmv


@pytest.fixture(scope="module")
def t21(exo):
    e = exo.Engine(-1, 4)
    m = e.add_member(exo.default_config())
    yield e, m
    e.close()


def pbar(m, n, x):
    norm = np.sqrt((2 * n + 1) / 2 * np.exp(gammaln(n - m + 1) - gammaln(n + m + 1)))
    return norm * lpmv(m, n, x)


def modes(T):
    for m in range(T + 1):
        for n in range(m, T + 1):
            yield m, n


@pytest.mark.parametrize("trunc,nlat", [(21, 32), (42, 64), (63, 96)])
def test_gaussian_grid_is_gauss_legendre(exo, trunc, nlat):
    """inigau (grid/index.ts:32-63) gives the Gauss-Legendre nodes and
    weights, north to south."""
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config(truncation=trunc, timestep_minutes=30))
        x, w = np.polynomial.legendre.leggauss(nlat)
        sid = e.read_table(m, "sid")
        zgw = e.read_table(m, "zgw").reshape(nlat, -1)[:, 0]
        assert np.all(np.diff(sid) < 0)
        np.testing.assert_allclose(sid, x[::-1], atol=1e-14)
        np.testing.assert_allclose(zgw, w[::-1], rtol=5e-12)
        assert abs(zgw.sum() - 2.0) < 1e-13


def test_legendre_functions_match_scipy(t21):
    """qi = orthonormal associated Legendre functions with the
    Condon-Shortley phase (legendre/index.ts:29-108)."""
    e, m = t21
    T, nlat, _, _ = e.dims(m)
    sid = e.read_table(m, "sid")
    qi = e.read_table(m, "qi_raw").reshape(nlat, -1)
    for w, (mm, n) in enumerate(modes(T)):
        np.testing.assert_allclose(qi[:, w], pbar(mm, n, sid), atol=5e-13)


def test_legendre_orthonormal_under_quadrature(t21):
    e, m = t21
    T, nlat, _, _ = e.dims(m)
    qi = e.read_table(m, "qi_raw").reshape(nlat, -1)
    wts = e.read_table(m, "zgw").reshape(nlat, -1)[:, 0]
    idx = list(modes(T))
    for mm in range(T + 1):
        cols = [w for w, (a, _) in enumerate(idx) if a == mm]
        G = qi[:, cols].T @ (wts[:, None] * qi[:, cols])
        np.testing.assert_allclose(G, np.eye(len(cols)), atol=1e-12)


def test_derivative_table_is_one_minus_mu2_dp(t21):
    """qj (before the physics filter) = (1 - mu^2) dP/dmu: the identity
    (1-x^2) P_n^m' = (n+m) P_{n-1}^m - n x P_n^m on unnormalised P."""
    e, m = t21
    T, nlat, _, _ = e.dims(m)
    sid = e.read_table(m, "sid")
    qj = e.read_table(m, "legendre_qj").reshape(nlat, -1)
    idx = list(modes(T))
    for w, (mm, n) in enumerate(idx):
        shape = np.exp(-8.0 * ((n + 1) / T) ** 8)  # skspgp, nfilter = 2
        norm = np.sqrt((2 * n + 1) / 2 * np.exp(gammaln(n - mm + 1) - gammaln(n + mm + 1)))
        prev = lpmv(mm, n - 1, sid) if n - 1 >= mm else 0.0
        d = norm * ((n + mm) * prev - n * sid * lpmv(mm, n, sid))
        np.testing.assert_allclose(qj[:, w] / shape, d, atol=1e-11)


def test_kernel_tables_fold_weights_and_filters(t21):
    """qc = qi w, qe = qj w/(1-mu^2), qm = qi w m/(1-mu^2),
    qq = qi w n(n+1)/(2(1-mu^2)), qu = qi m/(n(n+1)), qv = qj/(n(n+1)),
    with skgpsp on the first four and skspgp on the rest
    (legendre/index.ts:93-104, :202-221)."""
    e, m = t21
    T, nlat, _, _ = e.dims(m)
    sid = e.read_table(m, "sid")
    wts = e.read_table(m, "zgw").reshape(nlat, -1)[:, 0]
    qi = e.read_table(m, "qi_raw").reshape(nlat, -1)
    qj = e.read_table(m, "legendre_qj").reshape(nlat, -1)
    tab = {n: e.read_table(m, "legendre_" + n).reshape(nlat, -1) for n in ("qc", "qe", "qm", "qq", "qu", "qv", "qi")}
    cos2 = 1 - sid ** 2
    for w, (mm, n) in enumerate(modes(T)):
        f = np.exp(-8.0 * ((n + 1) / T) ** 8)
        inv = 1.0 / (n * (n + 1)) if n > 0 else 0.0
        dj = qj[:, w] / f
        np.testing.assert_allclose(tab["qc"][:, w], qi[:, w] * wts * f, atol=1e-14)
        np.testing.assert_allclose(tab["qi"][:, w], qi[:, w] * f, atol=1e-14)
        np.testing.assert_allclose(tab["qe"][:, w], dj * wts / cos2 * f, atol=1e-12)
        np.testing.assert_allclose(tab["qm"][:, w], qi[:, w] * wts / cos2 * mm * f, atol=1e-12)
        np.testing.assert_allclose(tab["qq"][:, w], qi[:, w] * wts / cos2 * n * (n + 1) * 0.5 * f, atol=1e-11)
        np.testing.assert_allclose(tab["qu"][:, w], qi[:, w] * inv * mm * f, atol=1e-14)
        np.testing.assert_allclose(tab["qv"][:, w], dj * inv * f, atol=1e-13)


def test_fourier_tables_closed_form(t21):
    e, m = t21
    T, nlat, nlon, _ = e.dims(m)
    nfc = 2 * (T + 1)
    fwd = e.read_table(m, "fourier_forward").reshape(nlon, nfc)
    inv = e.read_table(m, "fourier_inverse").reshape(nfc, nlon)
    lam = 2 * np.pi * np.arange(nlon) / nlon
    g = np.random.default_rng(1).standard_normal(nlon)
    c = np.fft.rfft(g) / nlon
    np.testing.assert_allclose(g @ fwd[:, 0::2], c.real[:T + 1], atol=1e-14)
    np.testing.assert_allclose(g @ fwd[:, 1::2], c.imag[:T + 1], atol=1e-14)
    # fc2gp of a truncated spectrum is the real synthesis with Im c0 ignored
    spec = np.zeros(nlon // 2 + 1, complex)
    spec[:T + 1] = c[:T + 1]
    spec[0] = spec[0].real
    back = np.fft.irfft(spec * nlon, nlon)
    four = np.empty(nfc)
    four[0::2] = c.real[:T + 1]
    four[1::2] = c.imag[:T + 1]
    np.testing.assert_allclose(four @ inv, back, atol=1e-12)
    assert np.all(inv[1, :] == 0.0) and np.allclose(inv[0, :], 1.0)
    del lam


def test_sigma_levels_closed_form(t21):
    """initpm, NEQSIG = 0 (dynamics/setup/index.ts:33-41)."""
    e, m = t21
    L = e.dims(m)[3]
    z = np.arange(1, L + 1) / L
    sh = 0.75 * z + 1.75 * z ** 3 - 1.5 * z ** 4
    np.testing.assert_allclose(e.read_table(m, "sigmah"), sh, atol=1e-15)
    ds = e.read_table(m, "dsigma")
    assert abs(ds.sum() - 1.0) < 1e-14
    sg = e.read_table(m, "sigma")
    np.testing.assert_allclose(sg, 0.5 * (sh + np.concatenate([[0.0], sh[:-1]])), atol=1e-15)


@pytest.mark.parametrize("phase", [0, 3])
def test_bm1_is_inverse_of_semi_implicit_matrix(t21, phase):
    """makebm: bm1[n] = (I / (n(n+1)) + delt^2 B)^-1 with B(l1, l2) =
    t0[l1] dsigma[l2] + sum_k g(k, l1) tau(l2, k) (dynamics/setup/
    index.ts:207-232); the inverse is checked with numpy's solver."""
    e, m = t21
    T, _, _, L = e.dims(m)
    k = e.constants(m, phase)
    g = e.read_table(m, "g").reshape(L, L)
    tau = e.read_table(m, "tau").reshape(L, L)
    t0 = e.read_table(m, "t0")
    ds = e.read_table(m, "dsigma")
    B = t0[:, None] * ds[None, :] + g @ tau
    bm1 = e.read_table(m, f"bm1_phase{phase}").reshape(T, L, L)
    for n in range(1, T + 1):
        A = np.eye(L) / (n * (n + 1)) + k["delt"] ** 2 * B
        np.testing.assert_allclose(bm1[n - 1], np.linalg.inv(A), rtol=1e-9, atol=1e-9 * np.abs(bm1[n - 1]).max())


def test_timestep_and_pumamod_constants(t21):
    """PUMAMOD.read (setup/pumamod/index.ts:103-232) for the T21 planet:
    45-minute step, 24 h solar day of a 360-day year, N2 atmosphere."""
    e, m = t21
    k = e.constants(m, 3)
    sidereal = 86400 * 359 / 360
    ww = 2 * 3.14159265359 / sidereal
    assert k["ww"] == pytest.approx(ww, rel=1e-14)
    assert k["cv"] == pytest.approx(6371220 * ww, rel=1e-14)
    assert k["ct"] == pytest.approx((6371220 * ww) ** 2 / (8314.46261815324 / 28.0134), rel=1e-13)
    assert k["deltsec"] == 2700.0
    assert k["delt"] == pytest.approx(2 * 3.14159265359 / 32, rel=1e-14)
    assert k["filter"] == 1.0 and e.constants(m, 0)["filter"] == 0.0
    assert e.constants(m, 0)["deltsec"] == 2700.0 / 8
    assert k["pnu21"] == pytest.approx(0.8)
    assert e.phases(m) == (4, 3)


def test_calendar_declination_closed_form(t21):
    """With e = 0 the sun's longitude advances uniformly and sin(decl) =
    sin(obliquity) sin(longitude) (calendar/solar/index.ts:63-93)."""
    e, m = t21
    eps = np.deg2rad(23.4)
    steps_per_year = 11520
    for step in (0, 1000, 5000, 11000):
        v = e.step_values(m, step)
        day = (step % steps_per_year) / steps_per_year
        mvelpp = np.deg2rad(102.7 + 180)
        lon = 0.0 + (day - 80.5 / 365) * 2 * np.pi
        assert v[0] == pytest.approx(np.sin(eps) * np.sin(lon), abs=1e-12)
        assert v[3] == pytest.approx(1.0, abs=1e-14)
        del mvelpp


def test_seeded_planet(t21):
    e, m = t21
    T, nlat, nlon, L = e.dims(m)
    land = e.read_table(m, "land_mask")
    assert abs(land.mean() - 0.4) < 1e-3
    ini = e.read_table(m, "initial")
    lay = {n: (o, c) for n, o, c in e.layout(m)}
    o, c = lay["sz"]
    sz = ini[o:o + c].reshape(L, -1)
    assert np.all(sz[:, 2] == pytest.approx(1.63299310207))
    o, c = lay["sd"]
    assert np.all(ini[o:o + c] == 0)
    # sea points carry the SST climatology at the surface level
    o, c = lay["dt"]
    ts = ini[o:o + c].reshape(L + 1, -1)[L]
    assert np.all(ts[land == 0] == 280.0)


def test_registry_sizes(exo):
    """Arena sizes of the registry (the reference quotes an 11 MB T21
    checkpoint, docs/climate/exoplasim-webgpu.md:703)."""
    with exo.Engine(-1, 2) as e:
        a = e.add_member(exo.default_config())
        assert e.arena_floats(a) == 2429594
        names = [n for n, _, _ in e.layout(a)]
        assert names[:4] == ["sd", "st", "sz", "sq"]
        assert names[-3:] == ["snapshot", "snapshot_tracking", "snapshot_step"]
        assert len(names) == len(set(names))
