"""The boundary data path on the CPU: the .sra writer (exo_boundary_sra,
host C++) against the oracle (oracle/boundary.py) and JavaScript's own
toFixed results, the raster validation of exo_boundary_from_terrain, and
the oracle's regrid pinned by closed forms:

* toFixed(3) rounds the exact binary value half up in magnitude
  (ECMA-262 21.1.3.3): 0.0625 -> "0.063", -0.0625 -> "-0.063", 1.0005 ->
  "1.000" (its double is below 1.0005), 0.9995 -> "1.000" (above),
  -0.0001 -> "-0.000", -0 -> "0.000";
* a uniform raster regrids to itself; a raster whose western hemisphere
  is land gives cells wholly west land fraction 1 and wholly east 0;
* regridding conserves the area-weighted land fraction: the cells' land
  fractions weighted by their areas (sin-latitude band x 1 / NLON) sum to
  the raster's area-weighted land fraction.
"""
import math

import numpy as np
import pytest

from oracle import boundary as ob

JS_TO_FIXED3 = {0.0625: "0.063", -0.0625: "-0.063", 1.0005: "1.000", 0.9995: "1.000", -0.0001: "-0.000",
                -0.0: "0.000", 0.0: "0.000", 0.1235: "0.123", 12345678.9: "12345678.900", 2.5: "2.500"}


def test_to_fixed_matches_javascript(exo):
    values = list(JS_TO_FIXED3)
    text = exo.boundary_sra(172, 4, values + [0.0] * (32 - len(values)))
    got = " ".join(text.splitlines()[1:]).split()[:len(values)]
    assert got == [JS_TO_FIXED3[v] for v in values]
    assert [ob.to_fixed3(v) for v in values] == [JS_TO_FIXED3[v] for v in values]


def test_sra_text_matches_the_oracle(exo):
    rng = np.random.default_rng(4)
    for nlat in (32, 64):
        values = rng.normal(0, 3000, 2 * nlat * nlat)
        values[:10] = [0.0005, -0.0005, 1e-9, -1e-9, 999999.9995, 0.5, -0.5, 100.0625, 7.0, -0.0]
        for code in (172, 129):
            assert exo.boundary_sra(code, nlat, values) == ob.to_sra(code, nlat, values)
    text = exo.boundary_sra(129, 32, np.zeros(2048))
    head = text.splitlines()[0]
    assert head == "".join(" " + str(v).rjust(11) for v in (129, 0, 20240101, 0, 64, 32, 0, 0))
    assert len(text.splitlines()) == 1 + 2048 // 8
    files = exo.boundary_files(32, np.zeros(2048), np.ones(2048))
    assert sorted(files) == ["N032_surf_0129.sra", "N032_surf_0172.sra"]
    assert ob.file_name(32, 172) in files


def test_terrain_validation(exo):
    host = exo.Engine(-1, 1)
    try:
        with pytest.raises(exo.ExoError, match="full-world 2:1 raster"):
            host.boundary_from_terrain(np.zeros((4, 6)), np.zeros((4, 6)), 32)
        bad = np.zeros((4, 8), dtype=np.float32)
        land = bad.copy()
        land[0, 0] = 1.5
        with pytest.raises(exo.ExoError, match="land coverage between 0 and 1"):
            host.boundary_from_terrain(bad, land, 32)
        with pytest.raises(exo.ExoError, match="host-only"):
            host.boundary_from_terrain(bad, np.zeros_like(bad), 32)
    finally:
        host.close()


def _cell_areas(nlat):
    lat = ob.gaussian_degrees(nlat)
    edges = np.concatenate([[90.0], (lat[:-1] + lat[1:]) / 2, [-90.0]])
    band = np.sin(np.radians(edges[:-1])) - np.sin(np.radians(edges[1:]))
    return np.repeat(band[:, None], 2 * nlat, axis=1) / (2 * nlat)


def test_oracle_regrid_closed_forms():
    h, w = 32, 64
    elev = np.full((h, w), 1000.0, dtype=np.float32)
    land = np.full((h, w), 0.3, dtype=np.float32)
    frac, geo = ob.from_terrain(elev, land, 32, 9.81)
    np.testing.assert_allclose(frac, np.float32(0.3), rtol=1e-12)
    np.testing.assert_allclose(geo, 9.81 * float(np.float32(0.3)) * 1000.0, rtol=1e-12)
    # western hemisphere land (raster columns from 180W; cells from 0E)
    land = np.zeros((h, w), dtype=np.float32)
    land[:, : w // 2] = 1.0
    frac, _ = ob.from_terrain(elev, land, 32, 9.81)
    lon = np.arange(64) * 360 / 64
    west = (lon - 360 / 128 >= 180) & (lon + 360 / 128 <= 360)
    east = (lon - 360 / 128 >= 0) & (lon + 360 / 128 <= 180)
    assert np.all(frac[:, west] == 1.0) and np.all(frac[:, east] == 0.0)
    # conservation of the area-weighted land fraction
    rng = np.random.default_rng(9)
    land = rng.uniform(0, 1, (h, w)).astype(np.float32)
    for nlat in (32, 64):
        frac, _ = ob.from_terrain(elev, land, nlat, 9.81)
        pixel_rows = np.sin(np.radians(90 - np.arange(h) * 180 / h)) - np.sin(np.radians(90 - (np.arange(h) + 1) * 180 / h))
        raster = np.sum(pixel_rows[:, None] * land.astype(np.float64)) / w
        cells = np.sum(_cell_areas(nlat) * frac)
        assert cells == pytest.approx(raster, rel=1e-12)
        assert math.isclose(np.sum(_cell_areas(nlat)), 2.0, rel_tol=1e-12)


def test_sra_round_trip_and_boundary_member(exo):
    """The .sra reader returns what the writer wrote (to its 3 decimals),
    rejects a text without a whole field with the setup's message, and a
    host member built from a land fraction and orography takes its land
    mask (fraction > 0.5) and surface geopotential from them."""
    rng = np.random.default_rng(2)
    nlat = 32
    frac = rng.uniform(0, 1, (nlat, 2 * nlat))
    geo = np.where(frac > 0.5, rng.uniform(0, 3e4, frac.shape), 0.0)
    files = exo.boundary_files(nlat, frac, geo)
    back_f = exo.boundary_read_sra(172, nlat, files["N032_surf_0172.sra"])
    back_g = exo.boundary_read_sra(129, nlat, files["N032_surf_0129.sra"])
    np.testing.assert_allclose(back_f, frac, atol=5e-4 + 1e-12)
    np.testing.assert_allclose(back_g, geo, atol=5e-4 + 1e-9)
    np.testing.assert_array_equal(back_f, [[float(ob.to_fixed3(v)) for v in row] for row in frac])
    with pytest.raises(exo.ExoError, match="N032_surf_0172.sra has 0 fields, expected 1"):
        exo.boundary_read_sra(172, nlat, files["N032_surf_0172.sra"][:500])
    host = exo.Engine(-1, 2)
    try:
        cfg = exo.default_config(truncation=21, levels=5)
        m = host.add_member(cfg, land_fraction=back_f, geopotential=back_g)
        initial = host.read_table(m, "initial")
        off, n = {name: (o, k) for name, o, k in host.layout(m)}["dls"]
        np.testing.assert_array_equal(initial[off:off + n], (back_f > 0.5).reshape(-1).astype(float))
        seeded = host.add_member(cfg)
        assert not np.array_equal(host.read_table(seeded, "initial")[off:off + n], initial[off:off + n])
        with pytest.raises(exo.ExoError):
            host.add_member(cfg, land_fraction=back_f[:10], geopotential=back_g[:10])
    finally:
        host.close()
