"""The terrain regrid on the GPU (k_boundary_cells) against the oracle's
GCM_TERRAIN_REGRID.cell (GPU). Random rasters (fractional and binary land,
elevations up to 8 km) onto the N32, N64 and N96 Gaussian grids, and a
raster finer than the cells and one coarser: both sum in FP64 in the
reference's order without fused multiply-adds, so they agree to the last
bits of the device sine (1e-12 relative)."""
import numpy as np
import pytest

from oracle import boundary as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nlat, height", [(32, 128), (64, 64), (96, 256), (32, 16)])
def test_regrid_matches_the_oracle(exo, nlat, height):
    rng = np.random.default_rng(nlat + height)
    width = 2 * height
    elevation = rng.uniform(-500, 8000, (height, width)).astype(np.float32)
    land = rng.uniform(0, 1, (height, width)).astype(np.float32)
    land[rng.uniform(size=land.shape) < 0.3] = 0.0
    land[rng.uniform(size=land.shape) < 0.3] = 1.0
    eng = exo.Engine(0, 1)
    try:
        fraction, geo = eng.boundary_from_terrain(elevation, land, nlat, 9.80665)
    finally:
        eng.close()
    want_f, want_g = ob.from_terrain(elevation, land, nlat, 9.80665)
    np.testing.assert_allclose(fraction, want_f, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(geo, want_g, rtol=1e-12, atol=1e-9)
    # and through the .sra text, digit for digit
    files = exo.boundary_files(nlat, fraction, geo)
    assert files[ob.file_name(nlat, 172)] == ob.to_sra(172, nlat, want_f)


def test_planet_from_terrain_steps_like_the_oracle(exo):
    """Terrain -> regrid (device) -> .sra text -> member: a T21 L10 planet
    built from a raster's land and orography takes its first step within
    1e-5 relative RMS of the FP64 oracle started from the same set-up, as
    the seeded planets do (tests/test_gpu_parity.py)."""
    from oracle_harness import OracleRun

    height = 128
    yy, xx = np.meshgrid(np.linspace(-1, 1, height), np.linspace(-1, 1, 2 * height), indexing="ij")
    elevation = (3000.0 * np.exp(-((xx + 0.3) ** 2 + yy ** 2) / 0.1) - 200.0).astype(np.float32)
    land = (elevation > 0).astype(np.float32)
    gpu = exo.Engine(0, 1)
    host = exo.Engine(-1, 1)
    try:
        fraction, geo = gpu.boundary_from_terrain(elevation, land, 32, 9.80665)
        files = exo.boundary_files(32, fraction, geo)
        frac_in = exo.boundary_read_sra(172, 32, files["N032_surf_0172.sra"])
        geo_in = exo.boundary_read_sra(129, 32, files["N032_surf_0129.sra"])
        assert 0.05 < np.mean(frac_in > 0.5) < 0.6 and geo_in.max() > 1e4
        cfg = exo.default_config(truncation=21, levels=10)
        g = gpu.add_member(cfg, land_fraction=frac_in, geopotential=geo_in)
        h = host.add_member(cfg, land_fraction=frac_in, geopotential=geo_in)
        oracle = OracleRun(host, h)
        gpu.step(1)
        gpu.sync()
        oracle.step(1, physics=True)
        k = host.constants(h, 0)
        for name in ("sd", "st", "sz", "sq", "sp"):
            a = gpu.read_field(g, name).astype(np.float64)
            b = oracle.field(name)
            assert np.all(np.isfinite(a)), name
            if name == "st":
                rel = np.sqrt(np.mean((a - b) ** 2)) * k["ct"] / np.sqrt(2.0) / 250.0
            else:
                rel = np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300)
            print(f"terrain planet step 1 {name}: rel {rel:.2e}")
            assert rel < 1e-5, (name, rel)
    finally:
        gpu.close()
        host.close()
