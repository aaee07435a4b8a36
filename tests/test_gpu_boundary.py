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
