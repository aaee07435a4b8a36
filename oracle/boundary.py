"""CPU restatement of the boundary data path: TEST INFRASTRUCTURE ONLY
(the tests compare exo_boundary_from_terrain and exo_boundary_sra with it;
nothing shipped imports it).

* cell() is GCM_TERRAIN_REGRID.cell (src/model/climate/gcm/terrain/regrid/
  index.ts:6-46): the raster pixels under a latitude-longitude box, each
  weighted by its sin-latitude band times its longitude overlap, summed in
  FP64 row by row.
* from_terrain() is GCM_BOUNDARY.fromTerrain (boundary/index.ts:12-41):
  cell() over the boxes of a Gaussian grid's cells (edges at the midpoints
  of neighbouring latitudes, the poles outermost), land fraction and
  g x mean elevation.
* to_fixed3() is JavaScript's Number.prototype.toFixed(3) (ECMA-262
  21.1.3.3) from the exact decimal value of the double; to_sra() and
  file_name() are toSra and files (boundary/index.ts:43-79).
"""
import math
from decimal import ROUND_HALF_UP, Decimal

import numpy as np


def gaussian_degrees(nlat):
    """Gauss-Legendre latitudes in degrees, north first."""
    x, _ = np.polynomial.legendre.leggauss(nlat)
    return np.degrees(np.arcsin(x[::-1]))


def cell(elevation, land, north, south, west, east):
    height, width = elevation.shape
    top = ((90 - north) * height) / 180
    bottom = ((90 - south) * height) / 180
    left = ((west + 180) * width) / 360
    right = ((east + 180) * width) / 360
    area = land_area = elevation_sum = 0.0
    y = max(0, math.floor(top))
    while y < math.ceil(bottom) and y < height:
        pixel_north = 90 - (max(top, y) * 180) / height
        pixel_south = 90 - (min(bottom, y + 1) * 180) / height
        latitude_weight = math.sin(pixel_north * math.pi / 180) - math.sin(pixel_south * math.pi / 180)
        for x in range(math.floor(left), math.ceil(right)):
            weight = latitude_weight * (min(right, x + 1) - max(left, x))
            l = float(land[y, x % width])
            area += weight
            land_area += weight * l
            elevation_sum += weight * l * float(elevation[y, x % width])
        y += 1
    return min(1.0, max(0.0, land_area / area)), elevation_sum / area


def from_terrain(elevation, land, nlat, gravity, degrees=None):
    lat = gaussian_degrees(nlat) if degrees is None else degrees
    nlon = 2 * nlat
    step = 360 / nlon
    fraction = np.zeros((nlat, nlon))
    geopotential = np.zeros((nlat, nlon))
    for row in range(nlat):
        north = 90 if row == 0 else (lat[row - 1] + lat[row]) / 2
        south = -90 if row == nlat - 1 else (lat[row] + lat[row + 1]) / 2
        for column in range(nlon):
            center = column * step
            f, e = cell(elevation, land, north, south, center - step / 2, center + step / 2)
            fraction[row, column] = f
            geopotential[row, column] = gravity * e
    return fraction, geopotential


def to_fixed3(x):
    x = float(x)
    if x == 0:
        return "0.000"
    body = str(Decimal(abs(x)).quantize(Decimal("0.001"), rounding=ROUND_HALF_UP))
    return ("-" if x < 0 else "") + body


def to_sra(code, nlat, values):
    nlon = 2 * nlat
    header = "".join(" " + str(v).rjust(11) for v in (code, 0, 20240101, 0, nlon, nlat, 0, 0))
    flat = np.asarray(values, dtype=np.float64).reshape(-1)
    lines = [header]
    for start in range(0, len(flat), 8):
        lines.append("".join(" " + to_fixed3(v).rjust(9) for v in flat[start:start + 8]))
    return "\n".join(lines) + "\n"


def file_name(nlat, code):
    return f"N{nlat:03d}_surf_0{code}.sra"
