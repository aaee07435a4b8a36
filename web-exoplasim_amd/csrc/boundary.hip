// The boundary data path either side of the step: a terrain raster
// regridded onto a Gaussian grid's cells (GCM_BOUNDARY.fromTerrain,
// src/model/climate/gcm/boundary/index.ts:12-41, with GCM_TERRAIN_REGRID.
// cell, terrain/regrid/index.ts:6-46) and written as ExoPlaSim .sra surface
// files (toSra and files, boundary/index.ts:43-79).
//
//   k_boundary_cells  one thread per grid cell: the cell's pixel
//                     rectangle summed row by row in FP64, each pixel
//                     weighted by its area (sin-latitude band x longitude
//                     overlap), in the reference's order and without
//                     fused multiply-adds, so the sums are the reference's
//                     up to the last bits of the device sine
//
// A 4096 x 2048 raster is 64 MB of input read once: HBM-bound byte work.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cctype>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/exoplasim_mi355x.h"
#include "engine.h"
#include "setup.h"

namespace {

constexpr int kMaxTerrainPixels = 8388608;  // terrain/index.ts:9

struct CellGrid {
  int nlat, nlon, width, height;
  double lon_step;
  const double* lat_deg;  // nlat, north first
};

}  // namespace

extern "C" __global__ __launch_bounds__(256) void k_boundary_cells(const float* __restrict__ elevation,
                                                                 const float* __restrict__ land, CellGrid g,
                                                                 double gravity, double* __restrict__ land_fraction,
                                                                 double* __restrict__ geopotential) {
#pragma clang fp contract(off)
  const int index = blockIdx.x * 256 + threadIdx.x;
  if (index >= g.nlat * g.nlon) return;
  const int row = index / g.nlon, column = index - row * g.nlon;
  const double north = row == 0 ? 90.0 : (g.lat_deg[row - 1] + g.lat_deg[row]) / 2;
  const double south = row == g.nlat - 1 ? -90.0 : (g.lat_deg[row] + g.lat_deg[row + 1]) / 2;
  const double center = column * g.lon_step;
  const double west = center - g.lon_step / 2, east = center + g.lon_step / 2;
  const double top = ((90 - north) * g.height) / 180;
  const double bottom = ((90 - south) * g.height) / 180;
  const double left = ((west + 180) * g.width) / 360;
  const double right = ((east + 180) * g.width) / 360;
  const double rad = M_PI / 180;
  double area = 0, land_area = 0, elevation_sum = 0;
  for (int y = (int)fmax(0.0, floor(top)); y < (int)ceil(bottom) && y < g.height; y++) {
    const double pixel_north = 90 - (fmax(top, (double)y) * 180) / g.height;
    const double pixel_south = 90 - (fmin(bottom, (double)(y + 1)) * 180) / g.height;
    const double latitude_weight = sin(pixel_north * rad) - sin(pixel_south * rad);
    for (int x = (int)floor(left); x < (int)ceil(right); x++) {
      const double weight = latitude_weight * (fmin(right, (double)(x + 1)) - fmax(left, (double)x));
      const int i = y * g.width + ((x % g.width) + g.width) % g.width;
      const double l = (double)land[i];
      area += weight;
      land_area += weight * l;
      elevation_sum += weight * l * (double)elevation[i];
    }
  }
  land_fraction[index] = fmin(1.0, fmax(0.0, land_area / area));
  geopotential[index] = gravity * (elevation_sum / area);
}

namespace {

// JavaScript's Number.prototype.toFixed(3) (ECMA-262 21.1.3.3): the sign
// comes off first, then the integer n nearest |x| * 1000 from |x|'s exact
// decimal value, the larger n on a tie; -0 prints as 0.000.
std::string to_fixed3(double x) {
  if (x == 0) return "0.000";
  char exact[1200];
  std::snprintf(exact, sizeof(exact), "%.400f", std::fabs(x));  // exact binary value
  std::string s(exact);
  const size_t dot = s.find('.');
  std::string whole = s.substr(0, dot), frac = s.substr(dot + 1);
  std::string keep = frac.substr(0, 3), rest = frac.substr(3);
  // rest against one half: above or tied rounds |x| up
  int cmp = 0;
  if (rest[0] > '5') cmp = 1;
  else if (rest[0] < '5') cmp = -1;
  else cmp = rest.find_first_not_of('0', 1) == std::string::npos ? 0 : 1;
  const bool up = cmp >= 0;
  std::string digits = whole + keep;
  if (up) {
    int i = (int)digits.size() - 1;
    while (i >= 0 && digits[i] == '9') digits[i--] = '0';
    if (i < 0) digits.insert(digits.begin(), '1');
    else digits[i]++;
  }
  const std::string body = digits.substr(0, digits.size() - 3) + "." + digits.substr(digits.size() - 3);
  return x < 0 ? "-" + body : body;
}

std::string pad_start(const std::string& s, size_t width) {
  return s.size() >= width ? s : std::string(width - s.size(), ' ') + s;
}

}  // namespace

extern "C" int exo_boundary_from_terrain(exo_engine* h, const float* elevation_m, const float* land, int width,
                                         int height, int nlat, double gravity_ms2, double* land_fraction,
                                         double* geopotential) {
  if (!h || !elevation_m || !land || !land_fraction || !geopotential)
    return exo::set_error(EXO_E_ARG, "missing argument");
  // terrain/index.ts:11-22
  if (height < 2 || width != height * 2 || (long long)width * height > kMaxTerrainPixels)
    return exo::set_error(EXO_E_ARG,
                          "Terrain must be a full-world 2:1 raster, at least 4 \xC3\x97 2 and at most 4096 \xC3\x97 "
                          "2048 pixels");
  if (nlat != 32 && nlat != 64 && nlat != 96) return exo::set_error(EXO_E_ARG, "no grid of " + std::to_string(nlat) + " latitudes");
  const size_t n = (size_t)width * height;
  // terrain/index.ts:26-44
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(elevation_m[i]) || !std::isfinite(land[i]) || land[i] < 0 || land[i] > 1)
      return exo::set_error(EXO_E_ARG, "Terrain requires finite elevations and land coverage between 0 and 1");
  exo::Engine& e = h->e;
  if (e.device < 0) return exo::set_error(EXO_E_STATE, "a host-only engine cannot regrid terrain");
  const int nlon = 2 * nlat, cells = nlat * nlon;
  const std::vector<double> degrees = exo::gaussian_degrees(nlat);
  float *d_elev = nullptr, *d_land = nullptr;
  double *d_lat = nullptr, *d_out = nullptr;
  try {
    HIP_OR_THROW(hipSetDevice(e.device));
    HIP_OR_THROW(hipMallocAsync((void**)&d_elev, n * sizeof(float), e.stream));
    HIP_OR_THROW(hipMallocAsync((void**)&d_land, n * sizeof(float), e.stream));
    HIP_OR_THROW(hipMallocAsync((void**)&d_lat, nlat * sizeof(double), e.stream));
    HIP_OR_THROW(hipMallocAsync((void**)&d_out, 2 * (size_t)cells * sizeof(double), e.stream));
    HIP_OR_THROW(hipMemcpyAsync(d_elev, elevation_m, n * sizeof(float), hipMemcpyHostToDevice, e.stream));
    HIP_OR_THROW(hipMemcpyAsync(d_land, land, n * sizeof(float), hipMemcpyHostToDevice, e.stream));
    HIP_OR_THROW(hipMemcpyAsync(d_lat, degrees.data(), nlat * sizeof(double), hipMemcpyHostToDevice, e.stream));
    const CellGrid g{nlat, nlon, width, height, 360.0 / nlon, d_lat};
    hipLaunchKernelGGL(k_boundary_cells, dim3((cells + 255) / 256), dim3(256), 0, e.stream, d_elev, d_land, g,
                       gravity_ms2, d_out, d_out + cells);
    HIP_OR_THROW(hipGetLastError());
    HIP_OR_THROW(hipMemcpyAsync(land_fraction, d_out, cells * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    HIP_OR_THROW(hipMemcpyAsync(geopotential, d_out + cells, cells * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    HIP_OR_THROW(hipFreeAsync(d_elev, e.stream));
    HIP_OR_THROW(hipFreeAsync(d_land, e.stream));
    HIP_OR_THROW(hipFreeAsync(d_lat, e.stream));
    HIP_OR_THROW(hipFreeAsync(d_out, e.stream));
    HIP_OR_THROW(hipStreamSynchronize(e.stream));
  } catch (const std::exception& ex) {
    return exo::set_error(EXO_E_DEVICE, ex.what());
  }
  return EXO_OK;
}

extern "C" int exo_boundary_sra(int code, int nlat, const double* values, char* out, size_t capacity,
                                size_t* length) {
  if (!values || !length || nlat < 1) return exo::set_error(EXO_E_ARG, "missing argument");
  const int nlon = 2 * nlat;
  // toSra (boundary/index.ts:43-57): a header of eight integers padded to
  // 11, then eight values a line, each toFixed(3) padded to 9
  const long long header[8] = {code, 0, 20240101, 0, nlon, nlat, 0, 0};
  std::string text;
  for (long long v : header) text += " " + pad_start(std::to_string(v), 11);
  const size_t count = (size_t)nlon * nlat;
  for (size_t start = 0; start < count; start += 8) {
    text += "\n";
    for (size_t i = start; i < start + 8 && i < count; i++) text += " " + pad_start(to_fixed3(values[i]), 9);
  }
  text += "\n";
  *length = text.size();
  if (!out) return EXO_OK;
  if (capacity < text.size()) return exo::set_error(EXO_E_ARG, "the .sra text needs " + std::to_string(text.size()) + " bytes");
  std::memcpy(out, text.data(), text.size());
  return EXO_OK;
}

extern "C" int exo_boundary_read_sra(int code, int nlat, const char* text, size_t length, double* out) {
  if (!text || !out || nlat < 1) return exo::set_error(EXO_E_ARG, "missing argument");
  const size_t points = 2 * (size_t)nlat * nlat, record = 8 + points;
  char name[64];
  std::snprintf(name, sizeof(name), "N%03d_surf_%04d.sra", nlat, code);
  // text.trim().split(/\s+/).map(Number): a token that is not wholly a
  // number is NaN
  std::vector<double> tokens;
  size_t i = 0;
  std::string token;
  while (i < length) {
    while (i < length && std::isspace((unsigned char)text[i])) i++;
    const size_t start = i;
    while (i < length && !std::isspace((unsigned char)text[i])) i++;
    if (i == start) break;
    token.assign(text + start, i - start);
    char* end = nullptr;
    const double v = std::strtod(token.c_str(), &end);
    tokens.push_back(end && *end == '\0' ? v : NAN);
    if (tokens.size() >= record) break;  // one field is all a one-month read takes
  }
  if (tokens.size() < record)
    return exo::set_error(EXO_E_FORMAT, std::string(name) + " has 0 fields, expected 1");
  std::copy(tokens.begin() + 8, tokens.begin() + record, out);
  return EXO_OK;
}
