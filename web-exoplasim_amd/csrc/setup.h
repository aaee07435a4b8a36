// Host set-up of one ExoPlaSim member in FP64: configuration to pumamod,
// Gaussian grid, Legendre/Fourier tables, semi-implicit matrices, physics
// tables and constants, the seeded synthetic planet and its initial state,
// and the per-step calendar values.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/exoplasim_mi355x.h"
#include "model.h"

namespace exo {

struct Layout {
  std::vector<size_t> offset, floats, plane;  // per ExoSlot
  std::vector<char> fourier;                  // per ExoSlot: shape F
  size_t total = 0;
};

struct TableLayout {
  std::vector<size_t> offset, count;  // per ExoTable
  size_t total = 0;
};

// Fixed-orbit calendar (radmod orb_params with nfixorb = 1) and the date
// rules of calendar/date/index.ts: the 360-day calendar, or the Gregorian
// one for a 365-day year.
struct Calendar {
  int steps_per_day = 0;        // ntspd
  int64_t steps_per_year = 0;   // n_steps_per_year
  int output_window = 0;        // nafter
  double step_minutes = 0;      // mpstep
  int days_per_year = 0;        // n_days_per_year
  int days_per_month = 0;       // n_days_per_month
  double solar_day = 0;
  double rotation_speed = 0;    // rotspd
  double toffo3 = 0;
  int keplerian = 0;
  int synchronous = 0;
  double fixed_longitude = 0;
  int carbon_interval = 1;
  // orbit
  double eccentricity = 0, obliquity_rad = 0, mean_anomaly_rad = 0;
  double perihelion_lon_rad = 0, mean_lon_equinox_rad = 0;
};

struct Member {
  exo_config cfg{};
  ExoDims dims{};
  int nsnap = 0;
  Layout layout;      // the reference's arena layout (checkpoints, field offsets)
  Layout dev;         // the device arena's: every slot on a 256-byte boundary
  TableLayout tlayout;
  std::vector<double> tables;  // FP64 tables, bm1 of the first phase
  // Phases in run order: start-up kits = n..1, then the main loop (kits 0).
  std::vector<std::array<double, CON_COUNT>> constants;
  std::vector<std::vector<double>> bm1;
  int startup_steps = 0;
  std::vector<double> initial;  // initial arena (FP64, arena layout)
  Calendar cal;
  int64_t start_step = 0;
  std::vector<double> land_mask;  // for tests and diagnostics
  std::vector<double> qc_raw, qi_raw;  // unfiltered legini tables
};

void config_default(exo_config* cfg);
// Builds everything the device needs for one member; throws
// std::runtime_error on an invalid configuration.
// The surface fields (exo_add_member_surface) replace the fallbacks of
// their codes; with codes 172 and 129 the planet's land mask and
// orography come from them (landini, setup/surface/index.ts:293-297: land
// where the fraction exceeds 0.5) instead of the seeded terrain.
void build_member(const exo_config& cfg, Member& m, const exo_surface_field* fields = nullptr, int count = 0);
// Calendar date {year, month, day, hour, minute} of main-loop step `step`
// (calendar/date/index.ts:46-82: 360-day calendar, or Gregorian for a
// 365-day year).
void member_date(const Member& m, int64_t step, int out[5]);
// Gaussian latitudes in degrees, north first (GCM_GRID.forLatitudes'
// latitudesDeg, grid/index.ts:32-82).
std::vector<double> gaussian_degrees(int nlat);
Layout make_layout(const ExoDims& d, int nsnap);
// The device arena's layout: the slots of `ref` in the same order, each
// starting on a 256-byte boundary, so that a wave's 64 consecutive floats
// of a grid plane are two whole 128-byte cache lines (the reference's
// offsets leave most slots a few floats off a line boundary).
Layout device_layout(const Layout& ref);

// plasim_namelist's level arrays and nhdiff at a resolution, before the
// namelist is read (setup/pumamod/index.ts:43-91, :113-127)
struct PlasimArrays {
  int nhdiff;
  std::vector<double> ndel, tdissd, tdissz, tdisst, tdissq, restim, t0, tfrc;
};
PlasimArrays plasim_arrays(int truncation, int levels);
TableLayout make_table_layout(const ExoDims& d);
// STEPPER.values for model step `step` into 16 doubles: radiation_step[5],
// surface_step[7], carbon_step[1], snapshot_step[3]. carbon_step is 1 on
// every carbon-interval-th call counted from the run start.
void step_values(const Member& m, int64_t step, int64_t carbon_count,
                 int publish, double* out16);

// FP64 transforms used at set-up (FOURIER.analyse, LEGENDRE.analyse and
// their inverses) over the member's grid.
void grid_to_spectral(const Member& m, const double* grid, double* spectral,
                      bool filtered);
void spectral_to_grid(const Member& m, const double* spectral, double* grid);

}  // namespace exo
