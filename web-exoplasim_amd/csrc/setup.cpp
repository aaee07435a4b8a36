// FP64 host set-up of one ExoPlaSim member. Each routine cites the
// reference TypeScript it restates; the Fortran names are kept where the
// reference keeps them.
#include "setup.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>

namespace exo {
namespace {

// pumamod's PI (setup/pumamod/index.ts:9-10, dynamics/setup/index.ts:15).
const double PI = 3.14159265359;
const double TWOPI = PI + PI;
const double EZ = 1.63299310207;  // 1/sqrt(3/8)
const double RV = 461.51;
const double ACPV = 1870;
const double EARTH_RADIUS_M = 6371220;
const int CLIM_MONTHS = 14;
const int SOIL_LAYERS = 5;

// ---------------------------------------------------------------- layout

int count_of(int shape_count_L, int) { return shape_count_L; }

// ------------------------------------------------- Gregorian calendar
// calendar/date/index.ts:10-29: the 365-day year is the proleptic
// Gregorian calendar counted in days from the start of year 0.

const double MEAN_GREGORIAN_YEAR = 365.2425;

int64_t floor_div(int64_t a, int64_t b) { return a / b - ((a % b != 0) && ((a < 0) != (b < 0))); }

bool gregorian_leap(int64_t year) { return year % 4 == 0 && (year % 100 != 0 || year % 400 == 0); }

int64_t gregorian_year_start(int64_t year) {
  return 365 * year + floor_div(year + 3, 4) - floor_div(year + 99, 100) + floor_div(year + 399, 400);
}

int gregorian_month_days(int64_t year, int month) {
  static const int days[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  return days[month - 1] + (month == 2 && gregorian_leap(year) ? 1 : 0);
}

// CALENDAR.startStep (calendar/date/index.ts:31-44): the first step of
// (year, month).
int64_t start_step_of(int64_t mtspd, int year, int month, int days_per_year, int days_per_month) {
  if (days_per_year != 365) return mtspd * ((int64_t)year * days_per_year + (int64_t)(month - 1) * days_per_month);
  int64_t days = gregorian_year_start(year);
  for (int at = 1; at < month; at++) days += gregorian_month_days(year, at);
  return mtspd * days;
}

}  // namespace

Layout make_layout(const ExoDims& d, int nsnap) {
  // planeFloats (layout/index.ts:14-20)
  const size_t S = (size_t)exo_nrsp(d.ntru);
  const size_t F = (size_t)d.nlat * exo_nfc(d.ntru);
  const size_t G = (size_t)d.nlon * d.nlat;
  const size_t T = 1;
  const int L = d.nlev;
  const int NSNAP = nsnap;
  (void)NSNAP;
  Layout lay;
  lay.offset.resize(SLOT_COUNT);
  lay.floats.resize(SLOT_COUNT);
  lay.plane.resize(SLOT_COUNT);
  lay.fourier.resize(SLOT_COUNT);
  size_t at = 0;
#define EXO_LAYOUT_SLOT(name, shape, count)          \
  {                                                  \
    const size_t plane = shape;                      \
    const size_t n = plane * (size_t)(count);        \
    lay.offset[SLOT_##name] = at;                    \
    lay.floats[SLOT_##name] = n;                     \
    lay.plane[SLOT_##name] = plane;                  \
    lay.fourier[SLOT_##name] = #shape[0] == 'F';     \
    at += n;                                         \
  }
  EXO_FIELDS(EXO_LAYOUT_SLOT)
#undef EXO_LAYOUT_SLOT
  lay.total = at;
  return lay;
}

Layout device_layout(const Layout& ref) {
  Layout d = ref;
  size_t at = 0;
  for (int s = 0; s < SLOT_COUNT; s++) {
    at = (at + 63) & ~(size_t)63;
    d.offset[s] = at;
    at += ref.floats[s];
  }
  d.total = (at + 63) & ~(size_t)63;
  return d;
}

TableLayout make_table_layout(const ExoDims& d) {
  const int T = d.ntru, NLAT = d.nlat, NLON = d.nlon, L = d.nlev;
  const int NCSP = exo_ncsp(T), NRSP = exo_nrsp(T), NFC = exo_nfc(T);
  const int NHOR = NLAT * NLON;
  TableLayout t;
  t.offset.resize(TAB_COUNT);
  t.count.resize(TAB_COUNT);
  size_t at = 0;
#define EXO_TABLE_SLOT(name, n)      \
  {                                  \
    t.offset[TAB_##name] = at;       \
    t.count[TAB_##name] = (size_t)(n); \
    at += (size_t)(n);               \
  }
  EXO_TABLES(EXO_TABLE_SLOT)
#undef EXO_TABLE_SLOT
  t.total = at;
  (void)NHOR;
  return t;
}

namespace {

// ----------------------------------------------------------- pumamod

struct Pumamod {
  // planet (p_earth.f90 defaults, setup/pumamod/index.ts:18-40)
  double eccen = 0.016715, mvelp = 102.7, obliq = 23.441, rotspd = 1;
  double sidereal_year = 31558149, meananom0 = -2.9589, akap = 0.286;
  double alr = 0.0065, gascon = 287, ra1 = 610.78, ra2 = 17.2693882;
  double ra4 = 35.86, pnu = 0.1, ga = 9.80665, plarad = 6371220;
  double gsol0 = 1365;
  int nfixorb = 0, ngenkeplerian = 0;
  // derived
  double day_24hr = 86400, solar_day = 86400, sidereal_day = 86400;
  double ww = 0, acpd = 0, adv = 0, cv = 0, ct = 0, pnu21 = 0, rdbrv = 0;
  double tmelt = 273.16, alv = 2.5008e6, als = 2.8345e6;
  double psurf = 101100, plavor = EZ, mpstep = 0;
  int mtspd = 0, ntspd = 0, nafter = 0;
  int64_t n_steps_per_year = 0;
  int n_days_per_year = 360, n_start_year = 1, n_start_month = 1;
  int64_t n_start_step = 0;
  int n_days_per_month = 30, m_days_per_year = 360;
  int nkits = 3, kick = 1, ndheat = 1, nveg = 1, nqspec = 1;
  int nhdiff = 15, nhordif = 1, neqsig = 0;
  int nfilter = 0, ngptfilter = 0, nspvfilter = 0;
  double filterkappa = 8, nfilterexp = 8, landhoskn0 = 15;
  double dtep = 0, dtns = 0, dtrop = 12000, dttrp = 2, dttl = 0;
  double fixedlon = 0, tgr = 288, ptop = 7500;
  std::vector<double> ndel, tdissd, tdissz, tdisst, tdissq, restim, t0, tfrc;
};

int latitudes_for(int truncation) {
  if (truncation == 21) return 32;
  if (truncation == 42) return 64;
  if (truncation == 63) return 96;
  throw std::runtime_error("T" + std::to_string(truncation) + " is not built");
}

}  // namespace

PlasimArrays plasim_arrays(int truncation, int levels) {
  // plasimmod's level-array defaults and readnl's resolution defaults
  // before the namelist (setup/pumamod/index.ts:43-91, :113-127)
  const double day = 86400, frcmod = 1;
  const int L = levels;
  PlasimArrays a;
  a.nhdiff = 15;
  a.ndel.assign(L, 2);
  a.tdissd.assign(L, 0.2);
  a.tdissz.assign(L, 1.1);
  a.tdisst.assign(L, 5.6);
  a.tdissq.assign(L, 0.1);
  a.restim.assign(L, 0);
  a.t0.assign(L, 250);
  a.tfrc.assign(L, 0);
  if (L == 10)
    for (int l = 0; l < L; l++) a.tfrc[l] = (l == 0 ? 20 : l == 1 ? 100 : 0) * day * frcmod;
  if (truncation == 42) {
    a.nhdiff = 16;
    for (int l = 0; l < L; l++) {
      a.ndel[l] = 4;
      a.tdissq[l] = 0.1 * day;
      a.tdisst[l] = 0.76 * day;
      a.tdissz[l] = 0.3 * day;
      a.tdissd[l] = 0.06 * day;
    }
  }
  return a;
}

namespace {

// GCM_NAMELIST.edits (namelist/index.ts:40-182) applied over the T21 base
// namelists, then PUMAMOD.read (setup/pumamod/index.ts:103-232).
Pumamod pumamod_read(const exo_config& c) {
  Pumamod p;
  const int T = c.truncation, L = c.levels;
  // planet_namelist base: ECCEN 0, MVELP 102.7, OBLIQ 0, GSOL0 1367,
  // NFIXORB 1; then the edits.
  p.ga = c.gravity;
  p.plarad = c.radius_earth * EARTH_RADIUS_M;
  p.gsol0 = c.flux_wm2;
  p.eccen = c.eccentricity;
  p.obliq = c.obliquity_deg;
  p.nfixorb = c.fixed_orbit ? 1 : 0;
  p.ngenkeplerian = c.keplerian ? 1 : 0;
  p.gascon = c.gas_constant;
  p.sidereal_year = c.year_days * 86400.0;
  p.mvelp = 102.7;
  // the orbital-angle edits (namelist/index.ts:57-75)
  for (double a : {c.periapse_longitude_deg, c.initial_mean_anomaly_deg})
    if (!std::isnan(a) && (!std::isfinite(a) || a < -360 || a > 360))
      throw std::runtime_error("Orbital angles must be finite and between -360 and 360 degrees");
  if (!std::isnan(c.periapse_longitude_deg)) {
    if (c.periapse_longitude_deg < 0) throw std::runtime_error("Periapse longitude must be between 0 and 360 degrees");
    p.mvelp = c.periapse_longitude_deg;
  }
  if (!std::isnan(c.initial_mean_anomaly_deg)) p.meananom0 = c.initial_mean_anomaly_deg;
  if (!c.fixed_orbit) throw std::runtime_error("fixedOrbit false (Berger's orbital series) is not built");
  int ndpy = (int)std::lround(c.year_days);
  if (c.rotation_period_days != 1) {
    p.rotspd = 1.0 / c.rotation_period_days;
    ndpy = std::max((int)std::floor(c.year_days / c.rotation_period_days / 12 + 0.5), 1) * 12;
  }
  // plasim_namelist: readnl defaults, resolution defaults, base + edits.
  const double day = 86400;
  {
    PlasimArrays a = plasim_arrays(T, L);
    p.nhdiff = a.nhdiff;
    p.ndel = a.ndel;
    p.tdissd = a.tdissd;
    p.tdissz = a.tdissz;
    p.tdisst = a.tdisst;
    p.tdissq = a.tdissq;
    p.restim = a.restim;
    p.t0 = a.t0;
    p.tfrc = a.tfrc;
  }
  p.kick = 1;
  p.nqspec = 1;
  if (c.vegetation < 0 || c.vegetation > 2) throw std::runtime_error("vegetation must be 0, 1 or 2");
  p.nveg = c.vegetation;
  p.neqsig = 0;
  if (c.filter < 0 || c.filter > 3) throw std::runtime_error("physicsFilter: NFILTER must be 0 to 3");
  p.nfilter = c.filter;  // the last of none / cesaro / exp / lh in physicsFilter
  p.ngptfilter = 1;      // the base namelist's, which "gp" only repeats
  p.nspvfilter = 1;      // likewise "sp"
  p.filterkappa = 8;
  p.nfilterexp = 8;
  p.landhoskn0 = 15;
  p.psurf = c.pressure_bar * 1e5;
  p.fixedlon = c.substellar_longitude_deg;
  p.n_days_per_year = ndpy;
  p.n_start_year = ndpy == 365 ? 1 : 0;
  p.n_start_month = 1;
  double mpstep = c.timestep_minutes;
  int nstpw = (int)std::floor(7200.0 / c.timestep_minutes);
  int nlowio = 1, nwpd = 0;

  double solar_day = day;
  double sidereal_day = solar_day * (ndpy - 1) / ndpy;
  if (p.rotspd != 1) {
    sidereal_day = day / p.rotspd;
    solar_day = ndpy != 1 ? sidereal_day * ndpy / (ndpy - 1) : sidereal_day;
  }
  const int m_days = (int)std::lround(ndpy * sidereal_day / day);
  p.ww = TWOPI / sidereal_day;
  p.acpd = p.gascon / p.akap;
  p.cv = p.plarad * p.ww;
  if (mpstep <= 0) {
    const int lat = latitudes_for(T);
    mpstep = lat <= 32 ? 45 : lat <= 48 ? 36 : std::trunc(30.0 * 64 / lat);
  }
  int mtspd = (int)std::trunc(std::round(day) / std::round(mpstep * 60));
  mtspd = mtspd + (mtspd % 2);
  mpstep = day / (mtspd * 60.0);
  const int ntspd = (int)std::trunc(std::round(solar_day) / std::round(mpstep * 60));
  int nafter = mtspd;
  if (nwpd > 0 && nwpd <= mtspd) nafter = mtspd / nwpd;
  if (nlowio > 0 && nstpw > 0) nafter = nstpw;
  p.day_24hr = day;
  p.solar_day = solar_day;
  p.sidereal_day = sidereal_day;
  p.adv = ACPV / p.acpd - 1;
  p.ct = p.cv * p.cv / p.gascon;
  p.pnu21 = 1 - 2 * p.pnu;
  p.rdbrv = p.gascon / RV;
  p.plavor = EZ;
  p.mpstep = mpstep;
  p.mtspd = mtspd;
  p.ntspd = ntspd;
  p.nafter = nafter;
  p.n_steps_per_year = std::llround(p.sidereal_year / (mpstep * 60));
  p.n_days_per_month = m_days / 12;
  p.m_days_per_year = m_days;
  // CALENDAR.startStep (calendar/date/index.ts:31-44)
  p.n_start_step = start_step_of(mtspd, p.n_start_year, p.n_start_month, ndpy, p.n_days_per_month);
  return p;
}

// ------------------------------------------------------------- grid

struct Grid {
  int nlat = 0, nlon = 0;
  std::vector<double> sines, weights, degrees;
};

// ql in gaussmod.f90's inigau (grid/index.ts:15-30)
double inigau_legendre(int k, double p) {
  const double angle = std::acos(p);
  double factor = 1, sum = 0, term = 0;
  for (int j = k; j >= 0; j -= 2) {
    term = factor * std::cos(angle * j);
    sum = sum + term;
    const double z4 = (k - j + 1) * (k + j) * 0.5;
    factor = factor * z4 / (z4 + (j - 1));
  }
  if (k % 2 == 0) sum = sum - 0.5 * term;
  double norm = std::sqrt(2.0);
  for (int j = 1; j <= k; j++) norm = norm * std::sqrt(1 - 0.25 / ((double)j * j));
  return norm * sum;
}

// inigau (grid/index.ts:32-63): Newton on the trigonometric series.
Grid gaussian_grid(int count) {
  Grid g;
  g.nlat = count;
  g.nlon = 2 * count;
  g.sines.assign(count, 0);
  g.weights.assign(count, 0);
  g.degrees.assign(count, 0);
  const double INIGAU_PI = 3.14159265358979;
  const double z0 = INIGAU_PI / (2 * count + 1);
  const double z1 = 1.0 / ((double)count * count * 8);
  const double z4 = 2.0 / ((double)count * count);
  for (int root = 0; root < count / 2; root++) {
    const double guess = z0 * (2 * (root + 1) - 0.5);
    double x = std::cos(guess + z1 / std::tan(guess));
    for (int it = 0; it < 50; it++) {
      const double value = inigau_legendre(count, x);
      const double z = x * value - std::sqrt((double)(count + count + 1) / (count + count - 1)) *
                                       inigau_legendre(count - 1, x);
      const double step = value * ((x * x - 1) / (count * z));
      x = x - step;
      if (std::fabs(step) < 1e-16) break;
    }
    const double z5 = inigau_legendre(count - 1, x) / std::sqrt(count - 0.5);
    const double lat = std::asin(x) * 180 / M_PI;
    const double w = z4 * (1 - x * x) / (z5 * z5);
    g.degrees[root] = lat;
    g.degrees[count - 1 - root] = -lat;
    g.sines[root] = x;
    g.sines[count - 1 - root] = -x;
    g.weights[root] = w;
    g.weights[count - 1 - root] = w;
  }
  return g;
}

// ----------------------------------------------------------- legendre

struct Legendre {
  int modes = 0;
  std::vector<double> qi, qj, qc, qe, qm, qq, qu, qv;  // [lat*NCSP + w]
};

// legini (legendre/index.ts:29-108): normalised associated Legendre
// functions by recurrence, with pli/pld 1-based as in Fortran.
Legendre legini(const Grid& grid, int T) {
  Legendre r;
  const int modes = exo_ncsp(T);
  r.modes = modes;
  const size_t size = (size_t)modes * grid.nlat;
  for (auto* v : {&r.qi, &r.qj, &r.qc, &r.qe, &r.qm, &r.qq, &r.qu, &r.qv}) v->assign(size, 0.0);
  std::vector<double> pli(modes + 1), pld(modes + 1);
  for (int lat = 0; lat < grid.nlat; lat++) {
    const double sin = grid.sines[lat];
    const double weight = grid.weights[lat];
    const double cos2 = 1 - sin * sin;
    const double wocs = weight / cos2;
    std::fill(pli.begin(), pli.end(), 0.0);
    std::fill(pld.begin(), pld.end(), 0.0);
    double f1m = std::sqrt(1.5);
    pli[1] = std::sqrt(0.5);
    pli[2] = f1m * sin;
    pld[1] = 0;
    int lm = 2;
    for (int m = 0; m <= T; m++) {
      if (m > 0) {
        lm++;
        const double f2m = -f1m * std::sqrt(cos2 / (m + m));
        f1m = f2m * std::sqrt((double)(m + m + 3));
        pli[lm] = f2m;
        if (lm < modes) {
          lm++;
          pli[lm] = f1m * sin;
          pld[lm - 1] = -m * f2m * sin;
        }
      }
      const double m2 = (double)m * m;
      for (int n = m + 2; n <= T; n++) {
        lm++;
        const double z1 = std::sqrt(((n - 1.0) * (n - 1.0) - m2) / (4.0 * (n - 1) * (n - 1) - 1));
        const double z2 = sin * pli[lm - 1] - z1 * pli[lm - 2];
        pli[lm] = z2 * std::sqrt((4.0 * n * n - 1) / ((double)n * n - m2));
        pld[lm - 1] = (1 - n) * z2 + n * z1 * pli[lm - 2];
      }
      if (lm < modes) {
        const double z3 = std::sqrt(((double)T * T - m2) / (4.0 * T * T - 1));
        pld[lm] = -T * sin * pli[lm] + (T + T + 1) * pli[lm - 1] * z3;
      } else {
        pld[lm] = -T * sin * pli[lm];
      }
    }
    int w = 0;
    for (int m = 0; m <= T; m++)
      for (int n = m; n <= T; n++) {
        const double inv = n > 0 ? 1.0 / ((double)n * (n + 1)) : 0;
        const size_t at = (size_t)lat * modes + w;
        const double p = pli[w + 1], d = pld[w + 1];
        r.qi[at] = p;
        r.qj[at] = d;
        r.qc[at] = p * weight;
        r.qu[at] = p * inv * m;
        r.qv[at] = d * inv;
        r.qe[at] = d * wocs;
        r.qq[at] = p * wocs * n * (n + 1) * 0.5;
        r.qm[at] = p * wocs * m;
        w++;
      }
  }
  return r;
}

// LEGENDRE.filters (legendre/index.ts:111-136): skgpsp and skspgp per
// total wavenumber.
void physics_filters(const Pumamod& p, int T, std::vector<double>& g2s, std::vector<double>& s2g) {
  g2s.assign(T + 1, 1);
  s2g.assign(T + 1, 1);
  for (int deg = 0; deg <= T; deg++) {
    const double n = deg + 1;
    double shape = 1;
    switch (p.nfilter) {
      case 0: shape = 1; break;
      case 1: shape = 1 - n / (T + 1); break;
      case 2: shape = std::exp(-p.filterkappa * std::pow(n / T, p.nfilterexp)); break;
      case 3: shape = std::exp(-std::pow(n * (n + 1) / (p.landhoskn0 * (p.landhoskn0 + 1)), 2)); break;
      case 4: shape = std::pow(1 - n / (T + 1), 2); break;
    }
    g2s[deg] = 1 - p.ngptfilter + p.ngptfilter * shape;
    s2g[deg] = 1 - p.nspvfilter + p.nspvfilter * shape;
  }
}

// ----------------------------------------------------- dynamics set-up

struct Dyn {
  std::vector<double> sigmah, sigma, dsigma, rdsig, damp, tfrc, tdissd, tdissz, tdisst, tdissq;
  std::vector<double> sak, nindex, spnorm, t0, tkp, t01s2, g, c, tau;
  double dtns = 0, dtep = 0, dttrp = 0;
};

// dayseccheck + the rate conversion of initpm (dynamics/setup/index.ts:
// 18-24, :40-45).
std::vector<double> rates(const std::vector<double>& values, double threshold, double day) {
  double largest = -1e300;
  for (double v : values) largest = std::max(largest, v);
  std::vector<double> secs = values;
  if (largest < threshold && largest > 0)
    for (auto& v : secs) v *= day;
  std::vector<double> out(values.size());
  for (size_t i = 0; i < values.size(); i++) out[i] = secs[i] > 0 ? day / (TWOPI * secs[i]) : 0;
  return out;
}

// initsi (dynamics/setup/index.ts:95-136); matrices column-major.
void initsi(Dyn& s, double akap) {
  const int n = (int)s.t0.size();
  auto at = [n](int i, int j) { return j * n + i; };
  s.tkp.assign(n, 0);
  s.t01s2.assign(n, 0);
  s.g.assign(n * n, 0);
  s.c.assign(n * n, 0);
  s.tau.assign(n * n, 0);
  for (int l = 0; l < n; l++) s.tkp[l] = akap * s.t0[l];
  for (int l = 0; l < n - 1; l++) s.t01s2[l] = s.t0[l + 1] - s.t0[l];
  s.t01s2[n - 1] = 0;
  std::vector<double> alp(n, 0);
  for (int l = 1; l < n; l++) alp[l] = std::log(s.sigmah[l]) - std::log(s.sigmah[l - 1]);
  s.g[at(0, 0)] = 1;
  for (int l = 1; l < n; l++) {
    s.g[at(l, l)] = 1 - alp[l] * s.sigmah[l - 1] / s.dsigma[l];
    for (int k = 0; k < l; k++) s.g[at(l, k)] = alp[l];
  }
  for (int l = 0; l < n; l++)
    for (int k = 0; k < n; k++) s.c[at(l, k)] = s.g[at(k, l)] * (s.dsigma[l] / s.dsigma[k]);
  double t01 = s.t01s2[0];
  double sig = s.sigmah[0];
  s.tau[at(0, 0)] = 0.5 * t01 * (sig - 1) + s.tkp[0];
  for (int k = 1; k < n; k++) s.tau[at(k, 0)] = 0.5 * t01 * s.dsigma[k];
  for (int l = 1; l < n; l++) {
    const double pt01 = t01, psig = sig;
    t01 = s.t01s2[l];
    sig = s.sigmah[l];
    for (int k = 0; k < n; k++) {
      const double below = k <= l ? 1 : 0;
      const double above = k < l ? 1 : 0;
      double value = pt01 * (psig - above);
      if (l < n - 1) value = value + t01 * (sig - below);
      value = value * s.rdsig[l] * s.dsigma[k];
      if (k <= l) value = value + s.tkp[l] * s.c[at(k, l)];
      s.tau[at(k, l)] = value;
    }
  }
}

// initpm for NEQSIG = 0 (dynamics/setup/index.ts:27-93).
Dyn initpm(const Pumamod& p, int T, int L) {
  Dyn s;
  s.sigmah.resize(L);
  for (int l = 0; l < L; l++) {
    const double z = (l + 1.0) / L;
    s.sigmah[l] = 0.75 * z + 1.75 * z * z * z - 1.5 * z * z * z * z;
  }
  s.dsigma.resize(L);
  s.rdsig.resize(L);
  s.sigma.resize(L);
  for (int l = 0; l < L; l++) {
    s.dsigma[l] = l == 0 ? s.sigmah[l] : s.sigmah[l] - s.sigmah[l - 1];
    s.rdsig[l] = 0.5 / s.dsigma[l];
    s.sigma[l] = l == 0 ? 0.5 * s.sigmah[l] : 0.5 * (s.sigmah[l - 1] + s.sigmah[l]);
  }
  const double threshold = p.day_24hr / p.mtspd;
  auto zeroed = [&](const std::vector<double>& v) {
    return p.nhordif != 0 ? v : std::vector<double>(v.size(), 0.0);
  };
  s.damp = rates(zeroed(p.restim), threshold, p.day_24hr);
  s.tfrc = rates(zeroed(p.tfrc), threshold, p.day_24hr);
  s.tdissd = rates(zeroed(p.tdissd), threshold, p.day_24hr);
  s.tdissz = rates(zeroed(p.tdissz), threshold, p.day_24hr);
  s.tdisst = rates(zeroed(p.tdisst), threshold, p.day_24hr);
  s.tdissq = rates(zeroed(p.tdissq), threshold, p.day_24hr);
  const int coeffs = exo_nrsp(T);
  s.sak.assign((size_t)coeffs * L, 0);
  for (int l = 0; l < L; l++) {
    const double power = p.ndel[l];
    const double scale = 1.0 / std::pow((double)(T - p.nhdiff), power);
    int r = 0;
    for (int m = 0; m <= T; m++)
      for (int n = m; n <= T; n++) {
        const double v = n >= p.nhdiff ? scale * std::pow((double)(n - p.nhdiff), power) : 0;
        s.sak[(size_t)l * coeffs + r] = v;
        s.sak[(size_t)l * coeffs + r + 1] = v;
        r += 2;
      }
  }
  s.nindex.assign(coeffs, 0);
  s.spnorm.assign(coeffs, 0);
  int r = 0;
  double norm = 1 / std::sqrt(2.0);
  for (int m = 0; m <= T; m++) {
    for (int n = m; n <= T; n++) {
      s.nindex[r] = s.nindex[r + 1] = n;
      s.spnorm[r] = s.spnorm[r + 1] = norm;
      r += 2;
    }
    norm = -norm;
  }
  s.t0.resize(L);
  for (int l = 0; l < L; l++) s.t0[l] = p.t0[l] / p.ct;
  s.dtns = p.dtns / p.ct;
  s.dtep = p.dtep / p.ct;
  s.dttrp = p.dttrp / p.ct;
  initsi(s, p.akap);
  return s;
}

// ludcmp + lubksb (Numerical Recipes, partial pivoting) as minvers uses
// them (dynamics/setup/index.ts:140-205); column-major, in place.
void invert(std::vector<double>& a, int n) {
  auto at = [n](int i, int j) { return j * n + i; };
  std::vector<double> scale(n);
  for (int i = 0; i < n; i++) {
    double largest = 0;
    for (int j = 0; j < n; j++) largest = std::max(largest, std::fabs(a[at(i, j)]));
    scale[i] = 1 / largest;
  }
  std::vector<int> index(n);
  for (int j = 0; j < n; j++) {
    for (int i = 1; i < j; i++) {
      double dot = 0;
      for (int k = 0; k < i; k++) dot += a[at(i, k)] * a[at(k, j)];
      a[at(i, j)] -= dot;
    }
    double largest = 0;
    int pivot = j;
    for (int i = j; i < n; i++) {
      if (j > 0) {
        double dot = 0;
        for (int k = 0; k < j; k++) dot += a[at(i, k)] * a[at(k, j)];
        a[at(i, j)] -= dot;
      }
      const double size = scale[i] * std::fabs(a[at(i, j)]);
      if (size >= largest) {
        pivot = i;
        largest = size;
      }
    }
    if (j != pivot) {
      for (int k = 0; k < n; k++) std::swap(a[at(pivot, k)], a[at(j, k)]);
      scale[pivot] = scale[j];
    }
    index[j] = pivot;
    if (a[at(j, j)] == 0) a[at(j, j)] = 5e-324;
    for (int i = j + 1; i < n; i++) a[at(i, j)] /= a[at(j, j)];
  }
  std::vector<double> inv((size_t)n * n);
  std::vector<double> b(n);
  for (int col = 0; col < n; col++) {
    for (int i = 0; i < n; i++) b[i] = i == col ? 1 : 0;
    int first = -1;
    for (int i = 0; i < n; i++) {
      const int l = index[i];
      double sum = b[l];
      b[l] = b[i];
      if (first >= 0) {
        for (int j = first; j < i; j++) sum -= a[at(i, j)] * b[j];
      } else if (sum != 0) {
        first = i;
      }
      b[i] = sum;
    }
    for (int i = n - 1; i >= 0; i--) {
      double sum = b[i];
      for (int j = i + 1; j < n; j++) sum -= a[at(i, j)] * b[j];
      b[i] = sum / a[at(i, i)];
    }
    for (int i = 0; i < n; i++) inv[(size_t)col * n + i] = b[i];
  }
  a = inv;
}

// makebm (dynamics/setup/index.ts:207-232).
std::vector<double> makebm(const Dyn& s, int T, double delt) {
  const int n = (int)s.t0.size();
  const double d2 = delt * delt;
  std::vector<double> base((size_t)n * n);
  for (int l1 = 0; l1 < n; l1++)
    for (int l2 = 0; l2 < n; l2++) {
      double dot = 0;
      for (int k = 0; k < n; k++) dot += s.g[l1 * n + k] * s.tau[k * n + l2];
      base[l1 * n + l2] = d2 * (s.t0[l1] * s.dsigma[l2] + dot);
    }
  std::vector<double> bm1((size_t)n * n * T);
  for (int deg = 1; deg <= T; deg++) {
    std::vector<double> m = base;
    for (int l = 0; l < n; l++) m[l * n + l] += 1.0 / ((double)deg * (deg + 1));
    invert(m, n);
    std::copy(m.begin(), m.end(), bm1.begin() + (size_t)(deg - 1) * n * n);
  }
  return bm1;
}

// setzt (dynamics/setup/index.ts:252-302).
std::vector<double> setzt(const Dyn& s, const Pumamod& p, int T) {
  const int L = (int)s.sigma.size();
  const double dttrp = s.dttrp * p.ct;
  const double alr = p.alr, tgr = p.tgr, dtrop = p.dtrop, gascon = p.gascon, ga = p.ga;
  auto profile = [&](double h) {
    return tgr - dtrop * alr + std::sqrt(std::pow(0.5 * alr * (h - dtrop), 2) + dttrp * dttrp) -
           0.5 * alr * (h - dtrop);
  };
  std::vector<double> trs(L);
  double sp = 1, tp = tgr, zp = 0;
  for (int l = L - 1; l >= 0; l--) {
    const double lr = std::log(sp / s.sigma[l]);
    const double first = profile(zp + gascon * tp / ga * lr);
    const double t = profile(zp + gascon * (0.5 * (tp + first)) / ga * lr);
    trs[l] = t;
    zp = zp + 0.5 * (t + tp) * gascon / ga * lr;
    tp = t;
    sp = s.sigma[l];
  }
  const double trop = std::pow((tgr - dtrop * alr) / tgr, ga / (alr * gascon));
  const double substellar = 1 / 3.463843223261725 / p.ct * p.dttl;
  const double subst1 = -substellar * std::cos(p.fixedlon * PI / 180);
  const double subst2 = substellar * std::sin(p.fixedlon * PI / 180);
  const int coeffs = exo_nrsp(T);
  const int sas = 2 * (T + 1);
  std::vector<double> sr((size_t)coeffs * L, 0);
  for (int l = 0; l < L; l++) {
    const double fac = std::max(std::sin(0.5 * PI * (s.sigma[l] - trop) / (1 - trop)), 0.0);
    const size_t at = (size_t)l * coeffs;
    sr[at] = std::sqrt(2.0) * (trs[l] / p.ct - s.t0[l]);
    sr[at + 2] = 1 / std::sqrt(6.0) * s.dtns * fac;
    sr[at + 4] = -2.0 / 3.0 * std::sqrt(0.4) * s.dtep * fac;
    sr[at + sas] = subst1 * fac;
    sr[at + sas + 1] = subst2 * fac;
  }
  return sr;
}

// ------------------------------------------------------------ random

// xoshiro256** (Blackman and Vigna) seeded through splitmix64: the
// synthetic planet's terrain and noise generator.
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    for (auto& w : s) w = splitmix(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return (next() >> 11) * 0x1.0p-53; }
  double normal() {
    double u1 = uniform();
    while (u1 <= 0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2 * std::log(u1)) * std::cos(2 * M_PI * u2);
  }
};

// libgfortran's random_seed(put=) and random_number for real(8), as the
// reference restates them (shared/random/gfortran/index.ts:8-51): the
// eight 32-bit seed words are taken in reverse order, paired into four
// little-endian 64-bit words and XORed with libgfortran's scrambling keys;
// each number is the top 53 bits of xoshiro256**'s next word times 2^-53.
struct GfortranRandom {
  uint64_t s[4];
  explicit GfortranRandom(const uint32_t seed[8]) {
    static const uint64_t keys[4] = {0xbd0c5b6e50c2df49ull, 0xd46061cd46e1df38ull,
                                     0xbb4f4d4ed6103544ull, 0x114a583d0756ad39ull};
    for (int i = 0; i < 4; i++)
      s[i] = ((uint64_t)seed[7 - 2 * i] | ((uint64_t)seed[6 - 2 * i] << 32)) ^ keys[i];
  }
  double next() {
    const uint64_t result = Rng::rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = Rng::rotl(s[3], 45);
    return (double)(result >> 11) * 0x1.0p-53;
  }
};

// initrandom (setup/state/index.ts:39-46): the namelist's SEED, or words
// from the clock, (clock + 37 i) truncated to 32 bits, when its first word
// is zero.
std::array<uint32_t, 8> initrandom_seed(const exo_config& cfg) {
  std::array<uint32_t, 8> words{};
  for (int i = 0; i < 8; i++)
    words[i] = cfg.seed[0] != 0 ? (uint32_t)cfg.seed[i] : (uint32_t)(uint64_t)(cfg.clock_ms + 37 * (i + 1));
  return words;
}

// noise (setup/state/index.ts:12-35): white noise of amplitude 1e-4 in
// ln(ps). Kick 1 perturbs every non-zonal coefficient in storage order,
// kick 2 the real and imaginary parts of the modes with n + m even (one
// draw for both), kick 3 sets mode (1, 1) alone.
void initfd_noise(double* sp, int T, int kick, const std::array<uint32_t, 8>& seed) {
  GfortranRandom random(seed.data());
  const double scale = 1e-4 / std::sqrt(2.0);
  const int ntp1 = T + 1, nrsp = (T + 1) * (T + 2);
  if (kick == 1) {
    for (int i = 2 * ntp1; i < nrsp; i++) sp[i] += scale * (random.next() - 0.5);
  } else if (kick == 2) {
    int r = 2 * ntp1 - 2;
    for (int m = 1; m <= T; m++)
      for (int n = m; n <= T; n++) {
        r += 2;
        if ((n + m) % 2 != 0) continue;
        const double value = scale * (random.next() - 0.5);
        sp[r] += value;
        sp[r + 1] += value;
      }
  } else if (kick == 3) {
    sp[2 * ntp1 + 2] = scale;
    sp[2 * ntp1 + 3] = scale * 0.5;
  }
}

}  // namespace

// ------------------------------------------------- set-up transforms

// FOURIER.analyse then LEGENDRE.analyse (fourier/index.ts:97-114,
// legendre/index.ts:334-361) in FP64. `filtered` applies skgpsp as the
// reference's set-up does for the orography; otherwise the plain
// quadrature projection.
void grid_to_spectral(const Member& m, const double* grid, double* sp, bool filtered) {
  const int T = m.dims.ntru, NLAT = m.dims.nlat, NLON = m.dims.nlon;
  const int NFC = exo_nfc(T), modes = exo_ncsp(T);
  const double* fwd = &m.tables[m.tlayout.offset[TAB_fourier_forward]];
  std::vector<double> four((size_t)NLAT * NFC);
  for (int row = 0; row < NLAT; row++)
    for (int c = 0; c < NFC; c++) {
      double sum = 0;
      for (int j = 0; j < NLON; j++) sum += grid[row * NLON + j] * fwd[j * NFC + c];
      four[row * NFC + c] = sum;
    }
  const double* qc = filtered ? &m.tables[m.tlayout.offset[TAB_legendre_qc]] : m.qc_raw.data();
  for (int i = 0; i < 2 * modes; i++) sp[i] = 0;
  for (int l = 0; l < NLAT / 2; l++) {
    const int north = l * NFC, south = (NLAT - 1 - l) * NFC;
    int w = 0;
    for (int mm = 0; mm <= T; mm++)
      for (int n = mm; n <= T; n++, w++) {
        const double sign = (mm + n) % 2 == 0 ? 1 : -1;
        const double weight = qc[(size_t)l * modes + w];
        for (int part = 0; part < 2; part++) {
          const double row = four[north + 2 * mm + part] + sign * four[south + 2 * mm + part];
          sp[2 * w + part] = sp[2 * w + part] + weight * row;
        }
      }
  }
}

// sp2fc with the unfiltered qi, then fc2gp (legendre/wgsl/inverse.wgsl,
// fourier/wgsl/inverse.wgsl) in FP64.
void spectral_to_grid(const Member& m, const double* sp, double* grid) {
  const int T = m.dims.ntru, NLAT = m.dims.nlat, NLON = m.dims.nlon;
  const int NFC = exo_nfc(T), modes = exo_ncsp(T);
  const double* qi = m.qi_raw.data();
  const double* inv = &m.tables[m.tlayout.offset[TAB_fourier_inverse]];
  std::vector<double> four(NFC);
  for (int lat = 0; lat < NLAT; lat++) {
    for (int mm = 0; mm <= T; mm++) {
      const int first = mm * (2 * (T + 1) - mm + 1) / 2;
      double re = 0, im = 0;
      for (int w = first; w < first + T + 1 - mm; w++) {
        const double q = qi[(size_t)lat * modes + w];
        re += q * sp[2 * w];
        im += q * sp[2 * w + 1];
      }
      four[2 * mm] = re;
      four[2 * mm + 1] = im;
    }
    for (int j = 0; j < NLON; j++) {
      double sum = 0;
      for (int k = 0; k < NFC; k++) sum += inv[(size_t)k * NLON + j] * four[k];
      grid[lat * NLON + j] = sum;
    }
  }
}

namespace {

// Saturation specific humidity at the namelist psurf (setup/surface/
// index.ts:187-196).
double saturation(double t, const Pumamod& p) {
  const double q = p.rdbrv * p.ra1 * std::exp(p.ra2 * (t - p.tmelt) / (t - p.ra4)) / p.psurf;
  return q / (1 - (1 / p.rdbrv - 1) * q);
}

struct Date {
  int year, month, day, hour, minute;
};
// calendar/date/index.ts:46-82. A 365-day year runs the proleptic
// Gregorian calendar from day 0 of year 0; any other year length the
// model's 360-day calendar of twelve 30-day months.
Date date_of(int64_t step, int steps_per_day, int days_per_year, double step_minutes) {
  if (days_per_year == 365) {
    const int64_t days = step / steps_per_day;
    int64_t year = (int64_t)std::trunc((double)days / MEAN_GREGORIAN_YEAR);
    while (gregorian_year_start(year) > days) year--;
    while (gregorian_year_start(year + 1) <= days) year++;
    int64_t day = days - gregorian_year_start(year) + 1;
    int month = 1;
    while (day > gregorian_month_days(year, month)) day -= gregorian_month_days(year, month++);
    const int64_t minutes = (step % steps_per_day) * 1440 / steps_per_day;
    return {(int)year, month, (int)day, (int)(minutes / 60), (int)(minutes % 60)};
  }
  const int64_t mtspd = (int64_t)days_per_year * steps_per_day / 360;
  const int64_t days = step / mtspd;
  const int64_t doy = days % 360;
  const int64_t minutes = (int64_t)((step % mtspd) * std::llround(step_minutes));
  return {(int)(days / 360 + 1), (int)(doy / 30 + 1), (int)(doy % 30 + 1), (int)(minutes / 60),
          (int)(minutes % 60)};
}

struct MonthInterp {
  int month, neighbour;
  double weight;
};
MonthInterp month_interp(int64_t step, const Calendar& c) {
  const Date d = date_of(step, c.steps_per_day, c.days_per_year, c.step_minutes);
  const double fday = d.day + ((d.hour * 60 + d.minute) * 60) / c.solar_day;
  const bool gregorian = c.days_per_year == 365;
  const double cur = gregorian ? gregorian_month_days(d.year, d.month) : c.days_per_month;
  const double median = 0.5 * (cur + 1);
  const int neighbour = fday > median ? d.month + 1 : d.month - 1;
  const double adj = gregorian ? gregorian_month_days(d.year, (neighbour + 11) % 12 + 1) : c.days_per_month;
  return {d.month, neighbour, std::fabs(fday - median) / (median + 0.5 * (adj + 1) - 1)};
}

// momint linear interpolation of a 14-month climatology (setup/surface/
// index.ts:106-130).
std::vector<double> monthly(const std::vector<double>& table, int64_t step, const Calendar& c, int points) {
  const MonthInterp mi = month_interp(step, c);
  std::vector<double> out(points);
  for (int h = 0; h < points; h++)
    out[h] = (1 - mi.weight) * table[(size_t)mi.month * points + h] +
             mi.weight * table[(size_t)mi.neighbour * points + h];
  return out;
}

}  // namespace

// ------------------------------------------------------ star spectrum

namespace {

#include "specblock.inc"

// solarini's split of a blackbody star (SPECTRUM.split, physics/radiation/
// spectrum/index.ts:58-168): the energy shares of the bands below and above
// 0.75 um, the Rayleigh coefficient relative to the Sun's, and each
// surface's spectral albedo averaged over the star in each band. The sums
// run in the reference's order: trapezoids over 1024 log-spaced points per
// band, the gap between the bands added to the first.
struct Spectrum {
  double zsolar1 = 0, zsolar2 = 0, rcoeff = 1;
  double dsnowalb[2], dsnowalbmn[2], dsnowalbmx[2], dicealbmn[2], dicealbmx[2], dglacalbmn[2];
  double dgroundalb[2], doceanalb[2];
};

Spectrum spectrum_split(double star_k, double min_wavelength_m, bool simple_albedo) {
  const double hc_over_k = 0.0143877735383, sun_k = 5772, default_zsolar1 = 0.517;
  const int n = 1024, band_two = 41;
  const double cg = hc_over_k / sun_k, cs = hc_over_k / star_k;
  const double lw0 = std::log10(min_wavelength_m), hinge = std::log10(7.5e-7);
  const double dl1 = (hinge - lw0) / n, dl2 = (-4 - hinge) / n;
  const auto pow5 = [](double x) {
    const double x2 = x * x;
    return x2 * x2 * x;
  };
  const auto pow4 = [](double x) {
    const double x2 = x * x;
    return x2 * x2;
  };
  // Planck's law up to a constant: 1 / scaled / (exp(c / lambda) - 1)
  const auto planck = [](double lambda, double scaled, double c) { return ((1 / scaled) * 1) / (std::exp(c / lambda) - 1); };
  std::vector<double> w1(n), w2(n), b1(n), b2(n), g1(n), g2(n);
  for (int k = 0; k < n; k++) {
    w1[k] = std::pow(10.0, lw0 + k * dl1);
    w2[k] = std::pow(10.0, hinge + k * dl2);
    const double s1 = pow5(1e6 * w1[k]), s2 = pow5(1e6 * w2[k]);
    g1[k] = planck(w1[k], s1, cg);
    g2[k] = planck(w2[k], s2, cg);
    b1[k] = planck(w1[k], s1, cs);
    b2[k] = planck(w2[k], s2, cs);
  }
  const auto trap = [](const std::vector<double>& v, const std::vector<double>& at, int k) {
    return 0.5 * (v[k] + v[k + 1]) * (at[k + 1] - at[k]);
  };
  // the lambda^-4 (Rayleigh cross-section) weighted trapezoid
  const auto cross = [&](const std::vector<double>& v, const std::vector<double>& at, int k) {
    return 0.5 * (v[k] / pow4(at[k] * 1e6) + v[k + 1] / pow4(at[k + 1] * 1e6)) * (at[k + 1] - at[k]);
  };
  double z1 = 0, z2 = 0, zg1 = 0, zg2 = 0, zc1 = 0, zc2 = 0, zgc1 = 0, zgc2 = 0;
  for (int k = 0; k < n - 1; k++) {
    z1 = z1 + trap(b1, w1, k);
    z2 = z2 + trap(b2, w2, k);
    zg1 = zg1 + trap(g1, w1, k);
    zg2 = zg2 + trap(g2, w2, k);
    zc1 = zc1 + cross(b1, w1, k);
    zc2 = zc2 + cross(b2, w2, k);
    zgc1 = zgc1 + cross(g1, w1, k);
    zgc2 = zgc2 + cross(g2, w2, k);
  }
  const int last = n - 1;
  const double gap = w2[0] - w1[last];
  z1 = z1 + 0.5 * (b1[last] + b2[0]) * gap;
  zc1 = zc1 + 0.5 * (b1[last] / pow4(w1[last] * 1e6) + b2[0] / pow4(w2[0] * 1e6)) * gap;
  zg1 = zg1 + 0.5 * (g1[last] + g2[0]) * gap;
  zgc1 = zgc1 + 0.5 * (g1[last] / pow4(w1[last] * 1e6) + g2[0] / pow4(w2[0] * 1e6)) * gap;
  Spectrum out;
  const double zchi = (zgc1 + zgc2) / (zg1 + zg2);
  out.rcoeff = ((zc1 + zc2) * default_zsolar1) / z1 / zchi;
  const double zdenom1 = 0.01 / z1, zdenom2 = 0.01 / z2;
  out.zsolar1 = z1 / (z1 + z2);
  out.zsolar2 = 1 - out.zsolar1;
  std::vector<double> b3(kSpecPoints);
  for (int k = 0; k < kSpecPoints; k++)
    b3[k] = planck(kSpec_wavelengths[k], pow5(kSpec_wavelengths[k]), 1e6 * cs);
  const auto band = [&](const double* albedo, double* dst) {
    double a1 = 0, a2 = 0;
    for (int k = 0; k < band_two; k++)
      a1 = a1 + 0.5 * (b3[k] * albedo[k] + b3[k + 1] * albedo[k + 1]) * 1e-6 *
                    (kSpec_wavelengths[k + 1] - kSpec_wavelengths[k]);
    for (int k = band_two; k < kSpecPoints - 1; k++)
      a2 = a2 + 0.5 * (b3[k] * albedo[k] + b3[k + 1] * albedo[k + 1]) * 1e-6 *
                    (kSpec_wavelengths[k + 1] - kSpec_wavelengths[k]);
    a1 = zdenom1 * a1;
    a2 = zdenom2 * a2;
    const double broadband = out.zsolar1 * a1 + out.zsolar2 * a2;
    dst[0] = simple_albedo ? broadband : a1;
    dst[1] = simple_albedo ? broadband : a2;
  };
  band(kSpec_iceblend, out.dsnowalb);
  band(kSpec_iceblendmin, out.dsnowalbmn);
  band(kSpec_iceblendmax, out.dsnowalbmx);
  band(kSpec_seaicemin, out.dicealbmn);
  band(kSpec_seaicemax, out.dicealbmx);
  band(kSpec_glacalbmin, out.dglacalbmn);
  band(kSpec_groundblend, out.dgroundalb);
  band(kSpec_oceanblend, out.doceanalb);
  return out;
}

}  // namespace

void config_default(exo_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->truncation = 21;
  c->levels = 10;
  c->timestep_minutes = 45;
  c->radius_earth = 1;
  c->gravity = 9.80665;
  c->rotation_period_days = 1;
  c->pressure_bar = 1;
  c->gas_constant = 8314.46261815324 / 28.0134;  // 1 bar of N2
  c->co2_ppm = 360;
  c->flux_wm2 = 1367;
  c->obliquity_deg = 23.4;
  c->eccentricity = 0;
  c->year_days = 360;
  c->ocean_depth_m = 50;
  c->ozone = 1;
  c->land_fraction = 0.4;
  c->terrain_max_m = 4000;
  c->initial_temperature = 250;
  c->initial_humidity = 1e-5;
  c->noise_kelvin = 1e-3;
  c->sea_surface_k = 280;
  c->terrain_seed = 20240521;
  c->noise_seed = 14963991;
  // GCM.DEFAULT_CONFIG (index.ts:4-44) for the rest
  c->vegetation = 2;
  c->initial_growth = 0.5;
  c->glaciers = 1;
  c->glacier_min_depth_m = 2;
  c->glacier_initial_height_m = -1;
  c->star_temperature_k = 5780;
  c->synchronous = 0;
  c->substellar_longitude_deg = 180;
  c->fixed_orbit = 1;
  c->keplerian = 0;
  c->periapse_longitude_deg = NAN;
  c->initial_mean_anomaly_deg = NAN;
  c->soil_depth_scale = 1;
  c->soil_water_capacity_m = NAN;
  c->soil_heat_capacity = NAN;
  c->land_albedo = NAN;
  c->ocean_albedo = NAN;
  c->snow_ice_albedo = NAN;
  c->wet_soil = 0;
  c->filter = 2;
  // the base plasim_namelist's KICK and SEED, surfmod's NSPINIT default
  c->initial_state = 0;
  c->kick = 1;
  static const int32_t seed[8] = {14963991, 2, 3, 4, 5, 6, 7, 8};
  std::copy(seed, seed + 8, c->seed);
  c->clock_ms = 0;
  c->nspinit = 0;
}

std::vector<double> gaussian_degrees(int nlat) { return gaussian_grid(nlat).degrees; }

void member_date(const Member& m, int64_t step, int out[5]) {
  const Date d = date_of(step, m.cal.steps_per_day, m.cal.days_per_year, m.cal.step_minutes);
  out[0] = d.year;
  out[1] = d.month;
  out[2] = d.day;
  out[3] = d.hour;
  out[4] = d.minute;
}

void build_member(const exo_config& cfg, Member& m, const exo_surface_field* fields, int count) {
  if (cfg.levels != 5 && cfg.levels != 10) throw std::runtime_error("Choose 5 or 10 atmospheric layers");
  if (!(cfg.timestep_minutes > 0)) throw std::runtime_error("timestep must be positive");
  if (!(cfg.land_fraction >= 0 && cfg.land_fraction <= 1)) throw std::runtime_error("land fraction must be in [0, 1]");
  if (cfg.initial_state != 0 && cfg.initial_state != 1) throw std::runtime_error("initial_state must be 0 or 1");
  if (cfg.kick < 0 || cfg.kick > 3) throw std::runtime_error("KICK must be 0 to 3");
  m.cfg = cfg;
  const int T = cfg.truncation, L = cfg.levels;
  const int NLAT = latitudes_for(T), NLON = 2 * NLAT, NHOR = NLAT * NLON;
  const int NCSP = exo_ncsp(T), NRSP = exo_nrsp(T), NFC = exo_nfc(T);
  m.dims = {T, NLAT, NLON, L};
  m.nsnap = exo_snapshot_count(L);
  m.layout = make_layout(m.dims, m.nsnap);
  m.dev = device_layout(m.layout);
  m.tlayout = make_table_layout(m.dims);
  m.tables.assign(m.tlayout.total, 0.0);
  auto tab = [&](ExoTable t) { return &m.tables[m.tlayout.offset[t]]; };

  Pumamod p = pumamod_read(cfg);
  const Grid grid = gaussian_grid(NLAT);

  // FOURIER.tables (fourier/index.ts:19-36)
  {
    double* fwd = tab(TAB_fourier_forward);
    double* inv = tab(TAB_fourier_inverse);
    for (int j = 0; j < NLON; j++) {
      const double lambda = 2 * M_PI * j / NLON;
      for (int mm = 0; mm <= T; mm++) {
        const double c = std::cos(mm * lambda), s = std::sin(mm * lambda);
        fwd[j * NFC + 2 * mm] = c / NLON;
        fwd[j * NFC + 2 * mm + 1] = -s / NLON;
        const double w = mm == 0 ? 1 : 2;
        inv[(size_t)2 * mm * NLON + j] = w * c;
        inv[(size_t)(2 * mm + 1) * NLON + j] = mm == 0 ? 0 : -w * s;
      }
    }
  }
  // LEGENDRE.kernelTables (legendre/index.ts:202-221)
  const Legendre leg = legini(grid, T);
  std::vector<double> g2s, s2g;
  physics_filters(p, T, g2s, s2g);
  auto filtered = [&](const std::vector<double>& src, const std::vector<double>& f, double* dst) {
    for (int lat = 0; lat < NLAT; lat++) {
      int w = 0;
      for (int mm = 0; mm <= T; mm++)
        for (int n = mm; n <= T; n++, w++) dst[(size_t)lat * NCSP + w] = src[(size_t)lat * NCSP + w] * f[n];
    }
  };
  m.qc_raw = leg.qc;
  m.qi_raw = leg.qi;
  filtered(leg.qc, g2s, tab(TAB_legendre_qc));
  filtered(leg.qe, g2s, tab(TAB_legendre_qe));
  filtered(leg.qm, g2s, tab(TAB_legendre_qm));
  filtered(leg.qq, g2s, tab(TAB_legendre_qq));
  filtered(leg.qi, s2g, tab(TAB_legendre_qi));
  filtered(leg.qj, s2g, tab(TAB_legendre_qj));
  filtered(leg.qu, s2g, tab(TAB_legendre_qu));
  filtered(leg.qv, s2g, tab(TAB_legendre_qv));

  // Dynamics set-up.
  Dyn dyn = initpm(p, T, L);
  const std::vector<double>& sigma = dyn.sigma;
  const std::vector<double>& sigmah = dyn.sigmah;
  const std::vector<double>& dsigma = dyn.dsigma;

  // Module set-ups (MODULE_SETUP, setup/modules/index.ts:13-305): each
  // module's namelist defaults, the configuration's edits over them
  // (namelist/index.ts:40-182) and what its *ini routine derives.
  struct {
    double gsol0, zsolar1 = 0.517, zsolar2 = 0.483, zsolars1 = 0, zsolars2 = 0, rcoeff = 1;
    double th2oc = 0.024, tswr1 = 0.077, tswr2 = 0.065, tswr3 = 0.0055, tpofmt = 1, acllwr = 0.1;
    double a0o3 = 0.25, a1o3 = 0.11, aco3 = 0.08, bo3 = 20000, co3 = 5000, toffo3 = 0.25;
    double dawn = 0, co2 = 360, starbbtemp = 5772, minwavel = 316.036116751;
    int nfixed = 0, nstartemp = 0, no3 = 1, nsimplealbedo = 1;
  } rad;
  rad.gsol0 = p.gsol0;
  rad.co2 = cfg.co2_ppm;
  rad.no3 = cfg.ozone ? 1 : 0;
  rad.nfixed = cfg.synchronous ? 1 : 0;
  if (cfg.star_temperature_k > 0) {
    rad.nstartemp = 1;
    rad.starbbtemp = cfg.star_temperature_k;
  }
  // radini's band albedos: the star's spectral averages (solarini), else
  // the grey defaults (setup/modules/index.ts:66-83)
  double dsnowalb[2] = {0.6, 0.6}, dsnowalbmn[2] = {0.4, 0.4}, dsnowalbmx[2] = {0.8, 0.8};
  double dicealbmn[2] = {0.5, 0.5}, dicealbmx[2] = {0.7, 0.7};
  double dglacalbmn[2] = {0.6, 0.6}, dgroundalb[2] = {0.2, 0.2};
  double doceanalb[2] = {0.069, 0.069};
  if (rad.nstartemp > 0) {
    const Spectrum sp = spectrum_split(rad.starbbtemp, rad.minwavel * 1e-9, rad.nsimplealbedo > 0.5);
    rad.zsolar1 = sp.zsolar1;
    rad.zsolar2 = sp.zsolar2;
    rad.zsolars1 = sp.zsolar1;
    rad.zsolars2 = sp.zsolar2;
    rad.rcoeff = sp.rcoeff;
    for (int b = 0; b < 2; b++) {
      dsnowalb[b] = sp.dsnowalb[b];
      dsnowalbmn[b] = sp.dsnowalbmn[b];
      dsnowalbmx[b] = sp.dsnowalbmx[b];
      dicealbmn[b] = sp.dicealbmn[b];
      dicealbmx[b] = sp.dicealbmx[b];
      dglacalbmn[b] = sp.dglacalbmn[b];
      dgroundalb[b] = sp.dgroundalb[b];
      doceanalb[b] = sp.doceanalb[b];
    }
  }
  (void)dsnowalb;
  struct {
    double rhbeta = 0, nbeta = 3, pdeep = 999999, pdeepth = 70000, rkshallow = 10, gamma;
    double clwcrit1 = -0.1, clwcrit2 = 0, rcritmod = 1, rcritslope = 0, clwfac;
  } rain;
  rain.gamma = (T == 42 && L == 10) ? 0.007 : 0.01;
  rain.clwfac = rain.clwcrit1 > rain.clwcrit2 ? 1 / (rain.clwcrit1 - rain.clwcrit2)
                : rain.clwcrit1 < rain.clwcrit2 ? -1 : 1;
  const double zumin = 1, vdiff_lamm = 160, vdiff_b = 5, vdiff_c = 5, vdiff_d = 5;
  struct {
    double nwetsoil = 0, albland = 0.22, albsmin = 0.4, albsmax = 0.8, albsminf = 0.3, albsmaxf = 0.4;
    double albgmin = 0.6, albgmax = 0.8, alblandmax = 0.4, drhsfull = 0.4, dztop = 0.2, dsmax = 5;
    double wsmax = 0.5, rhosnow = 330, soildiff = 1.8, sicediff = 2.03, snowdiff = 0.31;
    double soilcap = 2.4e6, sicecap = 2.07e6, snowcap = 0.6897e6, rinifor = 0.5, dz0land = 2;
    double drhsland = 0.25, dzglac = -1, nwatcini = 0, dwatcini = 0;
    double rlue = 3.4e-10, co2conv = 8.3e-4, tau_veg = 10, tau_soil = 42;
    double dsoilz[5] = {0.4, 0.8, 1.6, 3.2, 6.4};
  } land;
  // landmod_namelist edits (namelist/index.ts:114-163)
  for (double v : {cfg.soil_depth_scale, cfg.soil_water_capacity_m, cfg.soil_heat_capacity})
    if (!std::isnan(v) && (!std::isfinite(v) || v <= 0))
      throw std::runtime_error("Soil depth, water capacity and heat capacity must be positive");
  if (std::isnan(cfg.soil_depth_scale)) throw std::runtime_error("Soil depth, water capacity and heat capacity must be positive");
  for (double v : {cfg.land_albedo, cfg.ocean_albedo, cfg.snow_ice_albedo})
    if (!std::isnan(v) && (!std::isfinite(v) || v < 0 || v > 1))
      throw std::runtime_error("Albedo overrides must be between 0 and 1");
  land.nwetsoil = cfg.wet_soil ? 1 : 0;
  for (int i = 0; i < 5; i++) land.dsoilz[i] *= cfg.soil_depth_scale;
  if (!std::isnan(cfg.soil_water_capacity_m)) land.wsmax = cfg.soil_water_capacity_m;
  if (!std::isnan(cfg.soil_heat_capacity)) land.soilcap = cfg.soil_heat_capacity;
  struct {
    double albsea = 0.069, albice = 0.7, dz0sea = 1.5e-5, dz0ice = 0.001, drhssea = 1, drhsice = 1;
    double charnock = 0.018;
  } sea;
  if (!std::isnan(cfg.land_albedo)) {
    land.albland = cfg.land_albedo;
    dgroundalb[0] = dgroundalb[1] = cfg.land_albedo;
  }
  if (!std::isnan(cfg.ocean_albedo)) {
    sea.albsea = cfg.ocean_albedo;
    doceanalb[0] = doceanalb[1] = cfg.ocean_albedo;
  }
  if (!std::isnan(cfg.snow_ice_albedo)) {
    const double a = cfg.snow_ice_albedo;
    sea.albice = a;
    for (int b = 0; b < 2; b++)
      dicealbmn[b] = dicealbmx[b] = dsnowalbmn[b] = dglacalbmn[b] = dsnowalb[b] = a;
  }
  // landini: snow and glacier band albedos from the band maxima and minima,
  // vegetation time scales from years to seconds (setup/modules/index.ts:
  // 176-195)
  land.tau_veg = land.tau_veg * p.m_days_per_year * p.day_24hr;
  land.tau_soil = land.tau_soil * p.m_days_per_year * p.day_24hr;
  struct {
    double nice = 1, nseaice = 1, xmind = 0.1, xmaxd = 9, thicec = 0.5, tfreeze = 271.25;
    double cicemin = 0.5;
  } ice;
  struct {
    double ncarbon = 1, nsupply = 0, kact = 0.09, beta = 0.5, frequency = 4, vearth = 0.05;
    double pearth = 79, wmax = 1;
  } carbon;
  // carbonmod_namelist base values (t21 namelists.json)
  carbon.wmax = 80.9;
  carbon.frequency = std::min(carbon.frequency, (double)p.mtspd);
  const int carbon_interval = (int)std::trunc(p.mtspd / carbon.frequency);
  const double timeweight = 1 / (carbon.frequency * p.m_days_per_year);
  const double mldepth = cfg.ocean_depth_m;
  const int noromax = T;
  // vegmod_namelist (setup/modules/index.ts:234-247; RINIDAGG edited)
  struct {
    double ncveg = 1, forgrow = 1, rinidagg = 0.5, rinidsc = 1, rinidmr = 2, rinisoil = 0, riniveg = 0;
  } veg;
  veg.rinidagg = cfg.initial_growth;
  // glacier_namelist (setup/modules/index.ts:293-298 defaults; edits)
  const int nglacier = cfg.glaciers ? 1 : 0;
  const double glacelim = cfg.glacier_min_depth_m;
  const double icesheeth = cfg.glacier_initial_height_m;

  // ------------------------------------------------- physics tables
  {
    // SURFACE_FLUX.tables (physics/surface-flux/index.ts:31-68)
    const double VK = 0.4, ZTSCAL = 250;
    const double zlamh = vdiff_lamm * std::sqrt(3 * vdiff_d * 0.5);
    tab(TAB_flux_sigfac)[0] = std::pow(sigma[L - 1], -p.akap);
    tab(TAB_flux_lnsig)[0] = std::log(sigma[L - 1]);
    for (int l = 0; l < L; l++) {
      tab(TAB_vdiff_skap)[l] = std::pow(sigma[l], p.akap);
      tab(TAB_vdiff_skaph)[l] = std::pow(sigmah[l], p.akap);
    }
    for (int l = 0; l < L - 1; l++) {
      const double zzlev = -p.gascon * ZTSCAL * std::log(sigmah[l]) / p.ga;
      const double zdz = p.gascon * ZTSCAL * std::log(sigma[l + 1] / sigma[l]) / p.ga;
      const double zmixh = zlamh * VK * zzlev / (zlamh + VK * zzlev);
      const double zmixm = vdiff_lamm * VK * zzlev / (vdiff_lamm + VK * zzlev);
      const double zrdsig = 1 / (sigma[l + 1] - sigma[l]);
      tab(TAB_vdiff_mixh2)[l] = zmixh * zmixh;
      tab(TAB_vdiff_mixm2)[l] = zmixm * zmixm;
      tab(TAB_vdiff_rdsig)[l] = zrdsig;
      tab(TAB_vdiff_sigmah3)[l] = sigmah[l] * sigmah[l] * sigmah[l] * zrdsig;
      const double ratio = (std::pow((zzlev + zdz) / zzlev, 1.0 / 3.0) - 1) / zdz;
      tab(TAB_vdiff_kfac)[l] = std::sqrt(ratio * ratio * ratio / zzlev);
    }
    // RADIATION.tables (physics/radiation/index.ts:59-85)
    for (int lat = 0; lat < NLAT; lat++) {
      tab(TAB_sid)[lat] = grid.sines[lat];
      tab(TAB_cola)[lat] = std::sqrt(1 - grid.sines[lat] * grid.sines[lat]);
    }
    tab(TAB_rad_lnsig)[0] = std::log(sigma[0] / sigmah[0]) * 0.5;
    for (int l = 1; l < L; l++) tab(TAB_rad_lnsig)[l] = std::log(sigmah[l - 1] / sigmah[l]);
    for (int l = 0; l < L; l++) {
      tab(TAB_lwr_zzf1)[l] = sigma[l] * dsigma[l] / p.ga / 100000;
      tab(TAB_lwr_zzf3)[l] = -1.66 * rad.acllwr * 1000 * dsigma[l] / p.ga;
    }
    // PRECIPITATION.tables (physics/precipitation/index.ts:47-67)
    for (int l = 0; l < L; l++)
      tab(TAB_rcrit)[l] = std::max(0.85, std::max(sigma[l], 1 - sigma[l])) * rain.rcritmod *
                          (1 - 2 * rain.rcritslope * (L / 2.0 - (l + 1)) / L);
    for (int l = 0; l < L - 1; l++) tab(TAB_kuo_zfac)[l] = std::pow(sigma[l] / sigma[l + 1], p.akap);
    for (int i = 0; i < 5; i++) tab(TAB_dsoilz)[i] = land.dsoilz[i];
    // FIXER.tables (physics/fixer/index.ts:28-37) and rcsq (dynamics/setup/
    // index.ts:236-248)
    for (int lat = 0; lat < NLAT; lat++)
      for (int j = 0; j < NLON; j++) {
        tab(TAB_zgw)[lat * NLON + j] = grid.weights[lat];
        tab(TAB_rcsq)[lat * NLON + j] = 1 / (1 - grid.sines[lat] * grid.sines[lat]);
      }
    // DYNAMICS_SETUP.tables (dynamics/setup/index.ts:306-344)
    for (int l = 0; l < L; l++) {
      tab(TAB_t0)[l] = dyn.t0[l];
      tab(TAB_dsigma)[l] = dsigma[l];
      tab(TAB_sigma)[l] = sigma[l];
      tab(TAB_sigmah)[l] = sigmah[l];
      tab(TAB_rdsig)[l] = dyn.rdsig[l];
      tab(TAB_t01s2)[l] = dyn.t01s2[l];
      tab(TAB_tkp)[l] = dyn.tkp[l];
      tab(TAB_damp)[l] = dyn.damp[l];
      tab(TAB_tdisst)[l] = dyn.tdisst[l];
      tab(TAB_tdissq)[l] = dyn.tdissq[l];
      tab(TAB_tdissd)[l] = dyn.tdissd[l];
      tab(TAB_tdissz)[l] = dyn.tdissz[l];
      tab(TAB_tfrc)[l] = dyn.tfrc[l];
    }
    std::copy(dyn.c.begin(), dyn.c.end(), tab(TAB_c));
    std::copy(dyn.g.begin(), dyn.g.end(), tab(TAB_g));
    std::copy(dyn.tau.begin(), dyn.tau.end(), tab(TAB_tau));
    std::copy(dyn.nindex.begin(), dyn.nindex.end(), tab(TAB_nindex));
    std::copy(dyn.sak.begin(), dyn.sak.end(), tab(TAB_sakpp));
  }

  // ------------------------------------------------- calendar
  Calendar& cal = m.cal;
  cal.steps_per_day = p.ntspd;
  cal.steps_per_year = p.n_steps_per_year;
  cal.output_window = p.nafter;
  cal.step_minutes = p.mpstep;
  cal.days_per_year = p.n_days_per_year;
  cal.days_per_month = p.n_days_per_month;
  cal.solar_day = p.solar_day;
  cal.rotation_speed = p.rotspd;
  cal.toffo3 = rad.toffo3;
  cal.keplerian = p.ngenkeplerian;
  cal.synchronous = rad.nfixed;
  cal.fixed_longitude = p.fixedlon;
  cal.carbon_interval = carbon_interval;
  {
    // CALENDAR.orbit with a fixed orbit (calendar/index.ts:36-126)
    const double pi = 4 * std::atan(1.0), degrad = pi / 180;
    const double e = p.eccen, e2 = e * e, e3 = e2 * e;
    const double mvelpp = (p.mvelp + 180) * degrad;
    const double beta = std::sqrt(1 - e2);
    cal.eccentricity = e;
    cal.obliquity_rad = p.obliq * degrad;
    cal.mean_anomaly_rad = p.meananom0 * degrad;
    cal.perihelion_lon_rad = mvelpp;
    cal.mean_lon_equinox_rad =
        2 * ((0.5 * e + 0.125 * e3) * (1 + beta) * std::sin(mvelpp) -
             0.25 * e2 * (0.5 + beta) * std::sin(2 * mvelpp) +
             0.125 * e3 * (1.0 / 3 + beta) * std::sin(3 * mvelpp));
  }
  m.start_step = p.n_start_step;

  // ------------------------------------------------- synthetic planet
  // Terrain: a real spherical-harmonic random field with degree power
  // proportional to n^-2, on this truncation's grid; land is the highest
  // land_fraction of points, raised 0..terrain_max_m above the coast.
  std::vector<double> height(NHOR);
  {
    Rng rng(cfg.terrain_seed);
    std::vector<double> spec(NRSP, 0.0);
    int w = 0;
    for (int mm = 0; mm <= T; mm++)
      for (int n = mm; n <= T; n++, w++) {
        if (n == 0) continue;
        const double sd = std::sqrt(1.0 / ((double)n * n * (2 * n + 1)));
        spec[2 * w] = sd * rng.normal();
        spec[2 * w + 1] = mm == 0 ? 0 : sd * rng.normal();
      }
    spectral_to_grid(m, spec.data(), height.data());
  }
  // The .sra fields given, by code, read as get_surf_array reads them
  // (setup/surface/index.ts:76-107): the first `months` records; 12 of 14
  // gain December before and January after, one fills all 14.
  std::map<int, const exo_surface_field*> given;
  for (int i = 0; i < count; i++) {
    if (fields[i].months < 1) throw std::runtime_error("a surface field needs at least one month");
    if (!given.emplace(fields[i].code, &fields[i]).second)
      throw std::runtime_error("surface field " + std::to_string(fields[i].code) + " given twice");
  }
  const auto read = [&](int code, int months, std::vector<double>& out) {
    const auto it = given.find(code);
    if (it == given.end()) return false;
    const int found = std::min(months, it->second->months);
    out.assign((size_t)months * NHOR, 0.0);
    std::copy(it->second->values, it->second->values + (size_t)found * NHOR, out.begin());
    if (found == months) return true;
    auto at = [&](int month) { return out.begin() + (size_t)month * NHOR; };
    if (found == 12 && months == CLIM_MONTHS) {
      for (int mo = 12; mo >= 1; mo--) std::copy(at(mo - 1), at(mo), at(mo));
      std::copy(at(1), at(2), at(13));
      std::copy(at(12), at(13), at(0));
    } else if (found == 1 && months == CLIM_MONTHS) {
      for (int mo = 1; mo < CLIM_MONTHS; mo++) std::copy(at(0), at(1), at(mo));
    } else {
      char name[32];
      std::snprintf(name, sizeof name, "N%03d_surf_%04d.sra", NLAT, code);
      throw std::runtime_error(std::string(name) + " has " + std::to_string(found) + " fields, expected " +
                               std::to_string(months));
    }
    return true;
  };
  const auto field = [&](int code, double fallback) {
    std::vector<double> v;
    if (!read(code, 1, v)) v.assign(NHOR, fallback);
    return v;
  };
  const auto climatology = [&](int code, double fallback) {
    std::vector<double> v;
    if (!read(code, CLIM_MONTHS, v)) v.assign((size_t)CLIM_MONTHS * NHOR, fallback);
    return v;
  };

  // landini's land mask (172) and orography (129): from the fields, else
  // the seeded terrain
  std::vector<double> dls(NHOR, 0.0), orography(NHOR, 0.0);
  const bool seeded = !given.count(172) && !given.count(129);
  if (!seeded) {
    // boundary files 172 and 129 (setup/surface/index.ts:293-297)
    const std::vector<double> mask = field(172, 1);
    orography = field(129, 0);
    for (int h = 0; h < NHOR; h++) dls[h] = mask[h] > 0.5 ? 1 : 0;
  } else {
    std::vector<double> sorted = height;
    std::sort(sorted.begin(), sorted.end());
    const int nland = (int)std::lround(cfg.land_fraction * NHOR);
    const double threshold = nland > 0 ? sorted[NHOR - nland] : 1e300;
    const double top = sorted[NHOR - 1];
    for (int h = 0; h < NHOR; h++) {
      if (nland > 0 && height[h] >= threshold) {
        dls[h] = 1;
        const double z = top > threshold ? (height[h] - threshold) / (top - threshold) : 0;
        orography[h] = z * cfg.terrain_max_m * p.ga;  // geopotential, m^2/s^2
      }
    }
  }
  m.land_mask = dls;
  // the land mask as icemod and oceanmod read it (fallback 0, not 1)
  std::vector<double> xls(NHOR);
  {
    const std::vector<double> mask = seeded ? dls : field(172, 0);
    for (int h = 0; h < NHOR; h++) xls[h] = mask[h] > 0.5 ? 1 : 0;
  }

  // ------------------------------------------------- surface initial state
  // SURFACE_SETUP.initial (setup/surface/index.ts:267-643): landini,
  // glacierini and oroini, vegini, iceini and oceanini, seaini; every code
  // not given takes the reference's fallback. A missing surface
  // temperature climatology (169) is the configuration's uniform
  // sea_surface_k when that is a number.
  std::vector<double> A(m.layout.total, 0.0);
  auto slot = [&](ExoSlot s) { return &A[m.layout.offset[s]]; };
  const double tmelt = p.tmelt;
  std::vector<double> doro = orography;
  {
    double* a_dls = slot(SLOT_dls);
    for (int h = 0; h < NHOR; h++) a_dls[h] = dls[h];
    double* dz0 = slot(SLOT_dz0);
    double* drhs = slot(SLOT_drhs);
    double* dwatc = slot(SLOT_dwatc);
    double* dalb = slot(SLOT_dalb);
    double* dsalb = slot(SLOT_dsalb);
    double* dsnow = slot(SLOT_dsnow);
    double* dt = slot(SLOT_dt);
    double* dq = slot(SLOT_dq);
    const size_t surf = (size_t)L * NHOR;
    const std::vector<double> dz0clim = field(173, land.dz0land);
    const std::vector<double> dz0climo = field(1730, 0);
    std::vector<double> dglac = field(232, 0);
    if (land.dzglac > 0)
      for (int h = 0; h < NHOR; h++)
        if (orography[h] > land.dzglac * p.ga) dglac[h] = 1;
    for (auto& v : dglac) v = v > 0.5 ? 1 : 0;
    std::vector<double> dforest = field(212, land.rinifor);
    std::vector<double> dwmax = field(229, land.wsmax);
    for (auto& v : dwmax) v = std::max(v, 0.0);
    const std::vector<double> dtclsoil = field(209, tmelt);
    const std::vector<double> dalbcl = climatology(174, land.albland);
    const std::vector<double> dalbcl1 = climatology(175, dgroundalb[0]);
    const std::vector<double> dalbcl2 = climatology(176, dgroundalb[1]);
    double* land_dts = slot(SLOT_land_dts);
    double* land_dtsm = slot(SLOT_land_dtsm);
    double* land_dqs = slot(SLOT_land_dqs);
    for (int h = 0; h < NHOR; h++) land_dts[h] = land_dtsm[h] = land_dqs[h] = 1e20;
    std::vector<double> dsoilt((size_t)SOIL_LAYERS * NHOR, tmelt);
    double* dsnowt = slot(SLOT_dsnowt);
    double* dsnowz = slot(SLOT_dsnowz);
    for (int h = 0; h < NHOR; h++) dsnowt[h] = tmelt;
    for (int h = 0; h < NHOR; h++) {
      if (dls[h] <= 0) continue;
      land_dts[h] = dtclsoil[h];
      for (int l = 0; l < SOIL_LAYERS; l++) dsoilt[(size_t)l * NHOR + h] = land_dts[h];
      dwatc[h] = land.nwatcini > 0 ? land.dwatcini : land.drhsland * land.drhsfull * dwmax[h];
    }
    const int64_t step = m.start_step;
    std::vector<double> dalbclim = monthly(dalbcl, step + 1, cal, NHOR);
    std::vector<double> dalbclim1 = monthly(dalbcl1, step + 1, cal, NHOR);
    std::vector<double> dalbclim2 = monthly(dalbcl2, step + 1, cal, NHOR);
    if (land.nwetsoil > 0.5) {
      // getalb at a fresh start: the bucket is empty (bf = 0)
      const auto wet_soil = [](double bf, double dry, double ocean) {
        const double al = 1 / 5.2 * std::pow(std::pow(std::max(bf, 1e-3), (1 - 4.0) / 4) - 1, 1 / 4.0);
        const double dry_w = std::exp(-std::pow(bf * 25, 6)), wet_w = std::exp(-std::pow((1 - bf) * 30, 9));
        return dry * dry_w + al * (1 - dry_w - wet_w) + ocean * wet_w;
      };
      for (int h = 0; h < NHOR; h++) {
        if (dls[h] <= 0.5) continue;
        if (land_dts[h] > 273.15) {
          dalbclim[h] = wet_soil(0, land.alblandmax, doceanalb[0]);
          dalbclim1[h] = dalbclim[h];
          dalbclim2[h] = wet_soil(0, land.alblandmax, doceanalb[1]);
        } else {
          dalbclim[h] = land.albland;
          dalbclim1[h] = dgroundalb[0];
          dalbclim2[h] = dgroundalb[1];
        }
      }
    }
    const double albgmax1 = dsnowalbmx[0], albgmax2 = dsnowalbmx[1];
    const auto glacier_albedo = [&](double mx, double mn, double ts) {
      return std::max(mn, std::min(mx, mx - ((mx - mn) * (ts - 263.16)) / (tmelt - 263.16)));
    };
    double* soil[5] = {slot(SLOT_dtsoil), slot(SLOT_dtd2), slot(SLOT_dtd3), slot(SLOT_dtd4), slot(SLOT_dtd5)};
    for (int h = 0; h < NHOR; h++) {
      if (dls[h] <= 0) continue;
      const double ts = land_dts[h];
      land_dtsm[h] = ts;
      land_dqs[h] = saturation(ts, p);
      dsnow[h] = dsnowz[h];
      dalb[h] = dalbclim[h];
      dsalb[h] = dalbclim1[h];
      dsalb[NHOR + h] = dalbclim2[h];
      if (dwmax[h] > 0) drhs[h] = std::min(1.0, dwatc[h] / (land.drhsfull * dwmax[h]));
      dz0[h] = dz0clim[h];
      for (int l = 0; l < SOIL_LAYERS; l++) soil[l][h] = dsoilt[(size_t)l * NHOR + h];
      if (dglac[h] > 0.5) {
        dsnowz[h] = std::max(land.dsmax, 0.0);
        dsnow[h] = dsnowz[h];
        dalb[h] = glacier_albedo(land.albgmax, land.albgmin, ts);
        dsalb[h] = glacier_albedo(albgmax1, dglacalbmn[0], ts);
        dsalb[NHOR + h] = glacier_albedo(albgmax2, dglacalbmn[1], ts);
        drhs[h] = 1;
      }
      dt[surf + h] = ts;
      dq[surf + h] = land_dqs[h];
    }
    // glacierini and oroini: glacier ice raises the orography; rivers
    // follow it and its transform is so (setup/surface/index.ts:412-467)
    if (icesheeth >= 0) {
      for (int h = 0; h < NHOR; h++) {
        const bool covered = orography[h] != 0;
        dglac[h] = covered ? 1 : 0;
        dsnowz[h] = covered ? icesheeth : 0;
      }
    } else {
      for (int h = 0; h < NHOR; h++)
        if (dglac[h] > 0 && dsnowz[h] == 0) dsnowz[h] = glacelim;
    }
    if (nglacier == 1) {
      const double gm = 6.67408e-11 * 5.9722e24, re = 6.371e6, rho = 850;
      for (int h = 0; h < NHOR; h++) {
        const double ground = orography[h] * 1;
        const double dz = (re * re * ground) / (gm - re * ground);
        const double snow = dsnowz[h] / (rho / 1000);
        const double gravity = gm / ((re + dz) * (re + dz));
        doro[h] = ground + (gravity * snow - (gravity / (dz + re)) * snow * snow);
      }
    }
    // oroini's river directions (setup/surface/index.ts:199-262).
    std::vector<double> duroff(NHOR, 0.0), dvroff(NHOR, 0.0);
    {
      std::vector<double> ht(NHOR);
      for (int h = 0; h < NHOR; h++) ht[h] = dls[h] < 1 ? dls[h] - 1 : std::max(doro[h], 0.0);
      const double hmax = *std::max_element(ht.begin(), ht.end());
      const double hmin = *std::min_element(ht.begin(), ht.end());
      if ((hmax - hmin) / p.ga > 0.01) {
        bool raised = true;
        std::vector<double> next(NHOR);
        while (raised) {
          raised = false;
          for (int h = 0; h < NHOR; h++) {
            const int lon = h % NLON, lat = h / NLON;
            double around[4];
            int na = 0;
            around[na++] = ht[lat * NLON + (lon + 1) % NLON];
            around[na++] = ht[lat * NLON + (lon + NLON - 1) % NLON];
            if (lat > 0) around[na++] = ht[h - NLON];
            if (lat < NLAT - 1) around[na++] = ht[h + NLON];
            bool lowest = true;
            double mn = around[0];
            for (int k = 0; k < na; k++) {
              if (!(ht[h] <= around[k])) lowest = false;
              mn = std::min(mn, around[k]);
            }
            if (dls[h] > 0 && lowest) {
              next[h] = 1 + mn;
              raised = true;
            } else {
              next[h] = ht[h];
            }
          }
          ht = next;
        }
        auto rate = [](double slope, double spacing, double sign) {
          return sign * 4.2 / spacing * std::pow(std::fabs(slope), 0.18);
        };
        for (int lat = 0; lat < NLAT; lat++) {
          const double dx = TWOPI * std::cos(std::asin(grid.sines[lat])) * p.plarad / NLON;
          for (int lon = 0; lon < NLON; lon++) {
            const int h = lat * NLON + lon;
            const double slope = (ht[lat * NLON + (lon + 1) % NLON] - ht[h]) / dx;
            duroff[h] = rate(slope, dx, slope > 0 ? -1 : 1);
          }
        }
        for (int lat = 0; lat < NLAT - 1; lat++) {
          const double dy = (std::asin(grid.sines[lat + 1]) - std::asin(grid.sines[lat])) * p.plarad;
          for (int lon = 0; lon < NLON; lon++) {
            const int h = lat * NLON + lon;
            const double slope = (ht[h + NLON] - ht[h]) / dy;
            dvroff[h] = rate(slope, dy, slope < 0 ? -1 : 1);
          }
        }
      }
    }
    std::copy(duroff.begin(), duroff.end(), slot(SLOT_duroff));
    std::copy(dvroff.begin(), dvroff.end(), slot(SLOT_dvroff));
    if (nglacier == 1)
      for (int h = 0; h < NHOR; h++)
        if (dsnowz[h] > 30) dglac[h] = 1;
    // vegini (setup/surface/index.ts:471-479)
    const std::vector<double> dagg = field(1605, veg.rinidagg), dsc = field(1606, veg.rinidsc);
    const std::vector<double> dmr = field(1607, veg.rinidmr), dgrow = field(1604, veg.forgrow);
    for (int h = 0; h < NHOR; h++) {
      slot(SLOT_dpersist)[h] = dls[h] < 0.5 ? 0 : 1;
      slot(SLOT_dforest)[h] = dls[h] == 0 ? 0 : dforest[h];
      slot(SLOT_dglac)[h] = dglac[h];
      slot(SLOT_dwmax)[h] = dwmax[h];
      slot(SLOT_darea)[h] = grid.weights[h / NLON];
      slot(SLOT_dz0climo)[h] = dz0climo[h];
      slot(SLOT_dalbclim)[h] = dalbclim[h];
      slot(SLOT_dalbclim1)[h] = dalbclim1[h];
      slot(SLOT_dalbclim2)[h] = dalbclim2[h];
      slot(SLOT_dagg)[h] = dagg[h];
      slot(SLOT_dsc)[h] = dsc[h];
      slot(SLOT_dmr)[h] = dmr[h];
      slot(SLOT_dgrow)[h] = dgrow[h];
      slot(SLOT_xgw)[h] = grid.weights[h / NLON];
    }
    std::copy(dsoilt.begin(), dsoilt.end(), slot(SLOT_dsoilt));
    std::copy(dalbcl.begin(), dalbcl.end(), slot(SLOT_dalbcl));
    std::copy(dalbcl1.begin(), dalbcl1.end(), slot(SLOT_dalbcl1));
    std::copy(dalbcl2.begin(), dalbcl2.end(), slot(SLOT_dalbcl2));
    for (int l = 0; l < L; l++)
      for (int h = 0; h < NHOR; h++) slot(SLOT_dqco2)[(size_t)l * NHOR + h] = rad.co2 > 0 ? rad.co2 : 0;
    slot(SLOT_co2)[0] = rad.co2;

    // iceini's read_ice_surface and iceget, oceanini's oceanget
    // (setup/surface/index.ts:481-536)
    const double tfreeze = ice.nseaice == 0 ? 0 : ice.tfreeze;
    std::vector<double> xclsst;
    if (!read(169, CLIM_MONTHS, xclsst))
      xclsst.assign((size_t)CLIM_MONTHS * NHOR, std::isnan(cfg.sea_surface_k) ? -999.0 : cfg.sea_surface_k);
    std::vector<double> xclicec = climatology(210, -999), xcliced = climatology(211, -999);
    if (ice.nseaice == 0) {
      std::fill(xclicec.begin(), xclicec.end(), 0.0);
      std::fill(xcliced.begin(), xcliced.end(), 0.0);
    }
    const double cover_max = *std::max_element(xclicec.begin(), xclicec.end());
    if (ice.nice >= 0) {
      if (cover_max > 5)
        for (auto& v : xclicec) v *= 0.01;
      if (cover_max < 0)
        for (size_t i = 0; i < xclicec.size(); i++) xclicec[i] = (ice.nice > 0.5 && xclsst[i] <= tfreeze) ? 1 : 0;
      if (xcliced[0] < 0) {
        // make_ice_thickness (CCM3; setup/surface/index.ts:265-290)
        static const double fac[14] = {0.912, 0.942, 1, 1.058, 1.124, 1.161, 1.175,
                                        1.058, 0.931, 0.883, 0.88,  0.876, 0.912, 0.942};
        for (int mo = 0; mo < CLIM_MONTHS; mo++)
          for (int h = 0; h < NHOR; h++) {
            const double c = xclicec[(size_t)mo * NHOR + h];
            const bool north = (h / NLON) < NLAT / 2;
            double d = 0;
            if (north) {
              if (c >= 0.9) d = 3 * fac[mo];
              else if (c >= 0.1) d = (0.25 + (3 - 0.25) * (c - 0.1) / (0.9 - 0.1)) * fac[mo];
              else if (c > 0) d = 0.25 * fac[mo];
            } else if (c >= 0.25) d = 0.5;
            else if (c > 0) d = 0.25 + c;
            xcliced[(size_t)mo * NHOR + h] = d;
          }
      }
    }
    for (size_t i = 0; i < xclicec.size(); i++)
      if (ice.nseaice == 0 || xls[i % NHOR] >= 1) {
        xclicec[i] = 0;
        xcliced[i] = 0;
      }
    std::vector<double> xclssto = monthly(xclsst, step, cal, NHOR);
    std::vector<double> xclsst2 = monthly(xclsst, step + 1, cal, NHOR);
    std::vector<double> xcliced2 = monthly(xcliced, step + 1, cal, NHOR);
    for (auto& v : xcliced2) v = std::max(v, 0.0);
    std::vector<double> xclicec2 = monthly(xclicec, step + 1, cal, NHOR);
    for (int h = 0; h < NHOR; h++) xclicec2[h] = xcliced2[h] <= 0 ? 0 : std::max(xclicec2[h], 0.0);
    std::vector<double> xiced = xcliced2, xicec(NHOR);
    for (int h = 0; h < NHOR; h++) xicec[h] = xclicec2[h] >= ice.thicec ? 1 : 0;
    std::vector<double> yclsst;
    if (!read(169, CLIM_MONTHS, yclsst))
      yclsst.assign((size_t)CLIM_MONTHS * NHOR, std::isnan(cfg.sea_surface_k) ? 0.0 : cfg.sea_surface_k);
    for (auto& v : yclsst) v = std::max(v, tfreeze);
    std::vector<double> yclsst2 = monthly(yclsst, step + 1, cal, NHOR);
    // seaini on sea points
    double* dicec = slot(SLOT_dicec);
    double* diced = slot(SLOT_diced);
    double* dmld = slot(SLOT_dmld);
    double* sea_dts = slot(SLOT_sea_dts);
    double* sea_dqs = slot(SLOT_sea_dqs);
    double* sea_dsst = slot(SLOT_sea_dsst);
    for (int h = 0; h < NHOR; h++) {
      if (dls[h] >= 0.5) continue;
      const double ts = yclsst2[h];
      sea_dts[h] = ts;
      dicec[h] = xicec[h];
      diced[h] = xiced[h];
      dsnow[h] = 0;
      sea_dsst[h] = ts;
      dmld[h] = mldepth;
      dt[surf + h] = ts;
      sea_dqs[h] = saturation(ts, p);
      dq[surf + h] = sea_dqs[h];
      const double c = dicec[h];
      drhs[h] = sea.drhssea * (1 - c) + sea.drhsice * c;
      dalb[h] = sea.albsea * (1 - c) + c * std::min(sea.albice, 0.5 + 0.025 * (273 - ts));
      dsalb[h] = doceanalb[0] * (1 - c) + c * std::min(dicealbmx[0], dicealbmn[0] + 0.025 * (273 - ts));
      dsalb[NHOR + h] = doceanalb[1] * (1 - c) + c * std::min(dicealbmx[1], dicealbmn[1] + 0.025 * (273 - ts));
      dz0[h] = sea.dz0sea * (1 - c) + sea.dz0ice * c;
    }
    for (int h = 0; h < NHOR; h++) {
      slot(SLOT_xts)[h] = yclsst2[h];
      slot(SLOT_xsst)[h] = yclsst2[h];
      slot(SLOT_xmld)[h] = mldepth;
      slot(SLOT_xls)[h] = xls[h];
      slot(SLOT_xiced)[h] = xiced[h];
      slot(SLOT_xicec)[h] = xicec[h];
      slot(SLOT_xclsst2)[h] = xclsst2[h];
      slot(SLOT_xclicec2)[h] = xclicec2[h];
      slot(SLOT_xcliced2)[h] = xcliced2[h];
      slot(SLOT_xclssto)[h] = xclssto[h];
      slot(SLOT_ysst)[h] = yclsst2[h];
      slot(SLOT_ymld)[h] = mldepth;
      slot(SLOT_yls)[h] = xls[h];
      slot(SLOT_yclsst2)[h] = yclsst2[h];
    }
    std::copy(xclsst.begin(), xclsst.end(), slot(SLOT_xclsst));
    std::copy(xclicec.begin(), xclicec.end(), slot(SLOT_xclicec));
    std::copy(xcliced.begin(), xcliced.end(), slot(SLOT_xcliced));
    std::copy(yclsst.begin(), yclsst.end(), slot(SLOT_yclsst));
  }

  // Spectral orography so (oroini: FOURIER.analyse + LEGENDRE.analyse with
  // skgpsp, / cv^2, truncated above noromax).
  std::vector<double> so(NRSP, 0.0);
  grid_to_spectral(m, doro.data(), so.data(), true);
  for (auto& v : so) v /= p.cv * p.cv;
  {
    int r = 0;
    for (int mm = 0; mm <= T; mm++)
      for (int n = mm; n <= T; n++, r += 2)
        if (n > noromax) so[r] = so[r + 1] = 0;
  }
  std::copy(so.begin(), so.end(), tab(TAB_sop));
  const std::vector<double> sr = setzt(dyn, p, T);
  std::copy(sr.begin(), sr.end(), tab(TAB_srp));

  if (cfg.initial_state == 1) {
    // The reference's fresh start (setup/index.ts:434-449): initfd after
    // setzt (setup/state/index.ts:60-86), then ln(ps) from the orography
    // when nspinit > 0.
    double* sp = slot(SLOT_sp);
    if (cfg.nspinit > 0) {
      const double scale = p.cv * p.cv / (p.gascon * p.tgr);
      for (int i = 0; i < NRSP; i++) sp[i] = -so[i] * scale;
    } else {
      for (int i = 0; i < NRSP; i++) sp[i] = 0;
      initfd_noise(sp, T, cfg.kick, initrandom_seed(cfg));
    }
    std::copy(sp, sp + NRSP, slot(SLOT_spm));
    for (int l = 0; l < L; l++) {
      double* st = slot(SLOT_st) + (size_t)l * NRSP;
      double* sz = slot(SLOT_sz) + (size_t)l * NRSP;
      double* sq = slot(SLOT_sq) + (size_t)l * NRSP;
      double* sd = slot(SLOT_sd) + (size_t)l * NRSP;
      for (int i = 0; i < NRSP; i++) st[i] = sz[i] = sq[i] = sd[i] = 0;
      if (p.dttl != 0)
        std::copy(sr.begin() + (size_t)l * NRSP, sr.begin() + (size_t)(l + 1) * NRSP, st);
      else
        st[0] = sr[(size_t)l * NRSP];
      sz[2] = p.plavor;
    }
    const size_t n = (size_t)L * NRSP;
    std::copy(slot(SLOT_st), slot(SLOT_st) + n, slot(SLOT_stm));
    std::copy(slot(SLOT_sz), slot(SLOT_sz) + n, slot(SLOT_szm));
    std::copy(slot(SLOT_sq), slot(SLOT_sq) + n, slot(SLOT_sqm));
    std::copy(slot(SLOT_sd), slot(SLOT_sd) + n, slot(SLOT_sdm));
  } else {
  // Spectral initial state: the isothermal atmosphere at rest. ln(ps)
  // follows the orography (initfd with nspinit > 0, setup/index.ts:
  // 434-449); temperature is initial_temperature plus seeded grid-point
  // noise; humidity is initial_humidity in the model's q * ps units;
  // vorticity is the planetary solid-body mode (setup/state/index.ts:68-74).
  {
    double* sp = slot(SLOT_sp);
    const double scale = p.cv * p.cv / (p.gascon * p.tgr);
    for (int i = 0; i < NRSP; i++) sp[i] = -so[i] * scale;
    std::copy(sp, sp + NRSP, slot(SLOT_spm));
    std::vector<double> lnps(NHOR);
    spectral_to_grid(m, sp, lnps.data());
    Rng rng(cfg.noise_seed);
    std::vector<double> gridv(NHOR), spec(NRSP);
    for (int l = 0; l < L; l++) {
      double* st = slot(SLOT_st) + (size_t)l * NRSP;
      double* sz = slot(SLOT_sz) + (size_t)l * NRSP;
      double* sq = slot(SLOT_sq) + (size_t)l * NRSP;
      for (int h = 0; h < NHOR; h++) gridv[h] = (cfg.initial_temperature + cfg.noise_kelvin * rng.normal()) / p.ct - dyn.t0[l];
      grid_to_spectral(m, gridv.data(), spec.data(), false);
      for (int i = 0; i < NRSP; i++) st[i] = spec[i];
      for (int h = 0; h < NHOR; h++) gridv[h] = cfg.initial_humidity * std::exp(lnps[h]);
      grid_to_spectral(m, gridv.data(), spec.data(), false);
      for (int i = 0; i < NRSP; i++) sq[i] = spec[i];
      sz[2] = p.plavor;
    }
    const size_t n = (size_t)L * NRSP;
    std::copy(slot(SLOT_st), slot(SLOT_st) + n, slot(SLOT_stm));
    std::copy(slot(SLOT_sz), slot(SLOT_sz) + n, slot(SLOT_szm));
    std::copy(slot(SLOT_sq), slot(SLOT_sq) + n, slot(SLOT_sqm));
  }
  }
  m.initial = A;

  // ------------------------------------------------- constants per phase
  const double meanHeight = so[0] * p.cv * p.cv / std::sqrt(2.0);
  const double psurf = std::exp(std::log(p.psurf) - meanHeight / p.gascon / p.tgr);
  const int startup = std::max(p.nkits, 1);
  m.startup_steps = startup;
  m.constants.clear();
  m.bm1.clear();
  auto phase = [&](int kits) {
    std::array<double, CON_COUNT> k{};
    const double deltsec = p.day_24hr / p.mtspd / std::pow(2.0, kits);
    const double delt = TWOPI / p.ntspd / std::pow(2.0, kits);
    const bool coupled = kits == 0;
    k[CON_psurf] = psurf;
    k[CON_ww] = p.ww;
    k[CON_cv] = p.cv;
    k[CON_ct] = p.ct;
    k[CON_akap] = p.akap;
    k[CON_rdbrv] = p.rdbrv;
    k[CON_adv] = p.adv;
    k[CON_acpd] = p.acpd;
    k[CON_ga] = p.ga;
    k[CON_gascon] = p.gascon;
    k[CON_plavor] = p.plavor;
    k[CON_delt] = delt;
    k[CON_delt2] = delt + delt;
    k[CON_deltsec] = deltsec;
    k[CON_deltsec2] = deltsec + deltsec;
    k[CON_pnu] = p.pnu;
    k[CON_pnu21] = p.pnu21;
    k[CON_filter] = coupled ? 1 : 0;
    k[CON_ra1] = p.ra1;
    k[CON_ra2] = p.ra2;
    k[CON_ra4] = p.ra4;
    k[CON_tmelt] = p.tmelt;
    k[CON_alv] = p.alv;
    k[CON_als] = p.als;
    k[CON_solar_day] = p.solar_day;
    k[CON_zumin] = zumin;
    k[CON_vdiff_b] = vdiff_b;
    k[CON_vdiff_c] = vdiff_c;
    k[CON_vdiff_d] = vdiff_d;
    k[CON_ndheat] = p.ndheat > 0 ? 1 : 0;
    // radiation (physics/radiation/index.ts:151-234)
    k[CON_rad_synchronous] = rad.nfixed;
    k[CON_gsol0] = rad.gsol0;
    k[CON_zsolar1] = rad.zsolar1;
    k[CON_zsolar2] = rad.zsolar2;
    k[CON_zsolars1] = rad.zsolars1;
    k[CON_zsolars2] = rad.zsolars2;
    k[CON_tswr1] = rad.tswr1;
    k[CON_tswr2] = rad.tswr2;
    k[CON_tswr3] = rad.tswr3;
    const double zmu00 = 0.5;
    k[CON_zb3] = rad.tswr1 * std::sqrt(zmu00) / zmu00;
    k[CON_zb4] = rad.tswr2 * std::sqrt(zmu00);
    k[CON_zb5] = rad.tswr3 * zmu00 * zmu00;
    k[CON_rcoeff] = rad.rcoeff;
    k[CON_ln_diffuse_rayleigh] = std::log(1 - 0.144);
    k[CON_nstartemp] = rad.nstartemp;
    k[CON_ozone] = rad.no3;
    k[CON_a0o3] = rad.a0o3;
    k[CON_a1o3] = rad.a1o3;
    k[CON_aco3] = rad.aco3;
    k[CON_bo3] = rad.bo3;
    k[CON_co3] = rad.co3;
    k[CON_zconst_o3] = std::exp(-rad.bo3 / rad.co3);
    k[CON_zdawn] = std::sin(rad.dawn * PI / 180);
    k[CON_sbk] = 5.67e-8;
    k[CON_th2oc] = rad.th2oc;
    k[CON_tpofmt] = rad.tpofmt;
    const double zao30 = 0.209 * std::pow(7e-5, 0.436);
    const double zco20 = 0.0676 * std::pow(0.01022, 0.421);
    const double zh2o0a = 0.846 * std::pow(3.59e-5, 0.243);
    const double zh2o0 = 0.832 * std::pow(0.0286, 0.26);
    k[CON_zao30] = zao30;
    k[CON_zco20] = zco20;
    k[CON_zh2o0a] = zh2o0a;
    k[CON_zh2o0] = zh2o0;
    k[CON_zao3c] = 0.209 * std::pow(0.01 + 7e-5, 0.436) - zao30 - 0.0212 * std::log10(0.01);
    k[CON_zaco2c] = 0.0676 * std::pow(1.01022, 0.421) - zco20;
    k[CON_zah2oc] = 0.846 * std::pow(0.01 + 3.59e-5, 0.243) - zh2o0a - 0.24 * std::log10(0.02);
    k[CON_zth2oc] = 1 - (0.832 * std::pow(2 + 0.0286, 0.26) - zh2o0) + 0.1196 * std::log(2 - 0.6931);
    k[CON_zfco2_zzf2] = (100 / 1.9635) * (0.0440098 / 0.0289644) * 1e-6;
    // rainmod
    k[CON_rhbeta] = rain.rhbeta;
    k[CON_nbeta] = rain.nbeta;
    k[CON_pdeep] = rain.pdeep;
    k[CON_pdeepth] = rain.pdeepth;
    k[CON_rkshallow] = rain.rkshallow;
    k[CON_gamma] = rain.gamma;
    k[CON_zrfac] = p.solar_day * 1000;
    k[CON_clwfac] = rain.clwfac;
    k[CON_clwcrit1] = rain.clwcrit1;
    k[CON_clwcrit2] = rain.clwcrit2;
    // landmod (setup/modules/index.ts:134-195)
    k[CON_sicecap] = land.sicecap;
    k[CON_soilcap] = land.soilcap;
    k[CON_sicediff] = land.sicediff;
    k[CON_soildiff] = land.soildiff;
    k[CON_snowcap] = land.snowcap;
    k[CON_snowdiff] = land.snowdiff;
    k[CON_rhosnow] = land.rhosnow;
    k[CON_dztop] = land.dztop;
    k[CON_dsmax] = land.dsmax;
    k[CON_albsmax] = land.albsmax;
    k[CON_albsmin] = land.albsmin;
    k[CON_albsmaxf] = land.albsmaxf;
    k[CON_albsminf] = land.albsminf;
    k[CON_albsmax1] = dsnowalbmx[0];
    k[CON_albsmax2] = dsnowalbmx[1];
    k[CON_albsmaxf1] = 0.5 * dsnowalbmx[0];
    k[CON_albsmaxf2] = 0.5 * dsnowalbmx[1];
    k[CON_albsmin1] = dsnowalbmn[0];
    k[CON_albsmin2] = dsnowalbmn[1];
    k[CON_albsminf1] = 0.75 * 0.5 * dsnowalbmx[0];
    k[CON_albsminf2] = 0.75 * 0.5 * dsnowalbmx[1];
    k[CON_albgmax] = land.albgmax;
    k[CON_albgmin] = land.albgmin;
    k[CON_albgmax1] = dsnowalbmx[0];
    k[CON_albgmax2] = dsnowalbmx[1];
    k[CON_albgmin1] = dglacalbmn[0];
    k[CON_albgmin2] = dglacalbmn[1];
    k[CON_drhsfull] = land.drhsfull;
    k[CON_nwetsoil] = land.nwetsoil;
    k[CON_wsmax] = std::max(land.wsmax, 0.0);
    k[CON_albland] = land.albland;
    k[CON_alblandmax] = land.alblandmax;
    k[CON_dgroundalb1] = dgroundalb[0];
    k[CON_dgroundalb2] = dgroundalb[1];
    k[CON_doceanalb1] = doceanalb[0];
    k[CON_doceanalb2] = doceanalb[1];
    k[CON_nveg] = p.nveg;
    k[CON_nglacier] = nglacier > 0 ? 1 : 0;
    k[CON_glacelim] = glacelim;
    // simba's vegstep (setup/index.ts:256-284): the namelist's light-use
    // efficiency, CO2 conversion and time scales; atan(cveg_d) and
    // atan(-cveg_f) in FP64
    k[CON_rlue] = land.rlue;
    k[CON_co2conv] = land.co2conv;
    k[CON_tau_veg] = land.tau_veg;
    k[CON_tau_soil] = land.tau_soil;
    k[CON_ncveg] = veg.ncveg;
    k[CON_cveg_e] = std::atan(-1.9);
    k[CON_cveg_g] = std::atan(-3.0);
    // icemod (setup/modules/index.ts:212-232; setup/index.ts:28-40)
    k[CON_xdt] = p.solar_day / p.ntspd;
    k[CON_xmind] = ice.xmind;
    k[CON_xmaxd] = ice.xmaxd;
    k[CON_cicemin] = ice.cicemin;
    k[CON_thicec] = ice.thicec;
    k[CON_tfreeze] = ice.nseaice == 0 ? 0 : ice.tfreeze;
    k[CON_crhoi] = 920;
    k[CON_crhos] = 1030;
    k[CON_crhosn] = 330;
    k[CON_cpi] = 2070;
    k[CON_ckapi] = 2.03;
    k[CON_ckapsn] = 0.31;
    k[CON_clfi] = 3.28e5;
    k[CON_clfsn] = 3.337e5;
    k[CON_ice_coupled] = coupled ? 1 : 0;
    k[CON_dtmix] = p.solar_day / p.ntspd;
    k[CON_cps] = 4180;
    k[CON_charnock] = sea.charnock;
    k[CON_dz0sea] = sea.dz0sea;
    k[CON_dz0ice] = sea.dz0ice;
    k[CON_drhssea] = sea.drhssea;
    k[CON_drhsice] = sea.drhsice;
    k[CON_albsea] = sea.albsea;
    k[CON_albice] = sea.albice;
    k[CON_dicealbmx1] = dicealbmx[0];
    k[CON_dicealbmx2] = dicealbmx[1];
    k[CON_dicealbmn1] = dicealbmn[0];
    k[CON_dicealbmn2] = dicealbmn[1];
    // carbonmod
    k[CON_ncarbon] = carbon.ncarbon;
    k[CON_nsupply] = carbon.nsupply;
    k[CON_timeweight] = timeweight;
    k[CON_co2earth] = 330;
    k[CON_cbeta] = carbon.beta;
    k[CON_kact] = carbon.kact;
    k[CON_pearth] = carbon.pearth;
    k[CON_wmax] = carbon.wmax;
    k[CON_vearth] = carbon.vearth;
    k[CON_plarad] = p.plarad;
    k[CON_rad_earthsq] = 6371220.0 * 6371220.0;
    // hurricanemod (setup/index.ts:107-114: hurricane_namelist defaults)
    k[CON_nstormdiag] = cfg.storm_diagnostics != 0 ? 1 : 0;
    k[CON_ckcd] = 0.9;
    k[CON_vithresh] = 0.145;
    // snapshot (snapshot/index.ts:26-54): level nearest sigma 0.25
    int jet = 0;
    for (int l = 1; l < L; l++)
      if (std::fabs(sigma[l] - 0.25) < std::fabs(sigma[jet] - 0.25)) jet = l;
    k[CON_jet_level] = jet;
    k[CON_sigma_bottom] = sigma[L - 1];
    m.constants.push_back(k);
    m.bm1.push_back(makebm(dyn, T, delt));
  };
  for (int i = 0; i < startup; i++) phase(startup - i);
  phase(0);
  // The tables hold the first phase's bm1 (setup/index.ts:470-474).
  std::copy(m.bm1[0].begin(), m.bm1[0].end(), tab(TAB_bm1));
  (void)count_of;
}

void step_values(const Member& m, int64_t step, int64_t carbon_count, int publish, double* out) {
  const Calendar& c = m.cal;
  // SOLAR.position (calendar/solar/index.ts:25-93)
  const double calday = (double)(step % c.steps_per_year) / c.steps_per_year;
  const double minutes = std::trunc((double)(step % c.steps_per_day) * c.step_minutes);
  const double e = c.eccentricity;
  double lon, dist;
  if (c.keplerian) {
    // Kepler's equation by Newton's method (calendar/solar/index.ts:13-24,
    // :44-66)
    double anomaly_true = 0;
    dist = 1;
    if (!c.synchronous) {
      double mean = calday * TWOPI + c.mean_anomaly_rad;
      while (mean > TWOPI) mean -= TWOPI;
      while (mean < 0) mean += TWOPI;
      if (e > 0) {
        double ea = e < 0.5 ? mean : PI;
        for (int it = 0; it < 101; it++) {
          const double prev = ea;
          ea -= (ea - (mean + e * std::sin(ea))) / std::max(1e-15, 1 - e * std::cos(ea));
          if (std::fabs(ea - prev) <= 1e-14) break;
        }
        const double tangent = std::tan(ea * 0.5);
        const double magnitude = std::sqrt(((1 + e) / (1 - e)) * tangent * tangent);
        anomaly_true = 2 * std::atan(tangent < 0 ? -magnitude : magnitude);
        if (anomaly_true < 0) anomaly_true += TWOPI;
        anomaly_true = std::fmod(anomaly_true, TWOPI);
        const double inverse = 1 / (1 - e * std::cos(ea));
        dist = inverse * inverse;
      } else {
        anomaly_true = mean;
      }
    }
    lon = std::fmod(c.perihelion_lon_rad + anomaly_true, TWOPI);
  } else {
    const double ml = c.mean_lon_equinox_rad + (calday - 80.5 / 365) * 2 * M_PI;
    const double an = ml - c.perihelion_lon_rad;
    const double s = std::sin(an);
    lon = ml + e * (2 * s + e * (1.25 * std::sin(2 * an) + e * (13.0 / 12 * std::sin(3 * an) - 0.25 * s)));
    const double inv = (1 + e * std::cos(lon - c.perihelion_lon_rad)) / (1 - e * e);
    dist = inv * inv;
  }
  const double decl = std::asin(std::sin(c.obliquity_rad) * std::sin(lon));
  const double ra = std::atan2(std::cos(c.obliquity_rad) * std::sin(lon), std::cos(lon));
  // RADIATION.step (physics/radiation/index.ts:87-105)
  double phase = c.synchronous ? (1 - c.fixed_longitude / 360) * TWOPI
                               : minutes * c.rotation_speed * TWOPI / 1440;
  out[0] = std::sin(decl);
  out[1] = std::cos(decl);
  out[2] = phase - (c.keplerian ? ra : 0);
  out[3] = dist;
  out[4] = std::cos(TWOPI * (calday - c.toffo3));
  // SURFACE.step (physics/surface/index.ts:50-60)
  const MonthInterp now = month_interp(step, c), next = month_interp(step + 1, c);
  out[5] = now.month;
  out[6] = now.neighbour;
  out[7] = now.weight;
  out[8] = next.month;
  out[9] = next.neighbour;
  out[10] = next.weight;
  const int y0 = date_of(step, c.steps_per_day, c.days_per_year, c.step_minutes).year;
  const int y1 = date_of(step + 1, c.steps_per_day, c.days_per_year, c.step_minutes).year;
  out[11] = y1 != y0 ? 1 : 0;
  // CARBON.cycle (physics/carbon/index.ts:27-32)
  out[12] = ((carbon_count + 1) % c.carbon_interval) == 0 ? 1 : 0;
  // snapshot_step (stepper/index.ts:226-229, :322)
  out[13] = (double)step;
  out[14] = c.output_window;
  out[15] = publish;
}

}  // namespace exo
