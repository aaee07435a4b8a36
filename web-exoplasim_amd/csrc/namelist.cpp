// The namelist files of a run and their Fortran-style read, host only.
//
// GCM_NAMELIST.parse / read (src/model/climate/gcm/namelist/index.ts:
// 196-233) restated: a file's "KEY = value" lines, keys upper-cased, the
// later of two lines for one key winning; a value is a comma list of
// numbers (a Fortran 'd' exponent reads as 'e'), and a key's defaults are
// overwritten from the front.
//
// exo_config_from_namelists reads the files as the set-up does
// (setup/pumamod/index.ts:103-232, setup/modules/index.ts:41-305,
// setup/index.ts:107-112, each module's defaults for an absent key) and
// turns them back into the exo_config that GCM_NAMELIST.edits
// (namelist/index.ts:40-182) writes them from. A key the set-up reads that
// no exo_config field carries must hold the value this engine's set-up
// runs with (the base namelists' or the module's); another value is
// refused with the key named.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/exoplasim_mi355x.h"
#include "setup.h"

namespace exo {
int set_error(int code, const std::string& message);
}
using exo::PlasimArrays;
using exo::plasim_arrays;

namespace {

using Values = std::map<std::string, std::string>;

bool space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && space(s[a])) a++;
  while (b > a && space(s[b - 1])) b--;
  return s.substr(a, b - a);
}

// parse (namelist/index.ts:196-206): every line of the form
// `<spaces> NAME <spaces> = <spaces> value <spaces>`, NAME a letter or
// underscore then letters, digits and underscores; other lines (the group
// header, /END, blank lines) are skipped.
Values parse(const char* text) {
  Values out;
  if (!text) return out;
  const char* p = text;
  while (true) {
    const char* end = std::strchr(p, '\n');
    std::string line = end ? std::string(p, end - p) : std::string(p);
    if (!line.empty() && line.back() == '\r') line.pop_back();  // the split's \r?\n
    size_t i = 0;
    while (i < line.size() && space(line[i])) i++;
    if (i < line.size() && (std::isalpha((unsigned char)line[i]) || line[i] == '_')) {
      size_t j = i + 1;
      while (j < line.size() && (std::isalnum((unsigned char)line[j]) || line[j] == '_')) j++;
      size_t k = j;
      while (k < line.size() && space(line[k])) k++;
      if (k < line.size() && line[k] == '=') {
        std::string key = line.substr(i, j - i);
        for (char& c : key) c = (char)std::toupper((unsigned char)c);
        out[key] = trim(line.substr(k + 1));
      }
    }
    if (!end) break;
    p = end + 1;
  }
  return out;
}

bool digits(const std::string& s, size_t& i, int base) {
  const size_t start = i;
  while (i < s.size()) {
    const char c = (char)std::tolower((unsigned char)s[i]);
    const int v = std::isdigit((unsigned char)c) ? c - '0' : (c >= 'a' && c <= 'z') ? c - 'a' + 10 : 99;
    if (v >= base) break;
    i++;
  }
  return i > start;
}

// number (namelist/index.ts:208-210): JavaScript's Number() of the trimmed
// text with its first 'd' or 'D' read as 'e': empty text is 0; decimal
// literals with an optional sign and exponent, 0x / 0o / 0b integers and
// [+-]Infinity read; anything else is NaN.
double number(const std::string& raw) {
  std::string s = trim(raw);
  for (char& c : s)
    if (c == 'd' || c == 'D') {
      c = 'e';
      break;
    }
  if (s.empty()) return 0;
  if (s == "Infinity" || s == "+Infinity") return INFINITY;
  if (s == "-Infinity") return -INFINITY;
  if (s.size() > 2 && s[0] == '0' && std::strchr("xXoObB", s[1])) {
    const int base = (s[1] == 'x' || s[1] == 'X') ? 16 : (s[1] == 'o' || s[1] == 'O') ? 8 : 2;
    size_t i = 2;
    if (!digits(s, i, base) || i != s.size()) return NAN;
    double v = 0;
    for (size_t k = 2; k < s.size(); k++) {
      const char c = (char)std::tolower((unsigned char)s[k]);
      v = v * base + (std::isdigit((unsigned char)c) ? c - '0' : c - 'a' + 10);
    }
    return v;
  }
  size_t i = 0;
  if (s[i] == '+' || s[i] == '-') i++;
  const bool whole = digits(s, i, 10);
  bool frac = false;
  if (i < s.size() && s[i] == '.') {
    i++;
    frac = digits(s, i, 10);
  }
  if (!whole && !frac) return NAN;
  if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
    i++;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) i++;
    if (!digits(s, i, 10)) return NAN;
  }
  if (i != s.size()) return NAN;
  return std::strtod(s.c_str(), nullptr);
}

// read (namelist/index.ts:212-233) of one key: the defaults with the
// value's comma-separated numbers written over them from the front. Parts
// beyond the defaults' length are dropped (the set-up reads a fixed
// number of elements).
std::vector<double> read(const Values& v, const std::string& key, std::vector<double> value) {
  auto it = v.find(key);
  if (it == v.end()) return value;
  const std::string& text = it->second;
  size_t at = 0, i = 0;
  while (true) {
    const size_t comma = text.find(',', at);
    const std::string part = text.substr(at, comma == std::string::npos ? std::string::npos : comma - at);
    if (i < value.size()) value[i] = number(part);
    i++;
    if (comma == std::string::npos) break;
    at = comma + 1;
  }
  return value;
}

double read1(const Values& v, const std::string& key, double fallback) {
  return read(v, key, {fallback})[0];
}

std::string show(const std::vector<double>& a) {
  std::string s;
  char buf[40];
  for (size_t i = 0; i < a.size(); i++) {
    std::snprintf(buf, sizeof(buf), "%.17g", a[i]);
    s += (i ? "," : "") + std::string(buf);
  }
  return s;
}

struct Files {
  std::map<std::string, Values> files;
  const Values& operator[](const std::string& name) const {
    static const Values empty;
    auto it = files.find(name);
    return it == files.end() ? empty : it->second;
  }
};

[[noreturn]] void refuse(const std::string& file, const std::string& key, const std::vector<double>& got,
                         const std::vector<double>& runs) {
  throw std::runtime_error(file + ": " + key + " = " + show(got) + " is not built (this engine runs " +
                           show(runs) + ")");
}

// A key this engine holds at `runs`, read with the module's default.
void hold(const Files& f, const std::string& file, const std::string& key, const std::vector<double>& module_default,
          const std::vector<double>& runs) {
  const std::vector<double> got = read(f[file], key, module_default);
  if (got != runs) refuse(file, key, got, runs);
}

void hold1(const Files& f, const std::string& file, const std::string& key, double module_default, double runs) {
  hold(f, file, key, {module_default}, {runs});
}

// x with x * factor == target where the division's rounding missed it by
// an ulp or two (the set-up multiplies the exo_config field back).
double undo_scale(double target, double factor) {
  const double x = target / factor;
  if (x * factor == target) return x;
  double lo = x, hi = x;
  for (int i = 0; i < 4; i++) {
    lo = std::nextafter(lo, -INFINITY);
    hi = std::nextafter(hi, INFINITY);
    if (lo * factor == target) return lo;
    if (hi * factor == target) return hi;
  }
  return x;
}

double undo_reciprocal(double target) {
  const double x = 1 / target;
  if (1 / x == target) return x;
  double lo = x, hi = x;
  for (int i = 0; i < 4; i++) {
    lo = std::nextafter(lo, -INFINITY);
    hi = std::nextafter(hi, INFINITY);
    if (1 / lo == target) return lo;
    if (1 / hi == target) return hi;
  }
  return x;
}

// An albedo override: absent from every key, or the same value in each
// (two-band keys both bands).
double albedo_override(const Files& f, const std::vector<std::pair<std::string, std::string>>& keys) {
  int present = 0;
  double value = NAN;
  for (const auto& fk : keys) {
    const Values& v = f[fk.first];
    if (!v.count(fk.second)) continue;
    present++;
    const bool bands = fk.second[0] == 'D';
    const std::vector<double> got = read(v, fk.second, bands ? std::vector<double>{NAN, NAN} : std::vector<double>{NAN});
    if (std::isnan(value)) value = got[0];
    for (double g : got)
      if (!(g == value))
        throw std::runtime_error(fk.first + ": " + fk.second + " = " + show(got) +
                                 " is not built (one albedo for " + keys[0].second + " and its bands)");
  }
  if (present == 0) return NAN;
  if (present != (int)keys.size())
    throw std::runtime_error(keys[0].first + ": " + keys[0].second +
                             " and its band keys are set together or not at all in this engine");
  return value;
}

void from_namelists(int T, int L, const Files& f, exo_config& c) {
  if (T != 21 && T != 42 && T != 63) throw std::runtime_error("T" + std::to_string(T) + " is not built");
  if (L != 5 && L != 10) throw std::runtime_error("Choose 5 or 10 atmospheric layers");
  c.truncation = T;
  c.levels = L;

  // planet_namelist over p_earth's planet_ini (setup/pumamod/index.ts:18-40)
  const std::string PL = "planet_namelist";
  const Values& planet = f[PL];
  c.gravity = read1(planet, "GA", 9.80665);
  c.radius_earth = undo_scale(read1(planet, "PLARAD", 6371220), 6371220);
  c.flux_wm2 = read1(planet, "GSOL0", 1365);
  c.eccentricity = read1(planet, "ECCEN", 0.016715);
  c.obliquity_deg = read1(planet, "OBLIQ", 23.441);
  c.fixed_orbit = read1(planet, "NFIXORB", 0) != 0;
  c.keplerian = read1(planet, "NGENKEPLERIAN", 0) != 0;
  c.gas_constant = read1(planet, "GASCON", 287);
  const double sidereal_year = read1(planet, "SIDEREAL_YEAR", 31558149);
  // MVELP at p_earth's 102.7 (the base namelists') is the null periapse
  const double mvelp = read1(planet, "MVELP", 102.7);
  c.periapse_longitude_deg = mvelp == 102.7 ? NAN : mvelp;
  c.initial_mean_anomaly_deg = planet.count("MEANANOM0") ? read1(planet, "MEANANOM0", -2.9589) : NAN;
  const double rotspd = read1(planet, "ROTSPD", 1);
  c.rotation_period_days = rotspd == 1 ? 1 : undo_reciprocal(rotspd);
  for (const char* key : {"AKAP", "ALR", "RA1", "RA2", "RA4", "PNU"}) {
    static const std::map<std::string, double> earth = {{"AKAP", 0.286}, {"ALR", 0.0065}, {"RA1", 610.78},
                                                        {"RA2", 17.2693882}, {"RA4", 35.86}, {"PNU", 0.1}};
    hold1(f, PL, key, earth.at(key), earth.at(key));
  }

  // plasim_namelist over plasimmod's defaults (setup/pumamod/index.ts:43-91)
  const std::string PS = "plasim_namelist";
  const Values& plasim = f[PS];
  const double ndpy = read1(plasim, "N_DAYS_PER_YEAR", 360);
  if (rotspd == 1) {
    // edits write N_DAYS_PER_YEAR = yearDays and SIDEREAL_YEAR = yearDays days
    c.year_days = ndpy;
    if (ndpy * 86400 != sidereal_year)
      throw std::runtime_error(PL + ": SIDEREAL_YEAR = " + show({sidereal_year}) +
                               " is not built with N_DAYS_PER_YEAR = " + show({ndpy}) +
                               " (this engine runs a year of N_DAYS_PER_YEAR days)");
  } else {
    c.year_days = undo_scale(sidereal_year, 86400);
    const double derived =
        std::max(std::floor(c.year_days / c.rotation_period_days / 12 + 0.5), 1.0) * 12;
    if (ndpy != derived) refuse(PS, "N_DAYS_PER_YEAR", {ndpy}, {derived});
  }
  c.pressure_bar = undo_scale(read1(plasim, "PSURF", 101100), 1e5);
  c.substellar_longitude_deg = read1(plasim, "FIXEDLON", 0);
  c.vegetation = (int)read1(plasim, "NVEG", 1);
  if (c.vegetation != read1(plasim, "NVEG", 1) || c.vegetation < 0 || c.vegetation > 2)
    throw std::runtime_error(PS + ": NVEG must be 0, 1 or 2");
  double mpstep = read1(plasim, "MPSTEP", 0);
  hold1(f, PS, "NTSPD", 0, 0);
  if (mpstep <= 0) mpstep = T == 21 ? 45 : T == 42 ? 30 : 20;  // readnl's default for 32, 64, 96 latitudes
  c.timestep_minutes = mpstep;
  const double nfilter = read1(plasim, "NFILTER", 0);
  if (nfilter != 0 && nfilter != 1 && nfilter != 2 && nfilter != 3)
    throw std::runtime_error(PS + ": NFILTER = " + show({nfilter}) + " is not built (0 to 3)");
  c.filter = (int)nfilter;
  const double kick = read1(plasim, "KICK", 1);
  if (kick != 0 && kick != 1 && kick != 2 && kick != 3)
    throw std::runtime_error(PS + ": KICK = " + show({kick}) + " is not built (0 to 3)");
  c.kick = (int)kick;
  const std::vector<double> seed = read(plasim, "SEED", std::vector<double>(8, 0));
  for (int i = 0; i < 8; i++) {
    if (seed[i] != std::trunc(seed[i]) || seed[i] < INT32_MIN || seed[i] > INT32_MAX)
      throw std::runtime_error(PS + ": SEED = " + show(seed) + " is not eight 32-bit integers");
    c.seed[i] = (int32_t)seed[i];
  }
  const int ndays = (int)std::lround(rotspd == 1 ? c.year_days : ndpy);
  hold1(f, PS, "N_START_YEAR", 1, ndays == 365 ? 1 : 0);
  hold1(f, PS, "N_START_MONTH", 1, 1);
  hold1(f, PS, "NSTPW", 0, std::floor(7200.0 / mpstep));
  hold1(f, PS, "NLOWIO", 0, 1);
  hold1(f, PS, "NWPD", 0, 0);
  hold1(f, PS, "NGPTFILTER", 0, 1);
  hold1(f, PS, "NSPVFILTER", 0, 1);
  struct Held {
    const char* key;
    double value;
  };
  for (const Held& h : {Held{"NKITS", 3}, Held{"NDHEAT", 1}, Held{"NQSPEC", 1}, Held{"NHORDIF", 1},
                        Held{"NEQSIG", 0}, Held{"FILTERKAPPA", 8}, Held{"NFILTEREXP", 8}, Held{"LANDHOSKN0", 15},
                        Held{"DTEP", 0}, Held{"DTNS", 0}, Held{"DTROP", 12000}, Held{"DTTRP", 2}, Held{"DTTL", 0},
                        Held{"TGR", 288}, Held{"PTOP", 7500}, Held{"FRCMOD", 1}})
    hold1(f, PS, h.key, h.value, h.value);
  const PlasimArrays a = plasim_arrays(T, L);
  hold1(f, PS, "NHDIFF", a.nhdiff, a.nhdiff);
  hold(f, PS, "NDEL", a.ndel, a.ndel);
  hold(f, PS, "TDISSD", a.tdissd, a.tdissd);
  hold(f, PS, "TDISSZ", a.tdissz, a.tdissz);
  hold(f, PS, "TDISST", a.tdisst, a.tdisst);
  hold(f, PS, "TDISSQ", a.tdissq, a.tdissq);
  hold(f, PS, "RESTIM", a.restim, a.restim);
  hold(f, PS, "T0", a.t0, a.t0);
  hold(f, PS, "TFRC", a.tfrc, a.tfrc);

  // radmod_namelist (setup/modules/index.ts:13-92)
  const std::string RM = "radmod_namelist";
  const Values& rad = f[RM];
  c.synchronous = read1(rad, "NFIXED", 0) != 0;
  c.ozone = read1(rad, "NO3", 1) != 0;
  c.co2_ppm = read1(rad, "CO2", 360);
  const double nstartemp = read1(rad, "NSTARTEMP", 0);
  const double starbbtemp = read1(rad, "STARBBTEMP", 5772);
  if (nstartemp > 0 && !(starbbtemp > 0))
    throw std::runtime_error(RM + ": STARBBTEMP = " + show({starbbtemp}) + " is not built with NSTARTEMP > 0");
  c.star_temperature_k = nstartemp > 0 ? starbbtemp : 0;
  for (const Held& h : {Held{"NDCYCLE", 1}, Held{"MINWAVEL", 316.036116751}, Held{"NSIMPLEALBEDO", 1},
                        Held{"TH2OC", 0.024}, Held{"TSWR1", 0.077}, Held{"TSWR2", 0.065}, Held{"TSWR3", 0.0055},
                        Held{"TPOFMT", 1}, Held{"ACLLWR", 0.1}, Held{"A0O3", 0.25}, Held{"A1O3", 0.11},
                        Held{"ACO3", 0.08}, Held{"BO3", 20000}, Held{"CO3", 5000}, Held{"TOFFO3", 0.25},
                        Held{"DAWN", 0}})
    hold1(f, RM, h.key, h.value, h.value);

  // rainmod_namelist and fluxmod_namelist (setup/modules/index.ts:94-135)
  const double gamma = T == 42 && L == 10 ? 0.007 : 0.01;
  for (const Held& h : {Held{"RHBETA", 0}, Held{"NBETA", 3}, Held{"PDEEP", 999999}, Held{"PDEEPTH", 70000},
                        Held{"RKSHALLOW", 10}, Held{"GAMMA", gamma}, Held{"CLWCRIT1", -0.1}, Held{"CLWCRIT2", 0},
                        Held{"RCRITMOD", 1}, Held{"RCRITSLOPE", 0}})
    hold1(f, "rainmod_namelist", h.key, h.value, h.value);
  for (const Held& h : {Held{"ZUMIN", 1}, Held{"VDIFF_LAMM", 160}, Held{"VDIFF_B", 5}, Held{"VDIFF_C", 5},
                        Held{"VDIFF_D", 5}})
    hold1(f, "fluxmod_namelist", h.key, h.value, h.value);

  // landmod_namelist and seamod_namelist (setup/modules/index.ts:137-217)
  const std::string LM = "landmod_namelist", SM = "seamod_namelist";
  const Values& land = f[LM];
  c.wet_soil = read1(land, "NWETSOIL", 0) != 0;
  static const double kSoil[5] = {0.4, 0.8, 1.6, 3.2, 6.4};
  const std::vector<double> dsoilz = read(land, "DSOILZ", {0.4, 0.8, 1.6, 3.2, 6.4});
  c.soil_depth_scale = undo_scale(dsoilz[0], kSoil[0]);
  for (int i = 0; i < 5; i++)
    if (kSoil[i] * c.soil_depth_scale != dsoilz[i])
      throw std::runtime_error(LM + ": DSOILZ = " + show(dsoilz) +
                               " is not built (the layers 0.4, 0.8, 1.6, 3.2, 6.4 m times one scale)");
  c.soil_water_capacity_m = land.count("WSMAX") ? read1(land, "WSMAX", 0.5) : NAN;
  c.soil_heat_capacity = land.count("SOILCAP") ? read1(land, "SOILCAP", 2.4e6) : NAN;
  c.land_albedo = albedo_override(f, {{LM, "ALBLAND"}, {LM, "DGROUNDALB"}});
  c.ocean_albedo = albedo_override(f, {{SM, "ALBSEA"}, {SM, "DOCEANALB"}});
  c.snow_ice_albedo = albedo_override(
      f, {{SM, "ALBICE"}, {SM, "DICEALBMN"}, {SM, "DICEALBMX"}, {LM, "DSNOWALBMN"}, {LM, "DGLACALBMN"}});
  if (land.count("DSNOWALBMX"))
    throw std::runtime_error(LM + ": DSNOWALBMX is not built (the star's band maxima are kept)");
  for (const Held& h : {Held{"ALBSMIN", 0.4}, Held{"ALBSMAX", 0.8}, Held{"ALBSMINF", 0.3}, Held{"ALBSMAXF", 0.4},
                        Held{"ALBGMIN", 0.6}, Held{"ALBGMAX", 0.8}, Held{"ALBLANDMAX", 0.4}, Held{"DRHSFULL", 0.4},
                        Held{"DZTOP", 0.2}, Held{"DSMAX", 5}, Held{"RLUE", 3.4e-10}, Held{"CO2CONV", 8.3e-4},
                        Held{"TAU_VEG", 10}, Held{"TAU_SOIL", 42}, Held{"RHOSNOW", 330}, Held{"SOILDIFF", 1.8},
                        Held{"SICEDIFF", 2.03}, Held{"SNOWDIFF", 0.31}, Held{"SICECAP", 2.07e6},
                        Held{"SNOWCAP", 0.6897e6}, Held{"RINIFOR", 0.5}, Held{"DZ0LAND", 2}, Held{"DRHSLAND", 0.25},
                        Held{"DZGLAC", -1}, Held{"NWATCINI", 0}, Held{"DWATCINI", 0}})
    hold1(f, LM, h.key, h.value, h.value);
  for (const Held& h : {Held{"DZ0SEA", 1.5e-5}, Held{"DZ0ICE", 0.001}, Held{"DRHSSEA", 1}, Held{"DRHSICE", 1},
                        Held{"CHARNOCK", 0.018}})
    hold1(f, SM, h.key, h.value, h.value);

  // icemod, vegmod, carbonmod, oceanmod, surfmod, glacier
  // (setup/modules/index.ts:219-305) and hurricane (setup/index.ts:107-112)
  for (const Held& h : {Held{"NICE", 1}, Held{"NSEAICE", 1}, Held{"XMIND", 0.1}, Held{"XMAXD", 9},
                        Held{"THICEC", 0.5}, Held{"TFREEZE", 271.25}, Held{"CICEMIN", 0.5}})
    hold1(f, "icemod_namelist", h.key, h.value, h.value);
  const Values& veg = f["vegmod_namelist"];
  c.initial_growth = read1(veg, "RINIDAGG", 0.5);
  for (const Held& h : {Held{"NCVEG", 1}, Held{"FORGROW", 1}, Held{"RINIDSC", 1}, Held{"RINIDMR", 2},
                        Held{"RINISOIL", 0}, Held{"RINIVEG", 0}})
    hold1(f, "vegmod_namelist", h.key, h.value, h.value);
  for (const Held& h : {Held{"NCARBON", 1}, Held{"NSUPPLY", 0}, Held{"KACT", 0.09}, Held{"BETA", 0.5},
                        Held{"FREQUENCY", 4}, Held{"VEARTH", 0.05}, Held{"PEARTH", 79}})
    hold1(f, "carbonmod_namelist", h.key, h.value, h.value);
  hold1(f, "carbonmod_namelist", "WMAX", 1, 80.9);  // the base namelist's 80.9
  c.ocean_depth_m = read1(f["oceanmod_namelist"], "MLDEPTH", 50);
  const Values& surf = f["surfmod_namelist"];
  const double nspinit = read1(surf, "NSPINIT", 0);
  if (nspinit != std::trunc(nspinit)) throw std::runtime_error("surfmod_namelist: NSPINIT must be an integer");
  c.nspinit = (int)nspinit;
  hold1(f, "surfmod_namelist", "NOROMAX", T, T);
  const Values& glacier = f["glacier_namelist"];
  c.glaciers = read1(glacier, "NGLACIER", 1) != 0;
  c.glacier_min_depth_m = read1(glacier, "GLACELIM", 2);
  c.glacier_initial_height_m = read1(glacier, "ICESHEETH", -1);
  const Values& storm = f["hurricane_namelist"];
  c.storm_diagnostics = read1(storm, "NSTORMDIAG", 0) > 0;
  hold1(f, "hurricane_namelist", "CKCD", 0.9, 0.9);
  hold1(f, "hurricane_namelist", "VITHRESH", 0.145, 0.145);
}

}  // namespace

extern "C" int exo_namelist_read(const char* text, const char* const* keys, const int* lengths, int nkeys,
                                 double* values) {
  if (nkeys < 0 || (nkeys > 0 && (!keys || !lengths || !values))) return exo::set_error(EXO_E_ARG, "null argument");
  try {
    const Values v = parse(text);
    size_t at = 0;
    for (int k = 0; k < nkeys; k++) {
      if (!keys[k] || lengths[k] < 1) return exo::set_error(EXO_E_ARG, "each key needs a name and a length >= 1");
      std::string key(keys[k]);
      for (char& c : key) c = (char)std::toupper((unsigned char)c);
      std::vector<double> d(values + at, values + at + lengths[k]);
      d = read(v, key, d);
      std::copy(d.begin(), d.end(), values + at);
      at += (size_t)lengths[k];
    }
  } catch (const std::exception& x) {
    return exo::set_error(EXO_E_ARG, x.what());
  }
  return EXO_OK;
}

extern "C" int exo_config_from_namelists(int truncation, int levels, const char* const* names,
                                         const char* const* texts, int count, exo_config* cfg) {
  if (!cfg || count < 0 || (count > 0 && (!names || !texts))) return exo::set_error(EXO_E_ARG, "null argument");
  try {
    Files f;
    for (int i = 0; i < count; i++) {
      if (!names[i]) return exo::set_error(EXO_E_ARG, "null namelist name");
      f.files[names[i]] = parse(texts[i]);
    }
    exo_config c = *cfg;
    from_namelists(truncation, levels, f, c);
    *cfg = c;
  } catch (const std::exception& x) {
    return exo::set_error(EXO_E_ARG, x.what());
  }
  return EXO_OK;
}
