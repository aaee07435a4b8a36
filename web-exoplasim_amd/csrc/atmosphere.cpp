// GCM_ATMOSPHERE.resolve (src/model/climate/gcm/atmosphere/index.ts:
// 84-160): a gas mix, given as partial pressures in bar, to the surface
// pressure, the gas constant R / mu and the CO2 mixing ratio the
// namelists take (namelist/index.ts:40-43). Host only.
#include <cmath>
#include <string>

#include "../../include/exoplasim_mi355x.h"

namespace exo {
int set_error(int code, const std::string& message);
}

namespace {

// The gases of GCM_ATMOSPHERE.GASES (atmosphere/index.ts:14-64), in order:
// their labels (for the messages) and molar masses (g/mol).
struct Gas {
  const char* label;
  double molar_mass;
};
const Gas kGases[9] = {
    {"Hydrogen (H₂)", 2.01588},       {"Helium (He)", 4.002602},
    {"Nitrogen (N₂)", 28.0134},       {"Oxygen (O₂)", 31.9988},
    {"Argon (Ar)", 39.948},                {"Neon (Ne)", 20.1797},
    {"Krypton (Kr)", 83.798},              {"Carbon dioxide (CO₂)", 44.01},
    {"Water vapor (H₂O)", 18.01528},
};
constexpr int kCo2 = 7;
const double kUniversal = 8314.46261815324;
// ExoPlaSim's native atmosphere: 1.011 bar of Earth air, 300 ppm CO2
const double kNativeBar = 1.011;
const double kNativeGasConstant = kUniversal / 28.970253;
const double kNativeCo2 = 300;

}  // namespace

extern "C" int exo_atmosphere_resolve(const double* partial, int pressure_and_gas_constant, double pressure_bar,
                                      double gas_constant, double* out_bar, double* out_gas_constant,
                                      double* out_co2_ppm) {
  if (!partial || !out_bar || !out_gas_constant || !out_co2_ppm) return exo::set_error(EXO_E_ARG, "null argument");
  double specified = 0, weighted = 0;
  for (int i = 0; i < 9; i++) {
    const double p = partial[i];
    if (std::isnan(p)) continue;  // null
    if (!std::isfinite(p) || p < 0)
      return exo::set_error(EXO_E_ARG,
                            std::string(kGases[i].label) + " partial pressure must be finite and nonnegative.");
    specified += p;
    weighted += p * kGases[i].molar_mass;
  }
  if (!std::isfinite(specified) || !std::isfinite(weighted))
    return exo::set_error(EXO_E_ARG, "The specified atmosphere is too large to calculate.");
  double bar = specified != 0 ? specified : kNativeBar;
  double gas = specified > 0 ? kUniversal / (weighted / specified) : kNativeGasConstant;
  if (pressure_and_gas_constant) {
    bar = std::isnan(pressure_bar) ? kNativeBar : pressure_bar;
    gas = std::isnan(gas_constant) ? kNativeGasConstant : gas_constant;
  }
  if (!std::isfinite(bar) || bar <= 0 || !std::isfinite(gas) || gas <= 0)
    return exo::set_error(EXO_E_ARG, "Total pressure and gas constant must be finite and positive.");
  if (specified > bar + 1e-12 * std::fmax(1, bar))
    return exo::set_error(EXO_E_ARG, "The specified partial pressures exceed total surface pressure.");
  const double co2 = partial[kCo2];
  *out_bar = bar;
  *out_gas_constant = gas;
  *out_co2_ppm = std::isnan(co2) ? kNativeCo2 : std::fmin(1, co2 / bar) * 1e6;
  return EXO_OK;
}
