// Field readback and global diagnostics on gfx950.
//
//  k_pack          the reference's published field set (snapshot/wgsl/
//                  pack.wgsl:14-107): tracks surface-temperature extrema
//                  every main-loop step and, on publishing steps, writes
//                  the 101 (L10) / 76 (L5) snapshot planes in
//                  fields/index.ts:81-132 order.
//  k_global_means  area-weighted global means of surface temperature and
//                  TOA net flux (GCM_GRID.globalMean, grid/index.ts:99-111),
//                  accumulated in FP64 by one workgroup per member.
#include <string>

#include "engine.h"

namespace {

// ExoPlaSim's postprocessor: hfns = snow melt * L * rho_water + rss + rls +
// hfss + hfls (snapshot/wgsl/pack.wgsl:7).
constexpr float L_TIMES_RHOH2O = -333700000.0f;

}  // namespace

extern "C" __global__ __launch_bounds__(64) void k_pack(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int h = tile * 64 + threadIdx.x;
  const int NH = M.nhor, L = M.nlev;
  if (h >= NH) return;
  float* A = dev_arena(M);
  auto S = [&](int slot) { return A + M.slot[slot]; };
  auto put = [&](int index, float value) { A[M.slot[SLOT_snapshot] + index * NH + h] = value; };
  const float* ss = S(SLOT_snapshot_step);
  const unsigned step = (unsigned)ss[0];
  const unsigned window = (unsigned)ss[1];
  const float t = S(SLOT_dt)[L * NH + h];
  float* tr = S(SLOT_snapshot_tracking) + h;
  if (tr[7 * NH] == 0.0f) {
    tr[0] = t;
    tr[NH] = t;
    tr[7 * NH] = 1.0f;
  }
  tr[0] = fminf(tr[0], t);
  tr[NH] = fmaxf(tr[NH], t);
  if (step > 0u && window > 0u && step % window == 0u) {
    const float range = tr[NH] - tr[0];
    tr[2 * NH] = tr[0];
    tr[3 * NH] = tr[NH];
    tr[4 * NH] = range;
    tr[5 * NH] += range;
    tr[6 * NH] += 1.0f;
    tr[7 * NH] = 0.0f;
    tr[8 * NH] = 1.0f;
  }
  if (ss[2] == 0.0f) return;
  const int jet = (int)M.k[CON_jet_level];
  const float* dswfl = S(SLOT_dswfl);
  const float* dlwfl = S(SLOT_dlwfl);
  put(0, t);
  put(1, S(SLOT_dtsa)[h]);
  put(2, S(SLOT_du)[(L - 1) * NH + h]);
  put(3, S(SLOT_dv)[(L - 1) * NH + h]);
  put(4, S(SLOT_dcc)[L * NH + h]);
  put(5, S(SLOT_dprc)[h] + S(SLOT_dprl)[h]);
  put(6, S(SLOT_dp)[h]);
  put(7, S(SLOT_dqvi)[h]);
  put(8, S(SLOT_dicec)[h]);
  put(9, S(SLOT_dsnow)[h]);
  put(10, S(SLOT_dwatc)[h]);
  put(11, S(SLOT_dalb)[h]);
  put(12, S(SLOT_du)[jet * NH + h]);
  put(13, S(SLOT_dv)[jet * NH + h]);
  put(14, -dlwfl[h]);
  put(15, S(SLOT_dls)[h]);
  put(16, dswfl[h] + dlwfl[h]);
  put(17, S(SLOT_dsmelt)[h] * L_TIMES_RHOH2O + dswfl[L * NH + h] + dlwfl[L * NH + h] + S(SLOT_dshfl)[h] +
              S(SLOT_dlhfl)[h]);
  put(18, -S(SLOT_devap)[h]);
  put(19, S(SLOT_drunoff)[h]);
  put(20, S(SLOT_dforest)[h]);
  put(21, S(SLOT_dlai)[h]);
  put(22, S(SLOT_dprs)[h]);
  put(23, S(SLOT_dsmelt)[h]);
  put(24, S(SLOT_diced)[h]);
  put(25, S(SLOT_dglac)[h]);
  put(27, dswfl[L * NH + h]);
  put(28, dlwfl[L * NH + h]);
  put(29, S(SLOT_dshfl)[h]);
  put(30, S(SLOT_dlhfl)[h]);
  put(31, S(SLOT_dprc)[h]);
  put(32, S(SLOT_dprl)[h]);
  put(33, dswfl[h]);
  put(34, dlwfl[h]);
  const bool windowed = tr[8 * NH] != 0.0f;
  put(35, windowed ? tr[2 * NH] : tr[0]);
  put(36, windowed ? tr[3 * NH] : tr[NH]);
  put(37, windowed ? tr[4 * NH] : tr[NH] - tr[0]);
  put(38, tr[5 * NH]);
  put(39, tr[6 * NH]);
  put(40, S(SLOT_dnpp)[h]);
  put(41, S(SLOT_gmu0)[h]);
  tr[5 * NH] = 0.0f;
  tr[6 * NH] = 0.0f;
  for (int l = 0; l < L; l++) {
    const int field = 51 + 5 * l;
    put(field, S(SLOT_dt)[l * NH + h]);
    put(field + 1, S(SLOT_du)[l * NH + h]);
    put(field + 2, S(SLOT_dv)[l * NH + h]);
    put(field + 4, S(SLOT_dcc)[l * NH + h]);
  }
  put(42, S(SLOT_capen)[h]);
  put(43, S(SLOT_lnb)[h]);
  put(44, S(SLOT_mpoti)[h]);
  put(45, S(SLOT_chim)[h]);
  put(46, S(SLOT_laav)[h]);
  put(47, S(SLOT_venti)[h]);
  put(48, S(SLOT_vrmpi)[h]);
  put(49, S(SLOT_gpi)[h]);
  put(50, S(SLOT_shear)[h]);
  // relative humidity of the lowest level (saturation.wgsl:21-24)
  const int lowest = L - 1;
  const float tl = S(SLOT_dt)[lowest * NH + h];
  const float p = S(SLOT_dp)[h] * M.k[CON_sigma_bottom];
  const float rdbrv = M.k[CON_rdbrv];
  float qs = fminf(rdbrv * M.k[CON_ra1] * expf(M.k[CON_ra2] * (tl - M.k[CON_tmelt]) / (tl - M.k[CON_ra4])) / p, rdbrv);
  qs = qs / (1.0f - (1.0f / rdbrv - 1.0f) * qs);
  put(51 + 5 * lowest + 3, fmaxf(0.0f, S(SLOT_dq)[lowest * NH + h]) / fmaxf(1.17549435e-38f, qs));
}

// out[2 * i] = mean surface temperature, out[2 * i + 1] = mean TOA net flux
// of planned member i, FP64 accumulation in a fixed order.
extern "C" __global__ __launch_bounds__(256) void k_global_means(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                                  double* __restrict__ out) {
  const DevMember& M = mems[plan.member[blockIdx.x]];
  const int NH = M.nhor, L = M.nlev;
  const float* ts = dev_arena(M) + M.slot[SLOT_dt] + L * NH;
  const float* sw = dev_arena(M) + M.slot[SLOT_dswfl];
  const float* lw = dev_arena(M) + M.slot[SLOT_dlwfl];
  const float* w = dev_tab(M) + M.tabo[TAB_zgw];
  __shared__ double s0[256], s1[256], s2[256];
  double a = 0, b = 0, c = 0;
  for (int h = threadIdx.x; h < NH; h += 256) {
    const double wh = w[h];
    a += wh * ts[h];
    b += wh * ((double)sw[h] + (double)lw[h]);
    c += wh;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  s2[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s0[threadIdx.x] += s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = s0[0] / s2[0];
    out[2 * blockIdx.x + 1] = s1[0] / s2[0];
  }
}

// The published fields' ids in plane order (GCM_FIELDS ORDER, fields/
// index.ts:81-128): 42 surface fields, the 9 storm fields, then five per
// level from the top.
extern "C" const char* exo_snapshot_field_name(int levels, int index) {
  static const char* const surface[] = {
      "surfaceTemperature", "airTemperature", "windU", "windV", "cloudCover", "precipitation",
      "surfacePressure", "waterVapor", "seaIce", "snowDepth", "soilWater", "albedo", "jetU", "jetV",
      "outgoingLongwave", "landFraction", "toaNetRadiation", "surfaceNetHeat", "evaporation", "runoff",
      "forestCover", "leafAreaIndex", "snowfall", "snowmelt", "seaIceThickness", "glacierCover",
      "seaLevelPressure", "surfaceShortwave", "surfaceLongwave", "sensibleHeatFlux", "latentHeatFlux",
      "convectivePrecipitation", "largeScalePrecipitation", "toaShortwave", "toaLongwave",
      "surfaceTemperatureMin", "surfaceTemperatureMax", "diurnalTemperatureRange", "dailyTemperatureRangeSum",
      "dailyTemperatureRangeDays", "netPrimaryProduction", "solarZenithCosine", "cape",
      "neutralBuoyancyPressure", "potentialIntensity", "entropyDeficit", "absoluteVorticity",
      "ventilationIndex", "ventilationReducedIntensity", "genesisPotentialIndex", "windShear"};
  static const char* const layer[] = {"layerTemperature", "layerWindU", "layerWindV", "layerHumidity",
                                      "layerCloud"};
  static thread_local std::string name;
  const int n_surface = (int)(sizeof(surface) / sizeof(surface[0]));
  if ((levels != 5 && levels != 10) || index < 0 || index >= n_surface + 5 * levels) return nullptr;
  if (index < n_surface) return surface[index];
  const int k = index - n_surface;
  name = std::string(layer[k % 5]) + std::to_string(k / 5 + 1);
  return name.c_str();
}
