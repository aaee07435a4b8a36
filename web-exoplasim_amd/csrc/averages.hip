// Averages and balance of the published snapshots, on the device.
//
// The reference keeps them on the host (GCM_AVERAGES, src/model/climate/
// gcm/averages/index.ts:26-137; GCM_BALANCE, balance/index.ts:6-68): every
// published snapshot goes into an FP64 month sum; when the calendar month
// changes the sum becomes an FP32 monthly mean in a history of the last
// HISTORY_MONTHS months; when twelve closed months make a calendar year,
// the year's global balances are taken from its annual mean. Here the sum
// and the history live in HBM next to the member's state (a T63 L10
// history is about 1 GB), and each publishing step adds the new snapshot
// with one elementwise launch over every averaging member:
//
//   k_avg_update  open a month (sum = snapshot), add to it, or close it
//                 (mean = sum / samples into its history slot) and open the
//                 next with the snapshot
//   k_avg_year    the yearly balance: annual means of the TOA net
//                 radiation, surface net heat and surface temperature
//                 fields, then their global means (FP64)
//   k_avg_annual  GCM_AVERAGES.annual: every field's annual mean, the
//                 diurnal range as the mean of the months' daily ranges
//   k_avg_climate one calendar month's climatology over the last complete
//                 years (GCM_AVERAGES.view, averages/index.ts:177-222)
//
// A checkpoint's averager (the open month's FP64 sum, the closed months and
// the yearly balances) is read with exo_read_accumulator / exo_read_month /
// exo_yearly_balances and put back with exo_load_averager.
//
// Sums keep the reference's arithmetic: FP64 month sums over FP32
// snapshots, the mean rounded to FP32; annual means accumulate month / 12
// in FP64 and round to FP32 at every month, as a Float32Array does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/exoplasim_mi355x.h"
#include "engine.h"

namespace {

// Snapshot fields the averages treat specially (fields/index.ts:81-132).
constexpr int kSurfaceTemperature = 0;
constexpr int kToaNetRadiation = 16;
constexpr int kSurfaceNetHeat = 17;
constexpr int kDiurnalRange = 37;
constexpr int kDailyRangeSum = 38;
constexpr int kDailyRangeDays = 39;

// One annual-mean element: month values / 12 added in FP64 and rounded to
// FP32 after every month, oldest month first.
__device__ __forceinline__ float annual_value(const float* const* months, size_t at) {
  float v = 0.0f;
  for (int m = 0; m < EXO_MONTHS_PER_YEAR; m++) v = (float)((double)v + (double)months[m][at] / 12.0);
  return v;
}

// The daily temperature range of one month at a point: sum / days where
// the month counted days, else its diurnal range field (fields/index.ts:
// 492-497).
__device__ __forceinline__ float daily_range(const float* month, int nh, int h) {
  const float sum = month[kDailyRangeSum * nh + h], days = month[kDailyRangeDays * nh + h];
  return days > 0.0f ? (float)((double)sum / (double)days) : month[kDiurnalRange * nh + h];
}

}  // namespace

extern "C" __global__ __launch_bounds__(256) void k_avg_update(exo::AvgLaunch a) {
  int lo = 0, hi = a.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.first[mid] <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const size_t e = (size_t)(blockIdx.x - a.first[lo]) * 256 + threadIdx.x;
  if (e >= a.count[lo]) return;
  const float x = a.snap[lo][e];
  double* sum = a.sum[lo];
  if (a.mode[lo] == exo::kAvgAdd) {
    sum[e] += (double)x;
    return;
  }
  if (a.mode[lo] == exo::kAvgClose) a.out[lo][e] = (float)(sum[e] / (double)a.samples[lo]);
  sum[e] = (double)x;
}

// One workgroup per year: the year's twelve months (history slots), its
// member's latitude weights, and 3 doubles out.
extern "C" __global__ __launch_bounds__(256) void k_avg_year(const DevMember* __restrict__ mems, exo::YearLaunch y) {
  const int i = blockIdx.x;
  const DevMember& M = mems[y.desc[i]];
  const int NH = M.nhor;
  const double* w = y.weights[i];
  const float* months[EXO_MONTHS_PER_YEAR];
  for (int m = 0; m < EXO_MONTHS_PER_YEAR; m++) months[m] = y.months[i] + (size_t)y.slot[i][m] * y.floats[i];
  __shared__ double s[4][256];
  double toa = 0, surf = 0, ts = 0, wsum = 0;
  for (int h = threadIdx.x; h < NH; h += 256) {
    const double wh = w[h];
    toa += wh * annual_value(months, (size_t)kToaNetRadiation * NH + h);
    surf += wh * annual_value(months, (size_t)kSurfaceNetHeat * NH + h);
    ts += wh * annual_value(months, (size_t)kSurfaceTemperature * NH + h);
    wsum += wh;
  }
  s[0][threadIdx.x] = toa;
  s[1][threadIdx.x] = surf;
  s[2][threadIdx.x] = ts;
  s[3][threadIdx.x] = wsum;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k)
      for (int q = 0; q < 4; q++) s[q][threadIdx.x] += s[q][threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* out = y.out[i];
    out[0] = s[0][0] / s[3][0];
    out[1] = s[1][0] / s[3][0];
    out[2] = s[2][0] / s[3][0];
  }
}

extern "C" __global__ __launch_bounds__(256) void k_avg_annual(exo::YearLaunch y, int nh, float* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= y.floats[0]) return;
  const float* months[EXO_MONTHS_PER_YEAR];
  for (int m = 0; m < EXO_MONTHS_PER_YEAR; m++) months[m] = y.months[0] + (size_t)y.slot[0][m] * y.floats[0];
  const int field = (int)(e / nh), h = (int)(e % nh);
  if (field == kDailyRangeDays) {
    out[e] = 1.0f;
  } else if (field == kDiurnalRange || field == kDailyRangeSum) {
    float range = 0.0f;
    for (int m = 0; m < EXO_MONTHS_PER_YEAR; m++) range = (float)((double)range + (double)daily_range(months[m], nh, h) / 12.0);
    out[e] = range;
  } else {
    out[e] = annual_value(months, e);
  }
}

// One calendar month's climatology: the month's means of `years` years
// (history slots, oldest year first) summed in FP64, the sum / years
// rounded to FP32.
struct ClimateSlots {
  int years;
  int slot[EXO_CLIMATE_YEARS];
};

extern "C" __global__ __launch_bounds__(256) void k_avg_climate(const float* __restrict__ history, size_t floats,
                                                              ClimateSlots c, float* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= floats) return;
  double sum = 0.0;
  for (int y = 0; y < c.years; y++) sum += (double)history[(size_t)c.slot[y] * floats + e];
  out[e] = (float)(sum / c.years);
}

namespace exo {

size_t Averager::slot_of(size_t index) const {
  return (head + EXO_HISTORY_MONTHS - months.size() + index) % EXO_HISTORY_MONTHS;
}

Averager::~Averager() {
  if (d_sum) (void)hipFree(d_sum);
  if (d_months) (void)hipFree(d_months);
  if (d_years) (void)hipFree(d_years);
  if (d_weights) (void)hipFree(d_weights);
}

// After step k's pack: the averaging members that published at that step,
// with the calendar month of their snapshot.
void enqueue_averages(Engine& e, const std::vector<MemberDev*>& members, const std::vector<int64_t>& steps) {
  AvgLaunch a{};
  int blocks = 0;
  std::vector<std::pair<MemberDev*, int>> years;  // member, history index of its year's last month
  for (size_t i = 0; i < members.size(); i++) {
    MemberDev* m = members[i];
    Averager& v = *m->avg;
    int date[5];
    member_date(m->host, steps[i], date);
    const int n = a.n++;
    a.snap[n] = m->d_arena + m->host.dev.offset[SLOT_snapshot];
    a.sum[n] = v.d_sum;
    a.count[n] = v.floats;
    a.first[n] = blocks;
    blocks += (int)((v.floats + 255) / 256);
    if (!v.open) {
      a.mode[n] = kAvgOpen;
    } else if (v.year == date[0] && v.month == date[1]) {
      a.mode[n] = kAvgAdd;
      v.samples++;
      v.step = steps[i];
      continue;
    } else {
      a.mode[n] = kAvgClose;
      a.samples[n] = v.samples;
      a.out[n] = v.d_months + (size_t)v.head * v.floats;
      v.months.push_back(exo_month_info{v.year, v.month, v.samples, v.step});
      if (v.months.size() > EXO_HISTORY_MONTHS) v.months.pop_front();
      v.head = (v.head + 1) % EXO_HISTORY_MONTHS;
      // a calendar year closes with twelve months of one year ending in
      // month 12 (averages/index.ts:31-39)
      const size_t n_held = v.months.size();
      if (n_held >= EXO_MONTHS_PER_YEAR && v.months.back().month == EXO_MONTHS_PER_YEAR) {
        bool one_year = true;
        for (size_t k = n_held - EXO_MONTHS_PER_YEAR; k < n_held; k++) one_year &= v.months[k].year == v.months.back().year;
        if (one_year && (int)v.years.size() < kAvgYearCapacity) years.push_back({m, (int)n_held - 1});
      }
    }
    v.open = true;
    v.year = date[0];
    v.month = date[1];
    v.samples = 1;
    v.step = steps[i];
  }
  if (a.n > 0) launch(e, k_avg_update, dim3(blocks), dim3(256), a);
  YearLaunch y{};
  for (size_t i = 0; i < years.size(); i++) {
    auto [m, last] = years[i];
    Averager& v = *m->avg;
    const int n = y.n++;
    y.desc[n] = m->desc;
    y.months[n] = v.d_months;
    y.weights[n] = v.d_weights;
    y.floats[n] = v.floats;
    for (int k = 0; k < EXO_MONTHS_PER_YEAR; k++) y.slot[n][k] = (int)v.slot_of(last - (EXO_MONTHS_PER_YEAR - 1) + k);
    y.out[n] = v.d_years + 3 * v.years.size();
    v.years.push_back(v.months[last].year);
    if (y.n == kYearBatch || i + 1 == years.size()) {
      launch(e, k_avg_year, dim3(y.n), dim3(256), e.d_desc, y);
      y = YearLaunch{};
    }
  }
}

}  // namespace exo

using exo::MemberDev;

namespace {

MemberDev* averaging_member(exo_engine* h, int id) {
  if (!h || id < 0 || id >= (int)h->e.members.size() || !h->e.members[id]) return nullptr;
  return h->e.members[id].get();
}

// A cleared averager on the member's device (current device set by the
// caller): month sum, month history, yearly balances and the FP64 Gaussian
// weights the balance's global means use (GCM_GRID.globalMean,
// grid/index.ts:99-111), not the FP32 table. Throws on a failed allocation.
std::unique_ptr<exo::Averager> new_averager(MemberDev* m) {
  auto v = std::make_unique<exo::Averager>();
  v->floats = m->host.layout.floats[SLOT_snapshot];
  HIP_OR_THROW(hipMalloc(&v->d_sum, v->floats * sizeof(double)));
  HIP_OR_THROW(hipMalloc(&v->d_months, (size_t)EXO_HISTORY_MONTHS * v->floats * sizeof(float)));
  HIP_OR_THROW(hipMalloc(&v->d_years, 3 * sizeof(double) * exo::kAvgYearCapacity));
  const size_t nhor = (size_t)m->host.dims.nlat * m->host.dims.nlon;
  HIP_OR_THROW(hipMalloc(&v->d_weights, nhor * sizeof(double)));
  HIP_OR_THROW(hipMemcpy(v->d_weights, &m->host.tables[m->host.tlayout.offset[TAB_zgw]], nhor * sizeof(double),
                         hipMemcpyHostToDevice));
  return v;
}

}  // namespace

extern "C" int exo_set_averaging(exo_engine* h, int id, int enable) {
  MemberDev* m = averaging_member(h, id);
  if (!m) return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id));
  if (h->e.device < 0) return exo::set_error(EXO_E_STATE, "a host-only engine has no device averages");
  if (!enable) {
    if (m->avg) {
      (void)hipStreamSynchronize(h->e.stream);
      m->avg.reset();
    }
    return EXO_OK;
  }
  try {
    HIP_OR_THROW(hipSetDevice(h->e.device));
    HIP_OR_THROW(hipStreamSynchronize(h->e.stream));
    m->avg.reset();
    m->avg = new_averager(m);
  } catch (const std::exception& ex) {
    return exo::set_error(EXO_E_DEVICE, ex.what());
  }
  return EXO_OK;
}

extern "C" int exo_month_count(exo_engine* h, int id, int* count) {
  MemberDev* m = averaging_member(h, id);
  if (!m || !count) return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id));
  *count = m->avg ? (int)m->avg->months.size() : 0;
  return EXO_OK;
}

extern "C" int exo_read_month(exo_engine* h, int id, int index, exo_month_info* info, float* values, size_t count) {
  MemberDev* m = averaging_member(h, id);
  if (!m) return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id));
  if (!m->avg) return exo::set_error(EXO_E_STATE, "member " + std::to_string(id) + " is not averaging");
  const exo::Averager& v = *m->avg;
  if (index < 0 || index >= (int)v.months.size())
    return exo::set_error(EXO_E_ARG, "no closed month " + std::to_string(index));
  if (info) *info = v.months[index];
  if (!values) return EXO_OK;
  if (count != v.floats) return exo::set_error(EXO_E_ARG, "a month needs " + std::to_string(v.floats) + " values");
  if (hipSetDevice(h->e.device) != hipSuccess ||
      hipMemcpyAsync(values, v.d_months + v.slot_of(index) * v.floats, count * sizeof(float), hipMemcpyDeviceToHost,
                     h->e.stream) != hipSuccess ||
      hipStreamSynchronize(h->e.stream) != hipSuccess)
    return exo::set_error(EXO_E_DEVICE, "reading a monthly mean failed");
  return EXO_OK;
}

extern "C" int exo_read_annual(exo_engine* h, int id, float* values, size_t count) {
  MemberDev* m = averaging_member(h, id);
  if (!m || !values) return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id));
  if (!m->avg) return exo::set_error(EXO_E_STATE, "member " + std::to_string(id) + " is not averaging");
  const exo::Averager& v = *m->avg;
  if (v.months.size() < EXO_MONTHS_PER_YEAR) return exo::set_error(EXO_E_STATE, "fewer than 12 closed months");
  if (count != v.floats) return exo::set_error(EXO_E_ARG, "the annual mean needs " + std::to_string(v.floats) + " values");
  exo::Engine& e = h->e;
  try {
    HIP_OR_THROW(hipSetDevice(e.device));
    exo::YearLaunch y{};
    y.n = 1;
    y.months[0] = v.d_months;
    y.floats[0] = v.floats;
    const size_t last = v.months.size() - 1;
    for (int k = 0; k < EXO_MONTHS_PER_YEAR; k++) y.slot[0][k] = (int)v.slot_of(last - (EXO_MONTHS_PER_YEAR - 1) + k);
    float* d_out = nullptr;
    HIP_OR_THROW(hipMallocAsync((void**)&d_out, count * sizeof(float), e.stream));
    hipLaunchKernelGGL(k_avg_annual, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, e.stream, y,
                       m->host.dims.nlat * m->host.dims.nlon, d_out);
    HIP_OR_THROW(hipGetLastError());
    HIP_OR_THROW(hipMemcpyAsync(values, d_out, count * sizeof(float), hipMemcpyDeviceToHost, e.stream));
    HIP_OR_THROW(hipFreeAsync(d_out, e.stream));
    HIP_OR_THROW(hipStreamSynchronize(e.stream));
  } catch (const std::exception& ex) {
    return exo::set_error(EXO_E_DEVICE, ex.what());
  }
  return EXO_OK;
}

extern "C" int exo_yearly_balances(exo_engine* h, int id, exo_yearly_balance* out, int max, int* count) {
  MemberDev* m = averaging_member(h, id);
  if (!m || !count || max < 0 || (max > 0 && !out)) return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id));
  if (!m->avg) {
    *count = 0;
    return EXO_OK;
  }
  const exo::Averager& v = *m->avg;
  *count = (int)v.years.size();
  const int n = std::min(max, *count);
  if (n == 0) return EXO_OK;
  std::vector<double> host(3 * (size_t)n);
  if (hipSetDevice(h->e.device) != hipSuccess ||
      hipMemcpyAsync(host.data(), v.d_years, host.size() * sizeof(double), hipMemcpyDeviceToHost, h->e.stream) !=
          hipSuccess ||
      hipStreamSynchronize(h->e.stream) != hipSuccess)
    return exo::set_error(EXO_E_DEVICE, "reading the yearly balances failed");
  for (int i = 0; i < n; i++)
    out[i] = exo_yearly_balance{v.years[i], host[3 * i], host[3 * i + 1], host[3 * i + 2]};
  return EXO_OK;
}

extern "C" int exo_read_accumulator(exo_engine* h, int id, exo_month_info* info, double* sum, size_t count) {
  MemberDev* m = averaging_member(h, id);
  if (!m || !info) return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id));
  if (!m->avg) return exo::set_error(EXO_E_STATE, "member " + std::to_string(id) + " is not averaging");
  const exo::Averager& v = *m->avg;
  *info = v.open ? exo_month_info{v.year, v.month, v.samples, v.step} : exo_month_info{0, 0, 0, 0};
  if (!sum || !v.open) return EXO_OK;
  if (count != v.floats) return exo::set_error(EXO_E_ARG, "the month sum has " + std::to_string(v.floats) + " values");
  if (hipSetDevice(h->e.device) != hipSuccess ||
      hipMemcpyAsync(sum, v.d_sum, count * sizeof(double), hipMemcpyDeviceToHost, h->e.stream) != hipSuccess ||
      hipStreamSynchronize(h->e.stream) != hipSuccess)
    return exo::set_error(EXO_E_DEVICE, "reading the month sum failed");
  return EXO_OK;
}

extern "C" int exo_load_averager(exo_engine* h, int id, const exo_month_info* months, const float* month_values,
                                 int n_months, const exo_yearly_balance* years, int n_years,
                                 const exo_month_info* current, const double* current_sum, size_t floats) {
  MemberDev* m = averaging_member(h, id);
  if (!m) return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id));
  if (h->e.device < 0) return exo::set_error(EXO_E_STATE, "a host-only engine has no device averages");
  const size_t want = m->host.layout.floats[SLOT_snapshot];
  const bool open = current && current->samples > 0;
  if (floats != want) return exo::set_error(EXO_E_ARG, "a month holds " + std::to_string(want) + " values");
  if (n_months < 0 || n_months > EXO_HISTORY_MONTHS || (n_months > 0 && (!months || !month_values)))
    return exo::set_error(EXO_E_ARG, "at most " + std::to_string(EXO_HISTORY_MONTHS) + " closed months");
  if (n_years < 0 || n_years > exo::kAvgYearCapacity || (n_years > 0 && !years))
    return exo::set_error(EXO_E_ARG, "at most " + std::to_string(exo::kAvgYearCapacity) + " yearly balances");
  if (open && !current_sum) return exo::set_error(EXO_E_ARG, "an open month needs its sum");
  try {
    HIP_OR_THROW(hipSetDevice(h->e.device));
    HIP_OR_THROW(hipStreamSynchronize(h->e.stream));
    m->avg.reset();
    auto v = new_averager(m);
    if (n_months > 0)
      HIP_OR_THROW(hipMemcpy(v->d_months, month_values, (size_t)n_months * floats * sizeof(float), hipMemcpyHostToDevice));
    for (int i = 0; i < n_months; i++) v->months.push_back(months[i]);
    v->head = n_months % EXO_HISTORY_MONTHS;
    std::vector<double> balances(3 * (size_t)n_years);
    for (int i = 0; i < n_years; i++) {
      balances[3 * i] = years[i].toa_net_radiation;
      balances[3 * i + 1] = years[i].surface_net_heat;
      balances[3 * i + 2] = years[i].surface_temperature;
      v->years.push_back(years[i].year);
    }
    if (n_years > 0)
      HIP_OR_THROW(hipMemcpy(v->d_years, balances.data(), balances.size() * sizeof(double), hipMemcpyHostToDevice));
    if (open) {
      HIP_OR_THROW(hipMemcpy(v->d_sum, current_sum, floats * sizeof(double), hipMemcpyHostToDevice));
      v->open = true;
      v->year = current->year;
      v->month = current->month;
      v->samples = current->samples;
      v->step = current->step;
    }
    m->avg = std::move(v);
  } catch (const std::exception& ex) {
    m->avg.reset();
    return exo::set_error(EXO_E_DEVICE, ex.what());
  }
  return EXO_OK;
}

extern "C" int exo_read_climatology(exo_engine* h, int id, int month, exo_month_info* info, float* values,
                                    size_t count, int* years) {
  MemberDev* m = averaging_member(h, id);
  if (!m || !years || month < 1 || month > EXO_MONTHS_PER_YEAR)
    return exo::set_error(EXO_E_ARG, "no member " + std::to_string(id) + " or month " + std::to_string(month));
  if (!m->avg) return exo::set_error(EXO_E_STATE, "member " + std::to_string(id) + " is not averaging");
  const exo::Averager& v = *m->avg;
  // complete years: twelve closed months of one calendar year, all
  // distinct (averages/index.ts:178-190), oldest first, the last ten
  std::map<int, std::vector<size_t>> by_year;
  for (size_t i = 0; i < v.months.size(); i++) by_year[v.months[i].year].push_back(i);
  std::vector<size_t> chosen;  // history index of `month` in each complete year
  for (const auto& [year, held] : by_year) {
    if (held.size() != EXO_MONTHS_PER_YEAR) continue;
    unsigned seen = 0;
    size_t at = 0;
    for (size_t i : held) {
      seen |= 1u << v.months[i].month;
      if (v.months[i].month == month) at = i;
    }
    if (seen == 0x1FFEu) chosen.push_back(at);
  }
  if (chosen.size() > EXO_CLIMATE_YEARS) chosen.erase(chosen.begin(), chosen.end() - EXO_CLIMATE_YEARS);
  *years = (int)chosen.size();
  if (chosen.empty()) return EXO_OK;
  if (info) *info = v.months[chosen.back()];
  if (!values) return EXO_OK;
  if (count != v.floats) return exo::set_error(EXO_E_ARG, "a month needs " + std::to_string(v.floats) + " values");
  exo::Engine& e = h->e;
  try {
    HIP_OR_THROW(hipSetDevice(e.device));
    ClimateSlots c{};
    c.years = (int)chosen.size();
    for (int y = 0; y < c.years; y++) c.slot[y] = (int)v.slot_of(chosen[y]);
    float* d_out = nullptr;
    HIP_OR_THROW(hipMallocAsync((void**)&d_out, count * sizeof(float), e.stream));
    hipLaunchKernelGGL(k_avg_climate, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, e.stream, v.d_months,
                       v.floats, c, d_out);
    HIP_OR_THROW(hipGetLastError());
    HIP_OR_THROW(hipMemcpyAsync(values, d_out, count * sizeof(float), hipMemcpyDeviceToHost, e.stream));
    HIP_OR_THROW(hipFreeAsync(d_out, e.stream));
    HIP_OR_THROW(hipStreamSynchronize(e.stream));
  } catch (const std::exception& ex) {
    return exo::set_error(EXO_E_DEVICE, ex.what());
  }
  return EXO_OK;
}

// ExoPlaSim's Model._isbalanced as GCM_BALANCE.assess (balance/index.ts:
// 6-68): 10-year running means of |yearly balance|, least-squares slopes
// over 5 consecutive running means, the mean of the last 30 slopes against
// 5e-4 W/m^2 per year at the top of the atmosphere and at the surface.
namespace {

constexpr int kRunningMeanYears = 10, kSlopeWindow = 5, kSlopesAveraged = 30;
constexpr double kThreshold = 5e-4;
constexpr int kMinimumYears = kRunningMeanYears + kSlopeWindow - 1;

double mean_of(const double* x, int n) { return std::accumulate(x, x + n, 0.0) / n; }

double slope_of(const double* y, int n) {
  const double x_mean = (n + 1) / 2.0, y_mean = mean_of(y, n);
  double num = 0, den = 0;
  for (int i = 0; i < n; i++) {
    const double x = i + 1 - x_mean;
    num += x * (y[i] - y_mean);
    den += x * x;
  }
  return num / den;
}

double drift_of(const std::vector<double>& series) {
  std::vector<double> running, slopes;
  for (size_t end = kRunningMeanYears; end <= series.size(); end++)
    running.push_back(std::fabs(mean_of(series.data() + end - kRunningMeanYears, kRunningMeanYears)));
  for (size_t end = kSlopeWindow; end <= running.size(); end++)
    slopes.push_back(slope_of(running.data() + end - kSlopeWindow, kSlopeWindow));
  const size_t n = std::min<size_t>(kSlopesAveraged, slopes.size());
  return std::fabs(mean_of(slopes.data() + slopes.size() - n, (int)n));
}

}  // namespace

extern "C" int exo_balance_assess(const exo_yearly_balance* years, int count, exo_balance_assessment* out) {
  if (!out || count < 0 || (count > 0 && !years)) return exo::set_error(EXO_E_ARG, "null balance argument");
  if (count < kMinimumYears) {
    *out = exo_balance_assessment{EXO_BALANCE_SPINNING_UP, kMinimumYears - count, NAN, NAN};
    return EXO_OK;
  }
  std::vector<double> toa(count), surface(count);
  for (int i = 0; i < count; i++) {
    toa[i] = years[i].toa_net_radiation;
    surface[i] = years[i].surface_net_heat;
  }
  const double dt = drift_of(toa), ds = drift_of(surface);
  *out = exo_balance_assessment{dt < kThreshold && ds < kThreshold ? EXO_BALANCE_BALANCED : EXO_BALANCE_DRIFTING, 0,
                                dt, ds};
  return EXO_OK;
}
