// Engine internals shared by the step builders: member bookkeeping, the
// engine, and the launch helper that plans one launch over all active
// members.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cstdint>
#include <deque>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.h"
#include "setup.h"
#include "transforms.h"

extern "C" __global__ void k_leg_inv(const DevMember*, GroupPlan, InvJobs);
extern "C" __global__ void k_math_eval(int, const float*, const float*, float*, int);
extern "C" __global__ void k_dft_inv(const DevMember*, GroupPlan, InvJobs);
extern "C" __global__ void k_dft_fwd(const DevMember*, GroupPlan, FwdJobs);
extern "C" __global__ void k_leg_fwd(const DevMember*, GroupPlan, LegJobs);
extern "C" __global__ void k_leg_inv_split(const DevMember*, GroupPlan, InvJobs);
extern "C" __global__ void k_leg_fwd_split(const DevMember*, GroupPlan, LegJobs);
extern "C" __global__ void k_mkdheat(const DevMember*, LaunchPlan);
extern "C" __global__ void k_mean_dekin(const DevMember*, LaunchPlan);
extern "C" __global__ void k_load_step(const DevMember*, LaunchPlan, const float*);
extern "C" __global__ void k_load_step_at(const DevMember*, LaunchPlan, const float*, const int*, int);
extern "C" __global__ void k_step_index(int*, int);
extern "C" __global__ void k_calcgp(const DevMember*, LaunchPlan);
extern "C" __global__ void k_spectrala(const DevMember*, LaunchPlan);
extern "C" __global__ void k_spectrald_diffusion(const DevMember*, LaunchPlan);
extern "C" __global__ void k_spectrald_update(const DevMember*, LaunchPlan);
extern "C" __global__ void k_physics_in(const DevMember*, LaunchPlan);
extern "C" __global__ void k_fixer_column(const DevMember*, LaunchPlan);
extern "C" __global__ void k_fixer_rows(const DevMember*, LaunchPlan);
extern "C" __global__ void k_fixer_sum(const DevMember*, LaunchPlan);
extern "C" __global__ void k_surface_fluxes(const DevMember*, LaunchPlan);
extern "C" __global__ void k_shortwave(const DevMember*, LaunchPlan);
extern "C" __global__ void k_longwave(const DevMember*, LaunchPlan);
extern "C" __global__ void k_convection(const DevMember*, LaunchPlan);
extern "C" __global__ void k_dca_limit(const DevMember*, LaunchPlan);
extern "C" __global__ void k_rain(const DevMember*, LaunchPlan);
extern "C" __global__ void k_clouds(const DevMember*, LaunchPlan);
extern "C" __global__ void k_soil(const DevMember*, LaunchPlan);
extern "C" __global__ void k_mkradv(const DevMember*, LaunchPlan);
extern "C" __global__ void k_physics_c(const DevMember*, LaunchPlan);
extern "C" __global__ void k_ice_sums(const DevMember*, LaunchPlan);
extern "C" __global__ void k_physics_d(const DevMember*, LaunchPlan);
extern "C" __global__ void k_pack(const DevMember*, LaunchPlan);

namespace exo {

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Throws DeviceError with the call and HIP's message when a HIP call fails.
#define HIP_OR_THROW(expr)                                                              \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      throw exo::DeviceError(std::string(#expr) + ": " + hipGetErrorString(_e));        \
  } while (0)

// Every descriptor of the engine is taken.
struct FullError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Device-side averages of a member's published snapshots (averages.hip):
// the open month's FP64 sum, the history of closed monthly means (a ring
// of EXO_HISTORY_MONTHS slots) and the yearly balances (3 doubles each).
constexpr int kAvgYearCapacity = 4096;
struct Averager {
  double* d_sum = nullptr;
  float* d_months = nullptr;
  double* d_years = nullptr;
  double* d_weights = nullptr;  // FP64 Gaussian weight of every grid point
  size_t floats = 0;  // snapshot floats: fields x NHOR
  bool open = false;  // a month is being summed
  int year = 0, month = 0, samples = 0;
  int64_t step = 0;
  int head = 0;                       // history slot of the next closed month
  std::deque<exo_month_info> months;  // closed months held, oldest first
  std::vector<int> years;             // years whose balance is in d_years
  size_t slot_of(size_t index) const; // history slot of months[index]
  ~Averager();
};

// Kernel arguments of the averaging launches.
enum { kAvgOpen = 0, kAvgAdd = 1, kAvgClose = 2 };
struct AvgLaunch {
  int n;
  int first[EXO_MAX_MEMBERS + 1];  // first workgroup of entry i
  int mode[EXO_MAX_MEMBERS];
  int samples[EXO_MAX_MEMBERS];  // kAvgClose: snapshots in the closing month
  size_t count[EXO_MAX_MEMBERS];
  const float* snap[EXO_MAX_MEMBERS];
  double* sum[EXO_MAX_MEMBERS];
  float* out[EXO_MAX_MEMBERS];  // kAvgClose: the month's history slot
};
constexpr int kYearBatch = 16;  // years per k_avg_year launch (kernel argument size)
struct YearLaunch {
  int n;
  int desc[kYearBatch];
  const float* months[kYearBatch];  // history base
  const double* weights[kYearBatch];
  size_t floats[kYearBatch];
  int slot[kYearBatch][EXO_MONTHS_PER_YEAR];
  double* out[kYearBatch];
};

struct MemberDev {
  Member host;
  int precision = 0;
  float* d_arena = nullptr;
  float* d_tab_alloc = nullptr;  // the member's own tables (zgw on)
  float* d_tab = nullptr;        // base such that d_tab + tabo[t] addresses table t
  int shared_key = -1;           // truncation of its shared transform tables
  float* d_consts = nullptr;  // phases x CON_COUNT
  float* d_bm1 = nullptr;     // phases x bm1 count
  int loaded_phase = 0;       // phase whose constants/bm1 are on the device
  int phase = 0;              // phase of the next step
  int64_t step = 0;           // nstep of the next main-loop step
  int64_t carbon_count = 0;
  int desc = -1;              // descriptor index
  bool active = false;
  std::unique_ptr<Averager> avg;  // set while averaging (exo_set_averaging)
};

// Times every launch of one kernel with a pair of HIP events recorded on
// the engine's stream around it (begin[i], end[i] for launch i).
struct KernelProbe {
  const void* kernel = nullptr;
  std::vector<hipEvent_t> begin, end;
  size_t used = 0;
  size_t missed = 0;  // launches beyond the pool
};

// Fourier and Legendre tables of one truncation, shared by its members.
struct SharedTables {
  float* d = nullptr;
  int refs = 0;
  float lscale[8] = {};  // DevMember::lscale
};

struct Engine {
  int device = 0;
  std::map<int, SharedTables> transform_tables;
  hipStream_t stream = nullptr;
  int max_members = 0;
  DevMember* d_desc = nullptr;
  std::vector<DevMember> h_desc;
  std::vector<std::unique_ptr<MemberDev>> members;
  float* h_params = nullptr;
  float* d_params = nullptr;
  size_t params_cap = 0;
  hipEvent_t params_done = nullptr, t0 = nullptr, t1 = nullptr;
  bool timed = false;
  int publish_every = 0;
  KernelProbe probe;
  // Step graph: one captured step (calendar values, dynamics, physics,
  // snapshot pack) replayed while its key holds. The key is the active
  // descriptors, their loaded phases, the calendar buffer and `epoch`,
  // which admit_member and drop_member bump. The captured k_load_step_at
  // reads its calendar row at *d_step_index, which the graph's last node
  // advances.
  bool graphs = true;
  uint64_t epoch = 0;
  // Lanes of a captured step: with n > 1, the members split into n sets
  // of about equal work whose steps run on n streams, forked and joined
  // inside the graph, so one set's latency-bound launches overlap the
  // others'. Each member's launches keep their order and arithmetic, so
  // the results are the same for every n. exo_autotune picks by timing.
  static constexpr int kMaxLanes = 4;
  int lanes = 2;
  hipStream_t lane_stream[kMaxLanes - 1] = {};
  hipEvent_t lane_fork = nullptr;
  hipEvent_t lane_join[kMaxLanes - 1] = {};
  // Members per group of the grouped transforms, per (truncation, levels);
  // kLegGroupDefault where unset. exo_autotune picks them by timing; any value in
  // 1..kLegGroup gives bit-identical members, since every member's rows
  // are reduced on their own.
  std::map<std::pair<int, int>, int> group_cap;
  int group_cap_of(const DevMember& d) const {
    const auto it = group_cap.find({d.ntru, d.nlev});
    return it == group_cap.end() ? kLegGroupDefault : it->second;
  }
  std::vector<int64_t> graph_key;
  hipGraphExec_t graph = nullptr;
  int* d_step_index = nullptr;
  // exo_step_noalloc's calendar rows: the device address of the file-scope
  // scratch array on this engine's device (hipGetSymbolAddress at create),
  // a pinned host staging block of the same size allocated at create, and
  // the event after the last upload from it.
  float* d_scratch = nullptr;
  float* h_scratch = nullptr;
  hipEvent_t scratch_done = nullptr;
};

// Every launch of the step goes through here, so the probe can bracket
// the kernel it times.
template <typename... Params, typename... Args>
inline void launch(Engine& e, void (*k)(Params...), dim3 grid, dim3 block, Args... args) {
  KernelProbe& p = e.probe;
  const bool timed = p.kernel == reinterpret_cast<const void*>(k);
  const bool room = timed && p.used < p.begin.size();
  if (timed && !room) p.missed++;
  if (room) hipEventRecord(p.begin[p.used], e.stream);
  hipLaunchKernelGGL(k, grid, block, 0, e.stream, args...);
  if (room) hipEventRecord(p.end[p.used++], e.stream);
}

inline int tiles_column(const DevMember& d) { return (d.nhor + 63) / 64; }
inline int tiles_levels(const DevMember& d) { return (d.nlev * d.nhor + 63) / 64; }

template <typename Jobs>
int job_planes_total(const Jobs& jobs, int nlev) {
  int n = 0;
  for (int j = 0; j < jobs.n; j++) n += jobs.j[j].per_level ? nlev : 1;
  return n;
}

class Stepper {
 public:
  Stepper(Engine& e, const std::vector<int>& active) : e_(e), active_(active) {}

  template <typename F>
  LaunchPlan plan(F tiles) const {
    LaunchPlan p{};
    p.n = (int)active_.size();
    int at = 0;
    for (int i = 0; i < p.n; i++) {
      p.member[i] = active_[i];
      p.first[i] = at;
      at += tiles(e_.h_desc[active_[i]]);
    }
    p.first[p.n] = at;
    return p;
  }

  // The members include(d) selects, in groups of at most kLegGroup sharing
  // transform tables and level count (active order within a group);
  // blocks(d, count) workgroups per group. The groups are dealt heaviest
  // first (highest truncation, then most levels), so the long workgroups
  // start first instead of trailing the launch.
  template <typename I, typename F>
  GroupPlan group_plan(I include, F blocks) const {
    std::vector<std::vector<int>> groups;
    std::vector<char> taken(active_.size(), 0);
    for (size_t a = 0; a < active_.size(); a++) {
      const DevMember& d = e_.h_desc[active_[a]];
      if (taken[a] || !include(d)) continue;
      std::vector<int> g;
      const int cap = e_.group_cap_of(d);
      for (size_t b = a; b < active_.size() && (int)g.size() < cap; b++) {
        const DevMember& o = e_.h_desc[active_[b]];
        if (taken[b] || !include(o) || o.ltab != d.ltab || o.nlev != d.nlev) continue;
        taken[b] = 1;
        g.push_back(active_[b]);
      }
      groups.push_back(std::move(g));
    }
    std::stable_sort(groups.begin(), groups.end(), [&](const std::vector<int>& x, const std::vector<int>& y) {
      const DevMember &a = e_.h_desc[x[0]], &b = e_.h_desc[y[0]];
      return a.ntru != b.ntru ? a.ntru > b.ntru : a.nlev > b.nlev;
    });
    GroupPlan p{};
    int at = 0, listed = 0;
    for (const std::vector<int>& g : groups) {
      p.start[p.n] = listed;
      p.count[p.n] = (int)g.size();
      for (int id : g) {
        const DevMember& m = e_.h_desc[id];
        const DevMember& m0 = e_.h_desc[g[0]];
        if (std::memcmp(m.slot, m0.slot, sizeof(m.slot)) != 0)
          throw std::runtime_error("members of a transform group must share one arena layout");
        p.arena[listed] = m.arena;
        p.list[listed++] = id;
      }
      p.first[p.n] = at;
      at += blocks(e_.h_desc[g[0]], (int)g.size());
      p.n++;
    }
    p.first[p.n] = at;
    return p;
  }

  // The zonal transforms run in FP32 for both forms: every member, rows(d)
  // rows per member, 32 rows (two lanes a row) per one-wave workgroup.
  template <typename F>
  GroupPlan dft_plan(F rows) const {
    return group_plan([](const DevMember&) { return true; },
                      [&](const DevMember& d, int count) { return (count * rows(d) + 31) / 32; });
  }

  void column(void (*k)(const DevMember*, LaunchPlan)) {
    const LaunchPlan p = plan(tiles_column);
    launch(e_, k, dim3(p.first[p.n]), dim3(64), e_.d_desc, p);
  }
  // one thread per spectral coefficient and level, 256 a workgroup
  void coeff_levels(void (*k)(const DevMember*, LaunchPlan)) {
    const LaunchPlan p = plan([](const DevMember& d) { return (d.nlev * d.nrsp + 255) / 256; });
    launch(e_, k, dim3(p.first[p.n]), dim3(256), e_.d_desc, p);
  }
  // k_spectrala: 10 waves a workgroup over 640 / nlev coefficients
  void spectrala() {
    const LaunchPlan p = plan([](const DevMember& d) {
      const int per = 64 * (10 / d.nlev);
      return (d.nrsp + per - 1) / per;
    });
    launch(e_, k_spectrala, dim3(p.first[p.n]), dim3(640), e_.d_desc, p);
  }
  void levels(void (*k)(const DevMember*, LaunchPlan)) {
    const LaunchPlan p = plan(tiles_levels);
    launch(e_, k, dim3(p.first[p.n]), dim3(64), e_.d_desc, p);
  }
  // one workgroup of `threads` per member and level
  void member_levels(void (*k)(const DevMember*, LaunchPlan), int threads) {
    const LaunchPlan p = plan([](const DevMember& d) { return d.nlev; });
    launch(e_, k, dim3(p.first[p.n]), dim3(threads), e_.d_desc, p);
  }
  // one thread per point and plane (nlev + extra planes), 256 a workgroup
  void point_levels(void (*k)(const DevMember*, LaunchPlan), int extra = 0) {
    const LaunchPlan p = plan([extra](const DevMember& d) { return ((d.nlev + extra) * d.nhor + 255) / 256; });
    launch(e_, k, dim3(p.first[p.n]), dim3(256), e_.d_desc, p);
  }
  void mkdheat() {
    const LaunchPlan p = plan(tiles_levels);
    launch(e_, k_mkdheat, dim3(p.first[p.n]), dim3(64), e_.d_desc, p);
  }
  void inv(const InvJobs& jobs) {
    // the Legendre half per operand form (plain FP32, split FP16), each over
    // member groups: one workgroup per wavenumber pair (m, T - m) and job
    for (int form = 0; form < 2; form++) {
      const GroupPlan pg = group_plan([form](const DevMember& d) { return d.precision == form; },
                                      [&](const DevMember& d, int) { return (d.ntru + 2) / 2 * jobs.n; });
      if (pg.first[pg.n] > 0)
        launch(e_, form ? k_leg_inv_split : k_leg_inv, dim3(pg.first[pg.n]), dim3(512), e_.d_desc, pg, jobs);
    }
    // k_dft_inv: 32 rows (member x job x output x level x latitude) of a
    // group per workgroup, over every member of both forms
    const GroupPlan pd = dft_plan([&](const DevMember& d) {
      int rows = 0;
      for (int j = 0; j < jobs.n; j++)
        rows += (jobs.j[j].kind == INV_PLAIN ? 1 : 2) * job_planes(jobs.j[j].per_level, d.nlev) * d.nlat;
      return rows;
    });
    launch(e_, k_dft_inv, dim3(pd.first[pd.n]), dim3(64), e_.d_desc, pd, jobs);
  }
  void fwd(const FwdJobs& jobs) {
    // 32 rows (member x job x level x latitude) of a group per workgroup
    const GroupPlan p = dft_plan([&](const DevMember& d) {
      int rows = 0;
      for (int j = 0; j < jobs.n; j++) rows += job_planes(jobs.j[j].per_level, d.nlev) * d.nlat;
      return rows;
    });
    launch(e_, k_dft_fwd, dim3(p.first[p.n]), dim3(64), e_.d_desc, p, jobs);
  }
  void leg(const LegJobs& jobs) {
    // per operand form, over member groups: one workgroup per wavenumber
    // and job
    for (int form = 0; form < 2; form++) {
      const GroupPlan pg = group_plan([form](const DevMember& d) { return d.precision == form; },
                                      [&](const DevMember& d, int) { return (d.ntru + 1) * jobs.n; });
      if (pg.first[pg.n] > 0)
        launch(e_, form ? k_leg_fwd_split : k_leg_fwd, dim3(pg.first[pg.n]), dim3(512), e_.d_desc, pg, jobs);
    }
  }
  void load_step(const float* d_params) {
    const LaunchPlan p = plan([](const DevMember&) { return 1; });
    launch(e_, k_load_step, dim3(p.n), dim3(64), e_.d_desc, p, d_params);
  }
  // The calendar row at base + *index * per_step (the step graph's form).
  void load_step_at(const float* base, const int* index, int per_step) {
    const LaunchPlan p = plan([](const DevMember&) { return 1; });
    launch(e_, k_load_step_at, dim3(p.n), dim3(64), e_.d_desc, p, base, index, per_step);
  }

 private:
  Engine& e_;
  const std::vector<int>& active_;
};


inline InvJob inv_plain(int table, int src, int four, int dst, int per_level) {
  return InvJob{INV_PLAIN, table, src, -1, four, -1, dst, -1, per_level, 0};
}
inline InvJob inv_winds(int d, int z, int fu, int fv, int u, int v, int plavor) {
  return InvJob{INV_DV2UV, -1, d, z, fu, fv, u, v, 1, plavor};
}
inline FwdJob fwd_job(int src, int dst, int per_level) { return FwdJob{src, dst, per_level}; }
inline LegJob leg_plain(int src, int dst, int per_level) { return LegJob{LEG_PLAIN, -1, -1, -1, src, dst, -1, per_level}; }


template <typename Jobs, typename Job>
Jobs jobs_of(std::initializer_list<Job> list) {
  Jobs j{};
  for (const Job& x : list) j.j[j.n++] = x;
  return j;
}

void enqueue_physics(Engine& e, const std::vector<int>& active);
void enqueue_step(Engine& e, const std::vector<int>& active);

// Member life cycle. prepare_member runs the FP64 host set-up (no device
// work; safe to call from several threads); admit_member uploads it into a
// free descriptor and returns its id (FullError when none is free);
// member_bytes is the device memory admitting it takes, its truncation's
// shared tables included when no member holds them yet; drop_member frees
// it after the queued work.
std::unique_ptr<MemberDev> prepare_member(const exo_config& cfg, int precision,
                                          const exo_surface_field* fields = nullptr, int count = 0);
// Sets the calling thread's exo_last_error() text; returns `code`.
int set_error(int code, const std::string& message);
size_t member_bytes(const Engine& e, const MemberDev& md);
int admit_member(Engine& e, std::unique_ptr<MemberDev> md);
void drop_member(Engine& e, int id);
// The averaging members that published at the step just enqueued (their
// snapshot's step numbers): their month sums, closed months and yearly
// balances (averages.hip).
void enqueue_averages(Engine& e, const std::vector<MemberDev*>& members, const std::vector<int64_t>& steps);

}  // namespace exo

// The C-ABI's engine handle.
struct exo_engine {
  exo::Engine e;
};
