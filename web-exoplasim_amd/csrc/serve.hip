// Serving loop: continuous batching of independent runs on one engine.
//
// Runs arrive over time (exo_request.arrival_s). Host threads run each
// arrived run's FP64 set-up; the loop admits prepared runs while the
// engine has a free descriptor and the resident runs' device memory stays
// within the HBM budget (first fit, in arrival order), advances every
// resident run together, and retires each run at exactly its requested
// step count: a batched advance is as long as the shortest remaining run
// (capped), so runs that arrive meanwhile join at the next advance. A run
// computes the same alone and in the mix (every kernel treats members
// independently), and its state never leaves the device until it ends.
//
// The reference serves one run per page (runner/index.ts:61-78); this loop
// is the batched counterpart the engine's design allows, with no
// counterpart there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/exoplasim_mi355x.h"
#include "engine.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Prepared {
  int index;
  std::unique_ptr<exo::MemberDev> member;
  std::string error;
};

struct Resident {
  int index, id;
  int64_t remaining;
  size_t bytes;
};

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace

extern "C" int exo_serve(exo_engine* h, const exo_request* requests, int count, const exo_serve_options* options,
                         exo_result* results, double* wall_s) {
  if (!h || count < 0 || (count > 0 && (!requests || !results)))
    return exo::set_error(EXO_E_ARG, "null argument or negative request count");
  exo::Engine& e = h->e;
  if (e.device < 0) return exo::set_error(EXO_E_STATE, "a host-only engine cannot serve runs");
  for (auto& m : e.members)
    if (m) return exo::set_error(EXO_E_STATE, "exo_serve needs an engine without members");
  exo_serve_options opt{};
  if (options) opt = *options;
  if (opt.max_chunk_steps <= 0) opt.max_chunk_steps = 32;
  if (opt.prepare_threads <= 0) opt.prepare_threads = 4;
  size_t budget = opt.hbm_budget_bytes;
  if (budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(e.device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
      return exo::set_error(EXO_E_DEVICE, "hipMemGetInfo failed");
    budget = free_b / 10 * 9;
  }
  std::vector<int> order(count);
  for (int i = 0; i < count; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return requests[a].arrival_s < requests[b].arrival_s; });
  for (int i = 0; i < count; i++) {
    results[i] = exo_result{};
    results[i].arrival_s = requests[i].arrival_s;
    results[i].status = EXO_E_STATE;
  }

  const Clock::time_point t0 = Clock::now();
  std::mutex mu;
  std::condition_variable ready_cv, stop_cv;
  std::deque<Prepared> ready;
  std::atomic<int> next{0};
  std::atomic<bool> stop{false};
  std::vector<std::thread> workers;
  for (int w = 0; w < opt.prepare_threads; w++) {
    workers.emplace_back([&]() {
      for (;;) {
        const int k = next.fetch_add(1);
        if (k >= count || stop.load()) return;
        const int idx = order[k];
        {
          // the run has not arrived before its arrival time
          const Clock::time_point when =
              t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(requests[idx].arrival_s));
          std::unique_lock<std::mutex> lock(mu);
          if (stop_cv.wait_until(lock, when, [&]() { return stop.load(); })) return;
        }
        Prepared p{idx, nullptr, ""};
        try {
          if (requests[idx].steps <= 0) throw std::runtime_error("a run needs at least one step");
          p.member = exo::prepare_member(requests[idx].config, requests[idx].precision);
        } catch (const std::exception& ex) {
          p.error = ex.what();
        }
        std::lock_guard<std::mutex> lock(mu);
        ready.push_back(std::move(p));
        ready_cv.notify_one();
      }
    });
  }

  int status = EXO_OK;
  int finished = 0;
  size_t used = 0;
  std::vector<Resident> resident;
  std::vector<Prepared> pending;  // prepared, waiting for room, arrival order
  while (finished < count) {
    {
      std::unique_lock<std::mutex> lock(mu);
      if (resident.empty() && pending.empty() && ready.empty())
        ready_cv.wait_for(lock, std::chrono::milliseconds(50));
      while (!ready.empty()) {
        pending.push_back(std::move(ready.front()));
        ready.pop_front();
      }
    }
    std::stable_sort(pending.begin(), pending.end(), [&](const Prepared& a, const Prepared& b) {
      return requests[a.index].arrival_s < requests[b.index].arrival_s;
    });
    // admit, first fit in arrival order
    for (auto it = pending.begin(); it != pending.end();) {
      exo_result& r = results[it->index];
      if (!it->member) {
        r.status = EXO_E_ARG;
        r.finished_s = seconds_since(t0);
        finished++;
        it = pending.erase(it);
        continue;
      }
      const size_t bytes = exo::member_bytes(e, *it->member);
      if ((int)resident.size() >= e.max_members || used + bytes > budget) {
        ++it;
        continue;
      }
      try {
        const int id = exo::admit_member(e, std::move(it->member));
        resident.push_back(Resident{it->index, id, requests[it->index].steps, bytes});
        used += bytes;
        r.admitted_s = seconds_since(t0);
      } catch (const std::exception&) {
        r.status = EXO_E_DEVICE;
        r.finished_s = seconds_since(t0);
        finished++;
      }
      it = pending.erase(it);
    }
    if (resident.empty()) {
      // nothing resident and the oldest prepared run still does not fit:
      // it never will
      if (!pending.empty()) {
        exo_result& r = results[pending.front().index];
        r.status = EXO_E_ARG;
        r.finished_s = seconds_since(t0);
        finished++;
        pending.erase(pending.begin());
      }
      continue;
    }
    int64_t chunk = opt.max_chunk_steps;
    for (const Resident& x : resident) chunk = std::min(chunk, x.remaining);
    int rc = exo_step(h, (int)chunk);
    if (rc == EXO_OK) rc = exo_sync(h);
    if (rc != EXO_OK) {
      status = rc;
      break;
    }
    const int finished_before = finished;
    for (auto it = resident.begin(); it != resident.end();) {
      it->remaining -= chunk;
      if (it->remaining > 0) {
        ++it;
        continue;
      }
      exo_result& r = results[it->index];
      r.steps = requests[it->index].steps;
      r.status = exo_global_means(h, it->id, &r.tsurf, &r.toa_net);
      r.finished_s = seconds_since(t0);
      exo::drop_member(e, it->id);
      used -= it->bytes;
      finished++;
      it = resident.erase(it);
    }
    if (opt.progress && finished > finished_before)
      opt.progress(opt.progress_user, finished, count, seconds_since(t0));
  }
  {
    std::lock_guard<std::mutex> lock(mu);
    stop.store(true);
  }
  stop_cv.notify_all();
  for (auto& t : workers) t.join();
  for (const Resident& x : resident) exo::drop_member(e, x.id);
  if (wall_s) *wall_s = seconds_since(t0);
  return status;
}
