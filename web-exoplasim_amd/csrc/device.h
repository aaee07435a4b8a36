// Device-side view of the batch: one descriptor per member (planet) and a
// launch plan that maps a flat workgroup index to (member, tile). Every
// kernel of the step takes both, so one launch advances every member of a
// mixed T21/T42/T63, L5/L10 batch; a single run is a batch of one.
#pragma once
#include <hip/hip_runtime.h>

#include "model.h"

#define EXO_MAX_MEMBERS 64

struct DevMember {
  int ntru, nlat, nlon, nlev;
  int ncsp, nrsp, nfc, nhor;
  int precision;  // 0 plain FP32, 1 split-operand Legendre
  int pad_[3];
  float* arena;
  // The member's tables; tab + tabo[t] for every table t from zgw on.
  const float* tab;
  // The transform tables (Fourier and Legendre, tabo below zgw), one copy
  // shared by every member of the same truncation: ltab + tabo[t].
  const float* ltab;
  // Power-of-two scale of each Legendre table (qc, qe, qm, qq, qi, qj, qu,
  // qv) that puts its largest magnitude in [2^12, 2^13): the split form's
  // FP16 operands.
  float lscale[8];
  int slot[SLOT_COUNT];
  int tabo[TAB_COUNT];
  float k[CON_COUNT];
};

// first[i] is the first workgroup of member i; first[n] the total.
struct LaunchPlan {
  int n;
  int member[EXO_MAX_MEMBERS];  // descriptor index of each planned member
  int first[EXO_MAX_MEMBERS + 1];
};

__device__ __forceinline__ int plan_find(const LaunchPlan& p, int block, int* tile) {
  int lo = 0, hi = p.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.first[mid] <= block) lo = mid;
    else hi = mid - 1;
  }
  *tile = block - p.first[lo];
  return p.member[lo];
}

// Members one Legendre workgroup serves at most (FP32 form), and the
// group size where none is set or tuned.
constexpr int kLegGroup = 8;
constexpr int kLegGroupDefault = 4;

// Members grouped for the transform kernels: group i is list[start[i] ..
// start[i] + count[i]), members of one truncation and level count (one set
// of shared tables); first[i] is its first workgroup, first[n] the total.
// The members of a group share one arena layout (truncation and levels),
// so a slot's offset is read once from the group's first descriptor;
// arena[j] is the arena of list[j], so the transform kernels find a
// member's rows without reading its descriptor.
struct GroupPlan {
  int n;
  int first[EXO_MAX_MEMBERS + 1];
  int start[EXO_MAX_MEMBERS];
  int count[EXO_MAX_MEMBERS];
  int list[EXO_MAX_MEMBERS];
  float* arena[EXO_MAX_MEMBERS];
};

__device__ __forceinline__ int group_find(const GroupPlan& p, int block, int* tile) {
  int lo = 0, hi = p.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.first[mid] <= block) lo = mid;
    else hi = mid - 1;
  }
  *tile = block - p.first[lo];
  return lo;
}

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD
// b mod 8). Within each run of 64 blocks, XCD x is given the 8 consecutive
// tiles 8x..8x+7 instead, so neighbouring tiles, which share cache lines
// (of the Fourier rows of neighbouring wavenumbers, or of neighbouring
// rows' coefficient runs), meet in one L2 while heavy and light tiles stay
// spread over all XCDs. A last partial run keeps the plain order.
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int run = b >> 6;
  if ((run + 1) * 64 > n) return b;
  return (run << 6) + (b & 7) * 8 + ((b >> 3) & 7);
}

// A descriptor's device pointers, read as pointers into the global address
// space. Read as plain float*, they are generic, and every access through
// them is a flat_* instruction, which the hardware counts against the LDS
// and scalar-load counter as well as the vector-memory one, so each wait
// for a scalar or LDS result also waits for every flat load in flight.
// Read as global pointers, the accesses become global_* instructions.
typedef __attribute__((address_space(1))) float exo_gfloat;
typedef __attribute__((address_space(1))) const float exo_gcfloat;
__device__ __forceinline__ float* dev_arena(const DevMember& M) {
  return (float*)*static_cast<exo_gfloat* const*>(static_cast<const void*>(&M.arena));
}
__device__ __forceinline__ const float* dev_tab(const DevMember& M) {
  return (const float*)*static_cast<exo_gcfloat* const*>(static_cast<const void*>(&M.tab));
}
__device__ __forceinline__ const float* dev_ltab(const DevMember& M) {
  return (const float*)*static_cast<exo_gcfloat* const*>(static_cast<const void*>(&M.ltab));
}

// A workgroup barrier that orders LDS accesses only. __syncthreads() is a
// workgroup-scope fence on every address space, so each wave first waits
// for all of its global loads and stores in flight; here they stay in
// flight across the barrier. For workgroups that share data through LDS
// alone.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

#define SLOTP(M, name) (dev_arena(M) + (M).slot[SLOT_##name])
#define TABP(M, name) (dev_tab(M) + (M).tabo[TAB_##name])
#define KON(M, name) ((M).k[CON_##name])
