// Zonal DFTs of the ExoPlaSim step on gfx950 matrix cores.
//
// With the spectral truncation, a latitude row's zonal transform is a
// product with the truncated DFT matrix of its truncation (the FP32 tables
// FOURIER.tables builds, fourier/index.ts:19-36):
//   fc2gp (fourier/wgsl/inverse.wgsl:4-18):  g[x] = sum_k F[k] inv[k][x]
//   gp2fc (fourier/wgsl/forward.wgsl:4-18):  F[k] = sum_x g[x] fwd[x][k]
// and d/dlambda (dlambda.wgsl:5-17) is fc2gp of (-m Im F, m Re F). Rows of
// one job are contiguous in the arena (level-major planes of latitude
// rows), so a job is one [rows x K] x [K x N] product with the table as
// the right operand, shared by every row, job and member of a truncation.
//
// A workgroup takes 64 rows of one job (16 per wave) and all N columns;
// the reduction runs in chunks of 32: the rows' chunk and the table's
// chunk are staged in LDS, then each wave runs v_mfma_f32_16x16x4_f32
// over its 16 rows and every 16-column tile.
#include "transforms.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kRows = 64;   // rows per workgroup
constexpr int kChunk = 32;  // reduction chunk
constexpr int kMaxN = 192;  // NLON at T63 (inverse); NFC <= 128 (forward)
constexpr int kMaxTilesN = kMaxN / 16;

__host__ __device__ constexpr int a_stride(int r) { return (r + 29) / 32 * 32 + 2; }
__host__ __device__ constexpr int b_stride(int c) { return (c + 15) / 32 * 32 + 16; }
constexpr int kAs = kRows * a_stride(kChunk);
constexpr int kBs = kChunk * b_stride(kMaxN);

// Rows of an inverse job: outputs x planes x latitudes.
__device__ __forceinline__ int inv_rows(const DevMember& M, const InvJob& job) {
  return (job.kind == INV_PLAIN ? 1 : 2) * (job.per_level ? M.nlev : 1) * M.nlat;
}
__device__ __forceinline__ int fwd_rows(const DevMember& M, const FwdJob& job) {
  return (job.per_level ? M.nlev : 1) * M.nlat;
}

// The job and the row block of workgroup `tile` of a member.
template <typename Jobs, typename Rows>
__device__ __forceinline__ int find_block(const Jobs& jobs, int tile, Rows rows, int* block) {
  int j = 0;
  for (; j + 1 < jobs.n; j++) {
    const int blocks = (rows(jobs.j[j]) + kRows - 1) / kRows;
    if (tile < blocks) break;
    tile -= blocks;
  }
  *block = tile;
  return j;
}

// acc[nt] += A[rows 16 wave .. +16][0..R) x B[0..R)[16 nt .. +16], nt < NT
__device__ __forceinline__ void mfma_rows(const float* As, const float* Bs, int bs, int R, int NT, f4* acc) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  const float* a = As + (16 * wave + i) * a_stride(kChunk) + kk;
  const float* b = Bs + kk * bs + i;
  for (int r = 0; r < R; r += 4) {
    const float av = a[r];
#pragma unroll
    for (int nt = 0; nt < kMaxTilesN; nt++)
      if (nt < NT) acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b[r * bs + 16 * nt], acc[nt], 0, 0, 0);
  }
}

// Rows x columns of a chunk staged with lanes along the columns (32-wide
// chunks take two rows per wave at a time), each lane with 8 loads in
// flight before it stores.
template <typename Load, typename Store>
__device__ __forceinline__ void stage(int rows, int cols, Load load, Store store) {
  const bool narrow = cols <= 32;
  const int lane = threadIdx.x & 63;
  const int r0 = narrow ? (threadIdx.x >> 5) : (threadIdx.x >> 6);
  const int rstep = narrow ? 8 : 4;
  for (int c = narrow ? (lane & 31) : lane; c < cols; c += 64) {
    for (int base = r0; base < rows; base += 8 * rstep) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + rstep * j;
        v[j] = r < rows ? load(r, c) : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + rstep * j;
        if (r < rows) store(r, c, v[j]);
      }
    }
  }
}

}  // namespace

extern "C" __global__ __launch_bounds__(kThreads) void k_dft_inv(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                                  InvJobs jobs) {
  __shared__ float As[kAs];
  __shared__ float Bs[kBs];
  // per row of the block: arena offset of its Fourier row (-1 past the
  // job's rows), whether it is d/dlambda's, and its grid row's offset
  __shared__ int src_at[kRows], dst_at[kRows], derived[kRows];
  int tile, block;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int j = find_block(jobs, tile, [&](const InvJob& x) { return inv_rows(M, x); }, &block);
  const InvJob& job = jobs.j[j];
  const int nlat = M.nlat, nlon = M.nlon, nfc = M.nfc;
  const int plane_rows = (job.per_level ? M.nlev : 1) * nlat;  // rows of one output
  const int rows = inv_rows(M, job);
  const int row0 = block * kRows;
  const int K = nfc;  // a multiple of 4 at T21/T42/T63 (44, 86 -> 88 padded, 128)
  const int KP = (K + 3) & ~3;
  const int NT = nlon / 16;
  const int bs = b_stride(nlon);
  const float* inv = M.ltab + M.tabo[TAB_fourier_inverse];
  f4 acc[kMaxTilesN];
#pragma unroll
  for (int nt = 0; nt < kMaxTilesN; nt++) acc[nt] = f4{0, 0, 0, 0};
  if (threadIdx.x < kRows) {
    const int row = row0 + threadIdx.x;
    const int o = row / plane_rows, pr = row % plane_rows;
    const int src = (job.kind == INV_DV2UV && o == 1) ? job.four1 : job.four0;
    src_at[threadIdx.x] = row < rows ? M.slot[src] + pr * nfc : -1;
    dst_at[threadIdx.x] = row < rows ? M.slot[o == 0 ? job.dst0 : job.dst1] + pr * nlon : -1;
    derived[threadIdx.x] = job.kind == INV_DLAMBDA && o == 1;
  }
  __syncthreads();
  for (int k0 = 0; k0 < KP; k0 += kChunk) {
    const int kc = KP - k0 < kChunk ? KP - k0 : kChunk;
    // A: the block's Fourier rows (d/dlambda's built from its row)
    stage(
        kRows, kc,
        [&](int r, int c) {
          const int k = k0 + c, at = src_at[r];
          if (at < 0 || k >= K) return 0.0f;
          const float* F = M.arena + at;
          if (derived[r]) {
            const float m = (float)(k >> 1);
            return (k & 1) ? m * F[k - 1] : -m * F[k + 1];
          }
          return F[k];
        },
        [&](int r, int c, float v) { As[r * a_stride(kChunk) + c] = v; });
    stage(
        kc, nlon, [&](int c, int x) { return k0 + c < K ? inv[(k0 + c) * nlon + x] : 0.0f; },
        [&](int c, int x, float v) { Bs[c * bs + x] = v; });
    __syncthreads();
    mfma_rows(As, Bs, bs, kc, NT, acc);
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int at = dst_at[16 * wave + 4 * (lane >> 4) + r];
    if (at >= 0) {
      float* G = M.arena + at;
#pragma unroll
      for (int nt = 0; nt < kMaxTilesN; nt++)
        if (nt < NT) G[16 * nt + (lane & 15)] = acc[nt][r];
    }
  }
}

extern "C" __global__ __launch_bounds__(kThreads) void k_dft_fwd(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                                  FwdJobs jobs) {
  __shared__ float As[kAs];
  __shared__ float Bs[kBs];
  int tile, block;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int j = find_block(jobs, tile, [&](const FwdJob& x) { return fwd_rows(M, x); }, &block);
  const FwdJob& job = jobs.j[j];
  const int nlon = M.nlon, nfc = M.nfc;
  const int rows = fwd_rows(M, job);
  const int row0 = block * kRows;
  const int K = nlon;
  const int NT = (nfc + 15) / 16;
  const int bs = b_stride(NT * 16);
  const float* fwd = M.ltab + M.tabo[TAB_fourier_forward];
  const float* G = M.arena + M.slot[job.src];
  f4 acc[kMaxTilesN];
#pragma unroll
  for (int nt = 0; nt < kMaxTilesN; nt++) acc[nt] = f4{0, 0, 0, 0};
  for (int x0 = 0; x0 < K; x0 += kChunk) {
    stage(
        kRows, kChunk,
        [&](int r, int c) {
          const int row = row0 + r;
          return row < rows ? G[row * nlon + x0 + c] : 0.0f;
        },
        [&](int r, int c, float v) { As[r * a_stride(kChunk) + c] = v; });
    stage(
        kChunk, NT * 16, [&](int c, int k) { return k < nfc ? fwd[(x0 + c) * nfc + k] : 0.0f; },
        [&](int c, int k, float v) { Bs[c * bs + k] = v; });
    __syncthreads();
    mfma_rows(As, Bs, bs, kChunk, NT, acc);
    __syncthreads();
  }
  float* F = M.arena + M.slot[job.dst];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int row = row0 + 16 * wave + 4 * (lane >> 4) + r;
    if (row < rows) {
#pragma unroll
      for (int nt = 0; nt < kMaxTilesN; nt++) {
        const int k = 16 * nt + (lane & 15);
        if (nt < NT && k < nfc) F[row * nfc + k] = acc[nt][r];
      }
    }
  }
}
