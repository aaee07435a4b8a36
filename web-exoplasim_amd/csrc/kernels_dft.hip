// Zonal transforms of the ExoPlaSim step on gfx950: real FFTs, one latitude
// row per lane.
//
// With the spectral truncation, a row's transforms are the truncated DFTs
// of FOURIER.tables (fourier/index.ts:19-36):
//   fc2gp (fourier/wgsl/inverse.wgsl:4-18):  g[x] = Re F_0 + 2 Re sum_{m=1..T} F_m e^{i m lambda_x}
//   gp2fc (fourier/wgsl/forward.wgsl:4-18):  F_m = (1/N) sum_x g[x] e^{-i m lambda_x},  m <= T
// with lambda_x = 2 pi x / N, F_m = F[2m] + i F[2m+1] (Im F_0 ignored on the
// way in), and d/dlambda (dlambda.wgsl:5-17) is fc2gp of i m F_m. Here both
// run as real FFTs of length N = NLON through a complex FFT of half the
// length, M = N / 2 (32, 64 or 96 = 3 x 32), in the lane's registers:
//   inverse: Z[k] = (X[k] + X*[M-k]) + i w^k (X[k] - X*[M-k]), k < M (X the
//            truncated spectrum, w = e^{2 pi i / N}); z = sum_k Z[k] W^{kn}
//            (W = e^{2 pi i / M}); g[2n] = Re z[n], g[2n+1] = Im z[n];
//   forward: z[n] = g[2n] + i g[2n+1]; Z = sum_n z[n] W^{-kn};
//            X[k] = (Z[k] + Z*[M-k]) / 2 - i w^{-k} (Z[k] - Z*[M-k]) / 2.
// The complex FFTs are radix-2 decimation in time (input in bit-reversed
// register order), for M = 96 three 32-point FFTs joined by radix-3
// butterflies. The twiddles W^t and w^k come from the truncation's FP64
// set-up as an FP32 table (fft_twiddle) staged in LDS. The sums are the
// DFT's terms in another order: a few ulp of FP32 apart from the table
// products.
//
// Rows: lane = row; a workgroup (one wave) takes 64 rows of a member group
// (members of one truncation and level count), over the jobs of the launch,
// each row with its own source and destination. Fourier slots are
// coefficient-major (coefficient k of row r at k R + r, R the slot's rows),
// so a lane-per-row coefficient access is one contiguous run per wave.
// Grid rows are longitude-fastest rows of N floats; they pass through LDS
// in 32-value column chunks, each wave moving 2 rows x 128 bytes per access.
#include "transforms.h"

namespace {

constexpr int kThreads = 64;  // one wave: its LDS transpositions need no cross-wave barrier
constexpr int kRows = kThreads;  // one row per lane
constexpr int kMaxM = 96;
constexpr int kTs = 33;  // transposition rows: 32 values + 1 pad, conflict-free

// ------------------------------------------------------------- FFT core

__host__ __device__ constexpr int bitrev(int k, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r = (r << 1) | ((k >> i) & 1);
  return r;
}
__host__ __device__ constexpr int log2i(int n) { return n <= 1 ? 0 : 1 + log2i(n / 2); }

// Register index of input k of an M-point FFT: bit-reversed for a power of
// two; for M = 96, sub-FFT r = k mod 3 at 32 r, its input k / 3 bit-reversed.
template <int M>
__device__ __forceinline__ constexpr int slot_of(int k) {
  if constexpr ((M & (M - 1)) == 0)
    return bitrev(k, log2i(M));
  else
    return (k % 3) * (M / 3) + bitrev(k / 3, log2i(M / 3));
}

// In-place S-point radix-2 DIT FFT (S a power of two) of re/im[off..off+S),
// input bit-reversed, output natural; twiddle e^{sign 2 pi i j / S} is
// tw[(j * (MT / S))] of the MT-point table (cos, sin interleaved).
template <int S, int MT, int SIGN>
__device__ __forceinline__ void fft2(float* re, float* im, int off, const float* tw) {
#pragma unroll
  for (int size = 2; size <= S; size *= 2) {
    const int h = size / 2;
#pragma unroll
    for (int start = 0; start < S; start += size) {
#pragma unroll
      for (int j = 0; j < h; j++) {
        const int a = off + start + j, b = a + h;
        float tr = re[b], ti = im[b];
        if (j != 0) {
          const int t = j * (MT / size);
          const float wr = tw[2 * t], wi = SIGN * tw[2 * t + 1];
          const float xr = tr * wr - ti * wi;
          ti = tr * wi + ti * wr;
          tr = xr;
        }
        re[b] = re[a] - tr;
        im[b] = im[a] - ti;
        re[a] = re[a] + tr;
        im[a] = im[a] + ti;
      }
    }
  }
}

// In-place M-point FFT with e^{SIGN 2 pi i / M}, input in slot_of order.
template <int M, int SIGN>
__device__ __forceinline__ void fft(float* re, float* im, const float* tw) {
  if constexpr ((M & (M - 1)) == 0) {
    fft2<M, M, SIGN>(re, im, 0, tw);
  } else {
    constexpr int S = M / 3;
    fft2<S, M, SIGN>(re, im, 0, tw);
    fft2<S, M, SIGN>(re, im, S, tw);
    fft2<S, M, SIGN>(re, im, 2 * S, tw);
    // z[n1 + S n2] = sum_r (W^{r n1} A_r[n1]) W3^{r n2}
    const float h3 = SIGN * 0.866025403784438647f;  // sin(2 pi / 3)
#pragma unroll
    for (int n1 = 0; n1 < S; n1++) {
      const float t0r = re[n1], t0i = im[n1];
      float t1r = re[S + n1], t1i = im[S + n1], t2r = re[2 * S + n1], t2i = im[2 * S + n1];
      if (n1 != 0) {
        const float w1r = tw[2 * n1], w1i = SIGN * tw[2 * n1 + 1];
        const float w2r = tw[4 * n1], w2i = SIGN * tw[4 * n1 + 1];
        const float a = t1r * w1r - t1i * w1i;
        t1i = t1r * w1i + t1i * w1r;
        t1r = a;
        const float b = t2r * w2r - t2i * w2i;
        t2i = t2r * w2i + t2i * w2r;
        t2r = b;
      }
      const float sr = t1r + t2r, si = t1i + t2i, dr = t1r - t2r, di = t1i - t2i;
      const float mr = t0r - 0.5f * sr, mi = t0i - 0.5f * si;
      re[n1] = t0r + sr;
      im[n1] = t0i + si;
      re[S + n1] = mr - h3 * di;
      im[S + n1] = mi + h3 * dr;
      re[2 * S + n1] = mr + h3 * di;
      im[2 * S + n1] = mi - h3 * dr;
    }
  }
}

// ------------------------------------------------------------- rows

// Global-memory views of the arena: the row pointers pass through LDS,
// which would otherwise leave the compiler generic (flat) accesses.
typedef __attribute__((address_space(1))) float gfloat;

// Wavenumbers 0..T of the truncation whose grid has 2M longitudes
// (T21: 64, T42: 128, T63: 192).
template <int M>
__device__ __forceinline__ constexpr int wavenumbers() {
  return M == 32 ? 22 : M == 64 ? 43 : 64;
}

// Rows of an inverse job: outputs x planes x latitudes.
__device__ __forceinline__ int inv_rows(const DevMember& M, const InvJob& job) {
  return (job.kind == INV_PLAIN ? 1 : 2) * (job.per_level ? M.nlev : 1) * M.nlat;
}
__device__ __forceinline__ int fwd_rows(const DevMember& M, const FwdJob& job) {
  return (job.per_level ? M.nlev : 1) * M.nlat;
}

// Row `row` of a group's launch: member g = row / (rows per member), then
// the job and the row within it.
template <typename Jobs, typename Rows>
__device__ __forceinline__ int locate(const Jobs& jobs, int per_member, int row, Rows rows, int* g, int* r) {
  *g = row / per_member;
  int x = row - *g * per_member, j = 0;
  for (; j + 1 < jobs.n; j++) {
    const int n = rows(jobs.j[j]);
    if (x < n) break;
    x -= n;
  }
  *r = x;
  return j;
}

// The truncation's twiddles into LDS: W_M^t (t < M), then w_N^k (k < M).
__device__ __forceinline__ const float* twiddle_table(const DevMember& M) {
  return M.ltab + M.tabo[TAB_fft_twiddle];
}
__device__ __forceinline__ void stage_twiddles(const DevMember& M, float* tw) {
  const float* t = twiddle_table(M);
  for (int i = threadIdx.x; i < 2 * M.nlon; i += kThreads) tw[i] = t[i];
}

// ---------------------------------------------------------- inverse

// A lane past the launch's rows reads the twiddle table (stride 0) and
// stores nothing (null destination).
template <int M>
__device__ __forceinline__ void inverse_row(const float* source, int stride, bool derived, float* const* dstp,
                                            float* T, const float* tw) {
  constexpr int N = 2 * M, T1 = wavenumbers<M>();
  const gfloat* src = (const gfloat*)source;
  const float* wn = tw + N;  // w^k
  float re[M], im[M];
  // X[m] of this row, zero past T
  const auto x_of = [&](int m, float* xr, float* xi) {
    if (m >= T1) {
      *xr = 0.0f;
      *xi = 0.0f;
      return;
    }
    const float a = src[(2 * m) * stride], b = m == 0 ? 0.0f : src[(2 * m + 1) * stride];
    *xr = derived ? -(float)m * b : a;  // i m X for d/dlambda
    *xi = derived ? (float)m * a : b;
    if (m == 0) *xi = 0.0f;
  };
  // Z[k] and Z[M - k] from X[k] and X[M - k]
#pragma unroll
  for (int k = 0; k <= M / 2; k++) {
    float ar, ai, br, bi;
    x_of(k, &ar, &ai);
    if (k == 0) {
      br = 0.0f;
      bi = 0.0f;
    } else {
      x_of(M - k, &br, &bi);
    }
    // Z[k] = (X[k] + X*[M-k]) + i w^k (X[k] - X*[M-k])
    {
      const float sr = ar + br, si = ai - bi, dr = ar - br, di = ai + bi;
      const float wr = wn[2 * k], wi = wn[2 * k + 1];
      const float pr = dr * wr - di * wi, pi = dr * wi + di * wr;
      re[slot_of<M>(k)] = sr - pi;
      im[slot_of<M>(k)] = si + pr;
    }
    if (k != 0 && k != M / 2) {
      // Z[M-k] = (X[M-k] + X*[k]) + i w^{M-k} (X[M-k] - X*[k])
      const int j = M - k;
      const float sr = br + ar, si = bi - ai, dr = br - ar, di = bi + ai;
      const float wr = wn[2 * j], wi = wn[2 * j + 1];
      const float pr = dr * wr - di * wi, pi = dr * wi + di * wr;
      re[slot_of<M>(j)] = sr - pi;
      im[slot_of<M>(j)] = si + pr;
    }
  }
  fft<M, 1>(re, im, tw);
  // out through LDS in chunks of 32 values (16 complex): the wave then
  // writes 2 rows x 32 values per access
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < N / 32; c++) {
#pragma unroll
    for (int n = 0; n < 16; n++) {
      T[lane * kTs + 2 * n] = re[16 * c + n];
      T[lane * kTs + 2 * n + 1] = im[16 * c + n];
    }
    __syncthreads();
#pragma unroll 8
    for (int i = 0; i < 32; i++) {
      const int row = 2 * i + (lane >> 5), x = lane & 31;
      gfloat* G = (gfloat*)dstp[row];
      if (G) G[32 * c + x] = T[row * kTs + x];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------- forward

// Every lane has a source row (a lane past the launch's rows reads the
// twiddle table, 2N floats); only lanes with a destination store.
template <int M>
__device__ __forceinline__ void forward_row(const float* const* srcp, float* destination, int stride, float* T,
                                            const float* tw) {
  constexpr int N = 2 * M, T1 = wavenumbers<M>();
  const float* wn = tw + N;
  const int lane = threadIdx.x & 63;
  float re[M], im[M];
  // in through LDS in chunks of 32 values: z[n] = g[2n] + i g[2n+1]; each
  // chunk's 32 loads go in two batches of 16 to bound the live registers
#pragma unroll
  for (int c = 0; c < N / 32; c++) {
#pragma unroll
    for (int h = 0; h < 32; h += 16) {
      float v[16];
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const gfloat* G = (const gfloat*)srcp[2 * (h + i) + (lane >> 5)];
        v[i] = G[32 * c + (lane & 31)];
      }
#pragma unroll
      for (int i = 0; i < 16; i++) T[(2 * (h + i) + (lane >> 5)) * kTs + (lane & 31)] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < 16; n++) {
      re[slot_of<M>(16 * c + n)] = T[lane * kTs + 2 * n];
      im[slot_of<M>(16 * c + n)] = T[lane * kTs + 2 * n + 1];
    }
    __syncthreads();
  }
  fft<M, -1>(re, im, tw);
  if (!destination) return;
  gfloat* dst = (gfloat*)destination;
  const float scale = 1.0f / (float)N;
  // X[k] = (Z[k] + Z*[M-k]) / 2 - i w^{-k} (Z[k] - Z*[M-k]) / 2, k <= T
#pragma unroll
  for (int k = 0; k < T1; k++) {
    const int j = k == 0 ? 0 : M - k;
    const float zr = re[k], zi = im[k], yr = re[j], yi = -im[j];  // y = Z*[M-k]
    const float er = 0.5f * (zr + yr), ei = 0.5f * (zi + yi);
    const float dr = 0.5f * (zr - yr), di = 0.5f * (zi - yi);
    // -i w^{-k} d = -i (wr - i wi) d = (-wi - i wr) d
    const float wr = wn[2 * k], wi = wn[2 * k + 1];
    const float orr = -wi * dr + wr * di, oi = -wi * di - wr * dr;
    dst[(2 * k) * stride] = (er + orr) * scale;
    dst[(2 * k + 1) * stride] = (ei + oi) * scale;
  }
}

}  // namespace

// One wave per SIMD: the 96-point rows hold 192 values a lane.
extern "C" __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1))) void k_dft_inv(
    const DevMember* __restrict__ mems, GroupPlan plan, InvJobs jobs) {
  __shared__ float tw[4 * kMaxM];
  __shared__ float trans[64 * kTs];
  __shared__ float* dst[kRows];
  int block;
  const int grp = group_find(plan, blockIdx.x, &block);
  const DevMember& M = mems[plan.list[plan.start[grp]]];
  const int nlat = M.nlat, nlon = M.nlon;
  const auto rows_of = [&](const InvJob& x) { return inv_rows(M, x); };
  int per_member = 0;
  for (int j = 0; j < jobs.n; j++) per_member += rows_of(jobs.j[j]);
  const int total = plan.count[grp] * per_member;
  const int row = block * kRows + threadIdx.x;
  const float* src = twiddle_table(M);
  float* d = nullptr;
  int stride = 0;
  bool derived = false;
  if (row < total) {
    int g, r;
    const InvJob& job = jobs.j[locate(jobs, per_member, row, rows_of, &g, &r)];
    const DevMember& G = mems[plan.list[plan.start[grp] + g]];
    const int plane_rows = (job.per_level ? M.nlev : 1) * nlat;  // rows of one output
    const int o = r / plane_rows, pr = r - o * plane_rows;
    src = G.arena + G.slot[(job.kind == INV_DV2UV && o == 1) ? job.four1 : job.four0] + pr;
    d = G.arena + G.slot[o == 0 ? job.dst0 : job.dst1] + pr * nlon;
    derived = job.kind == INV_DLAMBDA && o == 1;
    stride = plane_rows;
  }
  dst[threadIdx.x] = d;
  stage_twiddles(M, tw);
  __syncthreads();
  if (nlon == 64)
    inverse_row<32>(src, stride, derived, dst, trans, tw);
  else if (nlon == 128)
    inverse_row<64>(src, stride, derived, dst, trans, tw);
  else
    inverse_row<96>(src, stride, derived, dst, trans, tw);
}

extern "C" __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1))) void k_dft_fwd(
    const DevMember* __restrict__ mems, GroupPlan plan, FwdJobs jobs) {
  __shared__ float tw[4 * kMaxM];
  __shared__ float trans[64 * kTs];
  __shared__ const float* src[kRows];
  int block;
  const int grp = group_find(plan, blockIdx.x, &block);
  const DevMember& M = mems[plan.list[plan.start[grp]]];
  const int nlon = M.nlon;
  const auto rows_of = [&](const FwdJob& x) { return fwd_rows(M, x); };
  int per_member = 0;
  for (int j = 0; j < jobs.n; j++) per_member += rows_of(jobs.j[j]);
  const int total = plan.count[grp] * per_member;
  const int row = block * kRows + threadIdx.x;
  const float* s = twiddle_table(M);
  float* d = nullptr;
  int stride = 0;
  if (row < total) {
    int g, r;
    const FwdJob& job = jobs.j[locate(jobs, per_member, row, rows_of, &g, &r)];
    const DevMember& G = mems[plan.list[plan.start[grp] + g]];
    s = G.arena + G.slot[job.src] + r * nlon;
    d = G.arena + G.slot[job.dst] + r;
    stride = rows_of(job);
  }
  src[threadIdx.x] = s;
  stage_twiddles(M, tw);
  __syncthreads();
  if (nlon == 64)
    forward_row<32>(src, d, stride, trans, tw);
  else if (nlon == 128)
    forward_row<64>(src, d, stride, trans, tw);
  else
    forward_row<96>(src, d, stride, trans, tw);
}
