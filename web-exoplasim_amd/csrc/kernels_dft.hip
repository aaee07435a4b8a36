// Zonal transforms of the ExoPlaSim step on gfx950: real FFTs, one latitude
// row per pair of lanes.
//
// With the spectral truncation, a row's transforms are the truncated DFTs
// of FOURIER.tables (fourier/index.ts:19-36):
//   fc2gp (fourier/wgsl/inverse.wgsl:4-18):  g[x] = Re F_0 + 2 Re sum_{m=1..T} F_m e^{i m lambda_x}
//   gp2fc (fourier/wgsl/forward.wgsl:4-18):  F_m = (1/N) sum_x g[x] e^{-i m lambda_x},  m <= T
// with lambda_x = 2 pi x / N, F_m = F[2m] + i F[2m+1] (Im F_0 ignored on the
// way in), and d/dlambda (dlambda.wgsl:5-17) is fc2gp of i m F_m. Here both
// run as real FFTs of length N = NLON through a complex FFT of half the
// length, M = N / 2 (32, 64 or 96):
//   inverse: Z[k] = (X[k] + X*[M-k]) + i w^k (X[k] - X*[M-k]), k < M (X the
//            truncated spectrum, w = e^{2 pi i / N}); z = sum_k Z[k] W^{kn}
//            (W = e^{2 pi i / M}); g[2n] = Re z[n], g[2n+1] = Im z[n];
//   forward: z[n] = g[2n] + i g[2n+1]; Z = sum_n z[n] W^{-kn};
//            X[k] = (Z[k] + Z*[M-k]) / 2 - i w^{-k} (Z[k] - Z*[M-k]) / 2.
//
// The M-point complex FFT is split over the two lanes of a row, p = 0, 1,
// by decimation in time: lane p transforms the H = M / 2 inputs of parity p,
//   A_p[n] = sum_j Z[2j + p] W^{2jn},
// in its registers (radix-2 DIT with bit-reversed input, for H = 48 three
// 16-point FFTs joined by radix-3 butterflies), and the last radix-2 stage
// swaps A between the pair through DPP:
//   lane 0 keeps z[n] = A_0[n] + W^n A_1[n], lane 1 z[n + H] = A_0[n] - W^n A_1[n].
// So the inverse's lane p needs only the wavenumbers of parity p (k and M - k
// share it), and the forward's X[k], X[M - k] pair up across the lanes with
// one more DPP swap. The twiddles W^t and w^k are FP64 values rounded to
// FP32 at compile time (constants in the instruction stream, no table). The
// sums are the DFT's terms in another order: a few ulp of FP32 apart from
// the table products.
//
// Rows: a workgroup (one wave) takes 32 rows of a member group (members of
// one truncation and level count), over the jobs of the launch, each row
// with its own source and destination. Fourier slots are coefficient-major
// (coefficient k of row r at k R + r, R the slot's rows), so a lane's
// coefficient access is two contiguous runs of 32 rows per wave. Grid rows
// (longitude fastest, N floats) pass through LDS: the inverse writes 128-byte
// runs of one row per half wave, the forward reads 256-byte runs of one row
// per wave access. Neighbouring row blocks share the cache lines at their
// coefficient runs' ends, so xcd_block gives them to one XCD's L2.
#include "transforms.h"

namespace {

constexpr int kThreads = 64;  // one wave: its LDS transpositions need no cross-wave barrier
constexpr int kRows = kThreads / 2;  // two lanes a row
constexpr int kTs = 33;  // inverse staging: 32 values + 1 pad per lane
constexpr int kLs = 68;  // forward staging: 64 values + 4 pad per row (conflict-free 8-byte reads)

// ------------------------------------------------------------- twiddles

// cos and sin of 2 pi t / n in FP64 at compile time: the angle reduced to
// the nearest quarter turn, then Taylor series on |y| <= pi / 4 (terms to
// y^23, truncation error below 1e-20).
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double sin_series(double y) {
  double term = y, sum = y;
  for (int k = 1; k < 12; k++) {
    term *= -y * y / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cos_series(double y) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 12; k++) {
    term *= -y * y / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
template <int n>
struct Cis {
  float c[n], s[n];
  constexpr Cis() : c(), s() {
    for (int t = 0; t < n; t++) {
      const int q = (4 * t + n / 2) / n;  // nearest quarter turn
      const double y = kTwoPi * ((double)(4 * t - q * n) / (4.0 * n));
      const double cy = cos_series(y), sy = sin_series(y);
      const double cc[4] = {cy, -sy, -cy, sy}, ss[4] = {sy, cy, -sy, -cy};
      c[t] = (float)cc[q % 4];
      s[t] = (float)ss[q % 4];
    }
  }
};
template <int n>
constexpr Cis<n> kCis{};

// e^{2 pi i t / n}, t a compile-time index after unrolling
template <int n>
__device__ __forceinline__ float cis_re(int t) {
  return kCis<n>.c[t];
}
template <int n>
__device__ __forceinline__ float cis_im(int t) {
  return kCis<n>.s[t];
}

// ------------------------------------------------------------- FFT core

__host__ __device__ constexpr int bitrev(int k, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r = (r << 1) | ((k >> i) & 1);
  return r;
}
__host__ __device__ constexpr int log2i(int n) { return n <= 1 ? 0 : 1 + log2i(n / 2); }

// Register index of input k of an H-point FFT: bit-reversed for a power of
// two; for H = 48, sub-FFT r = k mod 3 at 16 r, its input k / 3 bit-reversed.
template <int H>
__device__ __forceinline__ constexpr int slot_of(int k) {
  if constexpr ((H & (H - 1)) == 0)
    return bitrev(k, log2i(H));
  else
    return (k % 3) * (H / 3) + bitrev(k / 3, log2i(H / 3));
}

// In-place S-point radix-2 DIT FFT (S a power of two) of re/im[off..off+S),
// input bit-reversed, output natural; twiddle e^{sign 2 pi i j / size}.
template <int S, int SIGN>
__device__ __forceinline__ void fft2(float* re, float* im, int off) {
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
          const int t = j * (S / size);
          const float wr = cis_re<S>(t), wi = SIGN * cis_im<S>(t);
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

// In-place H-point FFT with e^{SIGN 2 pi i / H}, input in slot_of order.
template <int H, int SIGN>
__device__ __forceinline__ void fft(float* re, float* im) {
  if constexpr ((H & (H - 1)) == 0) {
    fft2<H, SIGN>(re, im, 0);
  } else {
    constexpr int S = H / 3;
    fft2<S, SIGN>(re, im, 0);
    fft2<S, SIGN>(re, im, S);
    fft2<S, SIGN>(re, im, 2 * S);
    // z[n1 + S n2] = sum_r (W_H^{r n1} A_r[n1]) W_3^{r n2}
    const float h3 = SIGN * 0.866025403784438647f;  // sin(2 pi / 3)
#pragma unroll
    for (int n1 = 0; n1 < S; n1++) {
      const float t0r = re[n1], t0i = im[n1];
      float t1r = re[S + n1], t1i = im[S + n1], t2r = re[2 * S + n1], t2i = im[2 * S + n1];
      if (n1 != 0) {
        const float w1r = cis_re<H>(n1), w1i = SIGN * cis_im<H>(n1);
        const float w2r = cis_re<H>(2 * n1), w2i = SIGN * cis_im<H>(2 * n1);
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

// The other lane of the pair (lane ^ 1), through DPP quad_perm [1, 0, 3, 2].
__device__ __forceinline__ float partner(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// The pair's last radix-2 stage: lane 0 keeps A_0[n] + W^n A_1[n], lane 1
// A_0[n] - W^n A_1[n] (W = e^{SIGN 2 pi i / M}, M = 2H).
template <int H, int SIGN>
__device__ __forceinline__ void join_pair(float* re, float* im, int p) {
#pragma unroll
  for (int n = 0; n < H; n++) {
    const float qr = partner(re[n]), qi = partner(im[n]);
    const float er = p ? qr : re[n], ei = p ? qi : im[n];
    const float orr = p ? re[n] : qr, oi = p ? im[n] : qi;
    float tr = orr, ti = oi;
    if (n != 0) {
      const float wr = cis_re<2 * H>(n), wi = SIGN * cis_im<2 * H>(n);
      tr = orr * wr - oi * wi;
      ti = orr * wi + oi * wr;
    }
    re[n] = p ? er - tr : er + tr;
    im[n] = p ? ei - ti : ei + ti;
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


// A row pointer of the wave's table, made wave-uniform (scalar registers).
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* const* table, int i) {
  const unsigned long long v = (unsigned long long)table[i];
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}

// ---------------------------------------------------------- inverse

// Lane p of a row. A pair past the launch's rows reads the group's first
// arena floats (stride 0) and stores nothing (null destination).
template <int M>
__device__ __forceinline__ void inverse_row(const float* source, int stride, bool derived, float* const* dstp,
                                            float* T) {
  constexpr int N = 2 * M, H = M / 2, T1 = wavenumbers<M>(), TR = T1 - 1;
  constexpr int J = TR / 2 + 1;  // wavenumbers 2j + p, j < J, cover 0..T
  const gfloat* src = (const gfloat*)source;
  const int lane = threadIdx.x & 63, p = lane & 1;
  // X[2j + p], zero past T (the address clamped to T)
  float xr[J], xi[J];
#pragma unroll
  for (int j = 0; j < J; j++) {
    const int m = 2 * j + p, mc = m <= TR ? m : TR;
    const float a = src[(2 * mc) * stride], b = src[(2 * mc + 1) * stride];
    const float fm = (float)m;
    float vr = derived ? -fm * b : a;  // i m X for d/dlambda
    float vi = derived ? fm * a : b;
    if (m == 0) vi = 0.0f;
    xr[j] = m <= TR ? vr : 0.0f;
    xi[j] = m <= TR ? vi : 0.0f;
  }
  const auto x_at = [&](int j, float* r, float* i) {  // X[2j + p] for a register index j
    *r = j < J ? xr[j] : 0.0f;
    *i = j < J ? xi[j] : 0.0f;
  };
  // Z[2j + p] = (X[k] + X*[M-k]) + i w^k (X[k] - X*[M-k]), k = 2j + p,
  // M - k = 2 (H - j - p) + p
  float re[H], im[H];
#pragma unroll
  for (int j = 0; j < H; j++) {
    float ar, ai, b0r, b0i, b1r, b1i;
    x_at(j, &ar, &ai);
    x_at(H - j, &b0r, &b0i);  // p = 0 (X[M] = 0 at j = 0)
    x_at(H - j - 1, &b1r, &b1i);
    const float br = p ? b1r : b0r, bi = p ? b1i : b0i;
    const float sr = ar + br, si = ai - bi, dr = ar - br, di = ai + bi;
    // w^k, k = 2j + p
    const float wr = p ? cis_re<N>(2 * j + 1) : cis_re<N>(2 * j);
    const float wi = p ? cis_im<N>(2 * j + 1) : cis_im<N>(2 * j);
    const float pr = dr * wr - di * wi, pi = dr * wi + di * wr;
    re[slot_of<H>(j)] = sr - pi;
    im[slot_of<H>(j)] = si + pr;
  }
  fft<H, 1>(re, im);
  join_pair<H, 1>(re, im, p);
  // lane p holds z[pH + n] = g[pM + 2n], g[pM + 2n + 1]; out through LDS in
  // chunks of 16 values of n (32 floats)
#pragma unroll
  for (int c = 0; c < H / 16; c++) {
#pragma unroll
    for (int n = 0; n < 16; n++) {
      T[lane * kTs + 2 * n] = re[16 * c + n];
      T[lane * kTs + 2 * n + 1] = im[16 * c + n];
    }
    lds_barrier();
    // row i's lanes 2i (first half of the row) and 2i + 1 (second half)
    const int h = lane >> 5, x = lane & 31;
#pragma unroll 8
    for (int i = 0; i < kRows; i++) {
      gfloat* G = (gfloat*)uniform_ptr(dstp, i);
      if (G) G[h * M + 32 * c + x] = T[(2 * i + h) * kTs + x];
    }
    lds_barrier();
  }
}

// ---------------------------------------------------------- forward

// Every row has a source (a pair past the launch's rows reads the group's
// first arena floats); only rows with a destination store.
template <int M>
__device__ __forceinline__ void forward_row(const float* const* srcp, float* destination, int stride, float* L) {
  constexpr int N = 2 * M, H = M / 2, T1 = wavenumbers<M>();
  const int lane = threadIdx.x & 63, p = lane & 1, r = lane >> 1;
  float re[H], im[H];
  // in through LDS, 64 values of each row at a time; lane p takes
  // z[2j + p] = g[4j + 2p] + i g[4j + 2p + 1]
#pragma unroll
  for (int c = 0; c < N / 64; c++) {
#pragma unroll
    for (int h = 0; h < kRows; h += 16) {
      float v[16];
#pragma unroll
      for (int i = 0; i < 16; i++) v[i] = ((const gfloat*)uniform_ptr(srcp, h + i))[64 * c + lane];
#pragma unroll
      for (int i = 0; i < 16; i++) L[(h + i) * kLs + lane] = v[i];
    }
    lds_barrier();
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const float2 z = *(const float2*)&L[r * kLs + 4 * j + 2 * p];
      re[slot_of<H>(16 * c + j)] = z.x;
      im[slot_of<H>(16 * c + j)] = z.y;
    }
    lds_barrier();
  }
  fft<H, -1>(re, im);
  join_pair<H, -1>(re, im, p);
  // lane p holds Z[pH + i]. Step i pairs lane 0's Z[i] with lane 1's
  // Z[M - i] (register H - i) and lane 1's Z[H + i] with lane 0's Z[H - i];
  // at i = 0 each lane pairs Z[pH] with itself.
  if (!destination) return;
  gfloat* dst = (gfloat*)destination;
  const float scale = 1.0f / (float)N;
#pragma unroll
  for (int i = 0; i < H; i++) {
    const float zr = re[i], zi = im[i];
    float yr = zr, yi = -zi;  // y = Z*[M - k]
    if (i != 0) {
      yr = partner(re[H - i]);
      yi = -partner(im[H - i]);
    }
    const float er = 0.5f * (zr + yr), ei = 0.5f * (zi + yi);
    const float dr = 0.5f * (zr - yr), di = 0.5f * (zi - yi);
    // -i w^{-k} d, k = pH + i, w^H = i: lane 0 (-wi - i wr) d, lane 1 (-wr + i wi) d
    const float wr = cis_re<N>(i), wi = cis_im<N>(i);
    const float cr = p ? -wr : -wi, ci = p ? wi : -wr;
    const float orr = cr * dr - ci * di, oi = cr * di + ci * dr;
    if (i < T1 - H || p == 0) {
      const int k = p * H + i;
      dst[(2 * k) * stride] = (er + orr) * scale;
      dst[(2 * k + 1) * stride] = (ei + oi) * scale;
    }
  }
}

// One wave per workgroup, 32 rows of a group whose grids have 2M longitudes
// (the kernel's LDS, shared by its three widths, passed in).
template <int M>
__device__ __forceinline__ void dft_inv(const DevMember* __restrict__ mems, const GroupPlan& plan,
                                        const InvJobs& jobs, int grp, int block, float* trans, float** dst) {
  const DevMember& M0 = mems[plan.list[plan.start[grp]]];
  const int nlat = M0.nlat, nlon = M0.nlon;
  const auto rows_of = [&](const InvJob& x) { return inv_rows(M0, x); };
  int per_member = 0;
  for (int j = 0; j < jobs.n; j++) per_member += rows_of(jobs.j[j]);
  const int total = plan.count[grp] * per_member;
  const int row = block * kRows + (threadIdx.x >> 1);
  const float* src = dev_arena(M0);
  float* d = nullptr;
  int stride = 0;
  bool derived = false;
  if (row < total) {
    int g, r;
    const InvJob& job = jobs.j[locate(jobs, per_member, row, rows_of, &g, &r)];
    const DevMember& G = mems[plan.list[plan.start[grp] + g]];
    const int plane_rows = (job.per_level ? M0.nlev : 1) * nlat;  // rows of one output
    const int o = r / plane_rows, pr = r - o * plane_rows;
    src = dev_arena(G) + G.slot[(job.kind == INV_DV2UV && o == 1) ? job.four1 : job.four0] + pr;
    d = dev_arena(G) + G.slot[o == 0 ? job.dst0 : job.dst1] + pr * nlon;
    derived = job.kind == INV_DLAMBDA && o == 1;
    stride = plane_rows;
  }
  if ((threadIdx.x & 1) == 0) dst[threadIdx.x >> 1] = d;
  lds_barrier();
  inverse_row<M>(src, stride, derived, dst, trans);
}

template <int M>
__device__ __forceinline__ void dft_fwd(const DevMember* __restrict__ mems, const GroupPlan& plan,
                                        const FwdJobs& jobs, int grp, int block, float* stage, const float** src) {
  const DevMember& M0 = mems[plan.list[plan.start[grp]]];
  const int nlon = M0.nlon;
  const auto rows_of = [&](const FwdJob& x) { return fwd_rows(M0, x); };
  int per_member = 0;
  for (int j = 0; j < jobs.n; j++) per_member += rows_of(jobs.j[j]);
  const int total = plan.count[grp] * per_member;
  const int row = block * kRows + (threadIdx.x >> 1);
  const float* s = dev_arena(M0);
  float* d = nullptr;
  int stride = 0;
  if (row < total) {
    int g, r;
    const FwdJob& job = jobs.j[locate(jobs, per_member, row, rows_of, &g, &r)];
    const DevMember& G = mems[plan.list[plan.start[grp] + g]];
    s = dev_arena(G) + G.slot[job.src] + r * nlon;
    d = dev_arena(G) + G.slot[job.dst] + r;
    stride = rows_of(job);
  }
  if ((threadIdx.x & 1) == 0) src[threadIdx.x >> 1] = s;
  lds_barrier();
  forward_row<M>(src, d, stride, stage);
}

}  // namespace

// One launch over every member of the batch, T21, T42 and T63 groups side by
// side (heaviest first): the launch's many workgroups overlap one another's
// loads, transforms and stores, which one launch per grid width, about one
// round of workgroups each, cannot.
extern "C" __global__ __launch_bounds__(kThreads) void k_dft_inv(const DevMember* __restrict__ mems, GroupPlan plan,
                                                                  InvJobs jobs) {
  __shared__ float trans[64 * kTs];
  __shared__ float* dst[kRows];
  int block;
  const int grp = group_find(plan, xcd_block(blockIdx.x, gridDim.x), &block);
  const int nlon = mems[plan.list[plan.start[grp]]].nlon;
  if (nlon == 64)
    dft_inv<32>(mems, plan, jobs, grp, block, trans, dst);
  else if (nlon == 128)
    dft_inv<64>(mems, plan, jobs, grp, block, trans, dst);
  else
    dft_inv<96>(mems, plan, jobs, grp, block, trans, dst);
}

extern "C" __global__ __launch_bounds__(kThreads) void k_dft_fwd(const DevMember* __restrict__ mems, GroupPlan plan,
                                                                  FwdJobs jobs) {
  __shared__ float stage[kRows * kLs];
  __shared__ const float* src[kRows];
  int block;
  const int grp = group_find(plan, xcd_block(blockIdx.x, gridDim.x), &block);
  const int nlon = mems[plan.list[plan.start[grp]]].nlon;
  if (nlon == 64)
    dft_fwd<32>(mems, plan, jobs, grp, block, stage, src);
  else if (nlon == 128)
    dft_fwd<64>(mems, plan, jobs, grp, block, stage, src);
  else
    dft_fwd<96>(mems, plan, jobs, grp, block, stage, src);
}
