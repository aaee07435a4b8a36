// Zonal DFTs of the ExoPlaSim step on gfx950 matrix cores.
//
// With the spectral truncation, a latitude row's zonal transform is a
// product with the truncated DFT matrix of its truncation (the FP32 tables
// FOURIER.tables builds, fourier/index.ts:19-36):
//   fc2gp (fourier/wgsl/inverse.wgsl:4-18):  g[x] = sum_k F[k] inv[k][x]
//   gp2fc (fourier/wgsl/forward.wgsl:4-18):  F[k] = sum_x g[x] fwd[x][k]
// and d/dlambda (dlambda.wgsl:5-17) is fc2gp of (-m Im F, m Re F); k = 2m
// + part. Half a period on, wavenumber m's table entries change sign by
// (-1)^m (inv[k][x + N/2] = (-1)^m inv[k][x]), so the inverse product is
// split by the parity of m into two half-width ones: E[x] = sum over even
// m, O[x] = sum over odd m (x < N/2); g[x] = E + O, g[x + N/2] = E - O:
// the same terms as the full product, in another order, for half its
// multiply-adds. The forward product splits the same way: half-row sums
// give the even wavenumbers, differences the odd ones. The tables are the
// right operand, shared by every row, job and member of a truncation.
//
// Fourier slots are coefficient-major (coefficient k of row r at k R + r,
// R the slot's rows): the inverse stages a chunk with lanes over its 128
// rows, and the forward writes its results through LDS one coefficient of
// 64 consecutive rows at a time, so both move contiguous runs.
//
// A workgroup (8 waves) takes 128 rows. Rows run over the jobs of a launch
// and the members of a group (members of one truncation and level count,
// so one table): each row has its own source and destination pointers.
// The reduction runs in chunks of 32: while the waves run
// v_mfma_f32_16x16x4_f32 over their 16 rows and the column tiles of chunk
// c from LDS, the loads of chunk c + 1 are in flight into registers.
#include "transforms.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 512;
constexpr int kRows = 128;  // rows per workgroup
constexpr int kChunk = 32;  // reduction chunk
constexpr int kInvTiles = 96 / 16;  // inverse: N/2 <= 96 columns per half
constexpr int kFwdTiles = 128 / 16;  // forward: NFC <= 128 columns

__host__ __device__ constexpr int a_stride(int r) { return (r + 29) / 32 * 32 + 2; }
__host__ __device__ constexpr int b_stride(int c) { return (c + 15) / 32 * 32 + 16; }
constexpr int kAs = kRows * a_stride(kChunk);
constexpr int kInvBs = kChunk * b_stride(96);
constexpr int kFwdBs = kChunk * b_stride(128);

// Chunk staging through registers: element e = threadIdx.x + kThreads j is
// row e >> LOGC, column e & (2^LOGC - 1).
constexpr int kAN = kRows * kChunk / kThreads;  // A chunk: 128 rows x 32
constexpr int kBN = kChunk * 128 / kThreads;    // B chunk: 32 x <= 128
static_assert(kRows * kChunk % kThreads == 0 && kChunk * 128 % kThreads == 0, "chunk staging");

template <int N, int LOGC, typename Load>
__device__ __forceinline__ void load_regs(int rows, int cols, Load load, float* v) {
#pragma unroll
  for (int j = 0; j < N; j++) {
    const int e = threadIdx.x + kThreads * j, r = e >> LOGC, c = e & ((1 << LOGC) - 1);
    v[j] = (r < rows && c < cols) ? load(r, c) : 0.0f;
  }
}

template <int N, int LOGC, typename Store>
__device__ __forceinline__ void store_regs(int rows, int cols, Store store, const float* v) {
#pragma unroll
  for (int j = 0; j < N; j++) {
    const int e = threadIdx.x + kThreads * j, r = e >> LOGC, c = e & ((1 << LOGC) - 1);
    if (r < rows && c < cols) store(r, c, v[j]);
  }
}

// A chunk staged with lanes over its 128 rows: element e = threadIdx.x +
// kThreads j is row e % 128, column e / 128, so a wave reads one
// coefficient of 64 consecutive rows, contiguous in a coefficient-major
// Fourier slot.
template <int N, typename Load>
__device__ __forceinline__ void load_rows(Load load, float* v) {
#pragma unroll
  for (int j = 0; j < N; j++) {
    const int e = threadIdx.x + kThreads * j;
    v[j] = load(e & (kRows - 1), e / kRows);
  }
}

template <int N, typename Store>
__device__ __forceinline__ void store_rows(Store store, const float* v) {
#pragma unroll
  for (int j = 0; j < N; j++) {
    const int e = threadIdx.x + kThreads * j;
    store(e & (kRows - 1), e / kRows, v[j]);
  }
}
static_assert(kRows * kChunk == kAN * kThreads && (kRows & (kRows - 1)) == 0, "row staging");

// Coefficient index k of the j-th coefficient of the wavenumbers of one
// parity (0: m = 0, 2, ..; 1: m = 1, 3, ..).
__device__ __forceinline__ int parity_k(int parity, int j) { return 4 * (j >> 1) + 2 * parity + (j & 1); }
// Coefficients of one parity: 2 x the number of such m <= T.
__device__ __forceinline__ int parity_count(int parity, int ntp1) { return 2 * ((ntp1 + 1 - parity) / 2); }

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

// acc[nt] += A[rows 16 wave .. +16][0..32) x B[0..32)[16 nt .. +16], nt < NT
template <int MAXT>
__device__ __forceinline__ void mfma_rows(const float* As, const float* Bs, int bs, int NT, f4* acc) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  const float* a = As + (16 * wave + i) * a_stride(kChunk) + kk;
  const float* b = Bs + kk * bs + i;
#pragma unroll 2
  for (int r = 0; r < kChunk; r += 4) {
    const float av = a[r];
#pragma unroll
    for (int nt = 0; nt < MAXT; nt++)
      if (nt < NT) acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b[r * bs + 16 * nt], acc[nt], 0, 0, 0);
  }
}

}  // namespace

extern "C" __global__ __launch_bounds__(kThreads) void k_dft_inv(const DevMember* __restrict__ mems, GroupPlan plan,
                                                                  InvJobs jobs) {
  __shared__ float As[kAs];
  __shared__ float Bs[kInvBs];
  // per row of the block: its Fourier row (null past the group's rows),
  // whether it is d/dlambda's, and its grid row
  __shared__ const float* src[kRows];
  __shared__ float* dst[kRows];
  __shared__ int derived[kRows];
  __shared__ int stride[kRows];  // rows of the row's Fourier slot: coefficient k at src + k * stride
  int block;
  const int grp = group_find(plan, blockIdx.x, &block);
  const DevMember& M = mems[plan.list[plan.start[grp]]];
  const int nlat = M.nlat, nlon = M.nlon, ntp1 = M.ntru + 1;
  const auto rows_of = [&](const InvJob& x) { return inv_rows(M, x); };
  int per_member = 0;
  for (int j = 0; j < jobs.n; j++) per_member += rows_of(jobs.j[j]);
  const int total = plan.count[grp] * per_member;
  const int row0 = block * kRows;
  const int half = nlon / 2, NT = half / 16;
  const int bs = b_stride(half);
  const int KE = parity_count(0, ntp1), KO = parity_count(1, ntp1);
  const int CE = (KE + kChunk - 1) / kChunk, chunks = CE + (KO + kChunk - 1) / kChunk;
  const float* inv = M.ltab + M.tabo[TAB_fourier_inverse];
  if (threadIdx.x < kRows) {
    const int row = row0 + threadIdx.x;
    const float* s = nullptr;
    float* d = nullptr;
    int dl = 0, st = 0;
    if (row < total) {
      int g, r;
      const InvJob& job = jobs.j[locate(jobs, per_member, row, rows_of, &g, &r)];
      const DevMember& G = mems[plan.list[plan.start[grp] + g]];
      const int plane_rows = (job.per_level ? M.nlev : 1) * nlat;  // rows of one output
      const int o = r / plane_rows, pr = r - o * plane_rows;
      s = G.arena + G.slot[(job.kind == INV_DV2UV && o == 1) ? job.four1 : job.four0] + pr;
      d = G.arena + G.slot[o == 0 ? job.dst0 : job.dst1] + pr * nlon;
      dl = job.kind == INV_DLAMBDA && o == 1;
      st = plane_rows;
    }
    src[threadIdx.x] = s;
    dst[threadIdx.x] = d;
    derived[threadIdx.x] = dl;
    stride[threadIdx.x] = st;
  }
  __syncthreads();
  // chunk c: parity c >= CE, coefficients j0 .. j0 + 32 of that parity
  const auto a_load = [&](int c) {
    const int parity = c >= CE, j0 = kChunk * (parity ? c - CE : c), n = parity ? KO : KE;
    return [=](int r, int x) {
      const int j = j0 + x;
      const float* F = src[r];
      if (!F || j >= n) return 0.0f;
      const int k = parity_k(parity, j), R = stride[r];
      if (derived[r]) {
        const float m = (float)(k >> 1);
        return (k & 1) ? m * F[(k - 1) * R] : -m * F[(k + 1) * R];
      }
      return F[k * R];
    };
  };
  const auto b_load = [&](int c) {
    const int parity = c >= CE, j0 = kChunk * (parity ? c - CE : c), n = parity ? KO : KE;
    return [=](int y, int x) { return j0 + y < n ? inv[parity_k(parity, j0 + y) * nlon + x] : 0.0f; };
  };
  f4 even[kInvTiles], odd[kInvTiles];
#pragma unroll
  for (int nt = 0; nt < kInvTiles; nt++) even[nt] = odd[nt] = f4{0, 0, 0, 0};
  float va[kAN], vb[kBN];
  load_rows<kAN>(a_load(0), va);
  load_regs<kBN, 7>(kChunk, half, b_load(0), vb);
  for (int c = 0; c < chunks; c++) {
    store_rows<kAN>([&](int r, int x, float v) { As[r * a_stride(kChunk) + x] = v; }, va);
    store_regs<kBN, 7>(kChunk, half, [&](int y, int x, float v) { Bs[y * bs + x] = v; }, vb);
    __syncthreads();
    if (c + 1 < chunks) {
      load_rows<kAN>(a_load(c + 1), va);
      load_regs<kBN, 7>(kChunk, half, b_load(c + 1), vb);
    }
    mfma_rows<kInvTiles>(As, Bs, bs, NT, c < CE ? even : odd);
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    float* G = dst[16 * wave + 4 * (lane >> 4) + r];
    if (G) {
#pragma unroll
      for (int nt = 0; nt < kInvTiles; nt++) {
        if (nt < NT) {
          const int x = 16 * nt + (lane & 15);
          G[x] = even[nt][r] + odd[nt][r];
          G[x + half] = even[nt][r] - odd[nt][r];
        }
      }
    }
  }
}

extern "C" __global__ __launch_bounds__(kThreads) void k_dft_fwd(const DevMember* __restrict__ mems, GroupPlan plan,
                                                                  FwdJobs jobs) {
  // A: the half-row sums (even m) and differences (odd m); B: the table's
  // coefficients of even m in columns 0..63 and of odd m in 64..127
  __shared__ float As[2 * kAs];
  __shared__ float Bs[kFwdBs];
  __shared__ const float* src[kRows];
  __shared__ float* dst[kRows];
  __shared__ int stride[kRows];  // rows of the row's Fourier slot
  int block;
  const int grp = group_find(plan, blockIdx.x, &block);
  const DevMember& M = mems[plan.list[plan.start[grp]]];
  const int nlon = M.nlon, nfc = M.nfc, ntp1 = M.ntru + 1;
  const auto rows_of = [&](const FwdJob& x) { return fwd_rows(M, x); };
  int per_member = 0;
  for (int j = 0; j < jobs.n; j++) per_member += rows_of(jobs.j[j]);
  const int total = plan.count[grp] * per_member;
  const int row0 = block * kRows;
  const int half = nlon / 2;  // a multiple of 32
  const int KE = parity_count(0, ntp1), KO = parity_count(1, ntp1);
  const int NTE = (KE + 15) / 16, NTO = (KO + 15) / 16;
  constexpr int bs = b_stride(128);
  const float* fwd = M.ltab + M.tabo[TAB_fourier_forward];
  if (threadIdx.x < kRows) {
    const int row = row0 + threadIdx.x;
    const float* s = nullptr;
    float* d = nullptr;
    int st = 0;
    if (row < total) {
      int g, r;
      const FwdJob& job = jobs.j[locate(jobs, per_member, row, rows_of, &g, &r)];
      const DevMember& G = mems[plan.list[plan.start[grp] + g]];
      s = G.arena + G.slot[job.src] + r * nlon;
      d = G.arena + G.slot[job.dst] + r;
      st = rows_of(job);
    }
    src[threadIdx.x] = s;
    dst[threadIdx.x] = d;
    stride[threadIdx.x] = st;
  }
  __syncthreads();
  const auto g_load = [&](int x0, int shift) {
    return [=](int r, int x) {
      const float* g = src[r];
      return g ? g[x0 + shift + x] : 0.0f;
    };
  };
  const auto b_load = [&](int x0) {
    return [=](int y, int j) {
      const int parity = j >> 6, jj = j & 63;
      return jj < (parity ? KO : KE) ? fwd[(x0 + y) * nfc + parity_k(parity, jj)] : 0.0f;
    };
  };
  f4 even[kFwdTiles / 2], odd[kFwdTiles / 2];
#pragma unroll
  for (int nt = 0; nt < kFwdTiles / 2; nt++) even[nt] = odd[nt] = f4{0, 0, 0, 0};
  float va[kAN], vc[kAN], vb[kBN];
  load_regs<kAN, 5>(kRows, kChunk, g_load(0, 0), va);
  load_regs<kAN, 5>(kRows, kChunk, g_load(0, half), vc);
  load_regs<kBN, 7>(kChunk, 128, b_load(0), vb);
  for (int x0 = 0; x0 < half; x0 += kChunk) {
#pragma unroll
    for (int j = 0; j < kAN; j++) {
      const float a = va[j], c = vc[j];
      va[j] = a + c;
      vc[j] = a - c;
    }
    store_regs<kAN, 5>(kRows, kChunk, [&](int r, int x, float v) { As[r * a_stride(kChunk) + x] = v; }, va);
    store_regs<kAN, 5>(kRows, kChunk, [&](int r, int x, float v) { As[kAs + r * a_stride(kChunk) + x] = v; }, vc);
    store_regs<kBN, 7>(kChunk, 128, [&](int y, int j, float v) { Bs[y * bs + j] = v; }, vb);
    __syncthreads();
    if (x0 + kChunk < half) {
      load_regs<kAN, 5>(kRows, kChunk, g_load(x0 + kChunk, 0), va);
      load_regs<kAN, 5>(kRows, kChunk, g_load(x0 + kChunk, half), vc);
      load_regs<kBN, 7>(kChunk, 128, b_load(x0 + kChunk), vb);
    }
    mfma_rows<kFwdTiles / 2>(As, Bs, bs, NTE, even);
    mfma_rows<kFwdTiles / 2>(As + kAs, Bs + 64, bs, NTO, odd);
    __syncthreads();
  }
  // Out through LDS, one parity at a time: the accumulators go to T[j][row]
  // (rows padded to 129 floats), then each wave writes one coefficient of
  // 64 consecutive rows, contiguous in the coefficient-major Fourier slot.
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float* Tt = As;
  static_assert(64 * (kRows + 1) <= 2 * kAs, "transpose buffer");
#pragma unroll
  for (int parity = 0; parity < 2; parity++) {
    const int K = parity ? KO : KE, NT = parity ? NTO : NTE;
#pragma unroll
    for (int nt = 0; nt < kFwdTiles / 2; nt++) {
      if (nt < NT) {
        const int j = 16 * nt + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; r++)
          Tt[j * (kRows + 1) + 16 * wave + 4 * (lane >> 4) + r] = parity ? odd[nt][r] : even[nt][r];
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < K * kRows; e += kThreads) {
      const int j = e / kRows, row = e & (kRows - 1);
      float* F = dst[row];
      if (F) F[parity_k(parity, j) * stride[row]] = Tt[j * (kRows + 1) + row];
    }
    __syncthreads();  // T is rewritten by the next parity
  }
}
