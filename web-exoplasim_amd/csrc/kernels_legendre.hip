// Legendre transforms of the ExoPlaSim step on gfx950 matrix cores.
//
// For one member, zonal wavenumber m and job, a transform is a dense
// product:
//   inverse (sp2fc, sp2fcdmu, dv2uv; legendre/wgsl/inverse.wgsl:6-27,
//   dv2uv.wgsl:9-43):   F[lat][c] = sum_k Q[lat][k] B[k][c]
//   forward (fc2sp, uv2dv, mktend, qtend; forward.wgsl:7-38,
//   uv2dv.wgsl:11-63, qtend.wgsl:9-46):   S[k][c] = sum_lat A[k][lat] B[lat][c]
// with c running over (level, output part). The reduction runs over
// "blocks", each one table against one arrangement of the inputs:
//   inverse plain:  Q = q (table of the job),       B = (Re S, Im S)
//   inverse dv2uv:  Q = qu, B = (Im D, -Re D, Im Z, -Re Z)
//                   Q = qv, B = (Re Z, Im Z, -Re D, -Im D)      -> (U, V)
//   forward plain:  A = qc,  B = (Re F, Im F)
//   forward uv2dv:  A = qe*, B = (-Re v, -Im v, Re u, Im u)
//                   A = qm,  B = (-Im u, Re u, -Im v, Re v)
//                   A = qq,  B = (Re ke, Im ke, 0, 0)  (with kinetic energy)
//                   -> (D, zeta)
//   forward qtend:  A = qe*, B = (Re v, Im v); A = qc, B = (Re s, Im s);
//                   A = qm,  B = (Im u, -Re u)
// The reference folds the forward sums over northern/southern row pairs
// (F_N +- F_S, sign (-1)^(n-m)); here the product runs over every
// latitude, with the southern row of A read as (-1)^(n-m) times the
// northern table value (negated again for the "difference" table qe*).
//
// A workgroup (4 waves) stages the job's inputs of wavenumber m once
// (spectral coefficients, or the Fourier coefficients of every row) and,
// per block, the table columns of m; B is never built: each lane reads
// its column's operand straight from the staged inputs with the column's
// sign. Every wave runs v_mfma_f32_16x16x4_f32 over its 16x16 output
// tiles, keeping the accumulators in registers across blocks.
#include "transforms.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kMaxLat = 96, kMaxK = 64, kMaxLev = 10;
// output tiles per wave: inverse dv2uv at T63 L10 is 6 x 3 tiles
constexpr int kMaxTiles = 5;

__device__ __forceinline__ int mode_first(int m, int ntp1) { return m * (2 * ntp1 - m + 1) / 2; }

template <typename Jobs>
__device__ __forceinline__ int job_of(const Jobs& jobs, int index) {
  return index < jobs.n ? index : jobs.n - 1;
}

// A rows are read 16 at a time with 4 columns per lane group (stride =
// 2 mod 32 spreads them over 32 banks).
__host__ __device__ constexpr int a_stride(int r) { return (r + 29) / 32 * 32 + 2; }

// Staged inputs: inverse rawS[(f * kMaxLev + p) * kRawK + 2k + part]
// (f = source field, k < K padded with zeros to a multiple of 4); forward
// raw[(f * kMaxLev + p) * kRawLat + 2 lat + part].
constexpr int kRawK = 2 * kMaxK;
constexpr int kRawLat = 2 * kMaxLat;
constexpr int kInvAs = kMaxLat * a_stride(kMaxK);
constexpr int kInvRaw = 2 * kMaxLev * kRawK;
constexpr int kFwdAs = kMaxK * a_stride(kMaxLat / 2);
constexpr int kFwdRaw = 3 * kMaxLev * kRawLat;

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD
// b mod 8). Within each run of 64 blocks, XCD x is given the 8 consecutive
// tiles 8x..8x+7 instead, so the tiles of neighbouring wavenumbers, which
// share cache lines of the Fourier rows, meet in one L2 while the heavy
// and light tiles stay spread over all XCDs. A last partial run keeps the
// plain order.
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int run = b >> 6;
  if ((run + 1) * 64 > n) return b;
  return (run << 6) + (b & 7) * 8 + ((b >> 3) & 7);
}

// LDS staging of a rows x cols block: lane = column (in runs of 64), the 4
// waves take rows r0, r0 + 4, ...; each lane issues 8 loads before it
// stores any, so one memory latency covers 8 rows.
template <typename Load, typename Store>
__device__ __forceinline__ void stage(int rows, int cols, Load load, Store store) {
  const int r0 = threadIdx.x >> 6;
  for (int c = threadIdx.x & 63; c < cols; c += 64) {
    for (int base = r0; base < rows; base += 32) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + 4 * j;
        v[j] = r < rows ? load(r, c) : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + 4 * j;
        if (r < rows) store(r, c, v[j]);
      }
    }
  }
}

// Where a product column's operand sits in the staged inputs (its offset
// for row 0; rows step by 2 floats) and its sign; sign 0 is a zero column.
struct Col {
  int base;
  float sign;
};

__device__ __forceinline__ Col col_of(int field, int p, int part, float sign, int row_floats) {
  return Col{(field * kMaxLev + p) * row_floats + part, sign};
}

// inverse: plain (S); dv2uv fields 0 = D, 1 = zeta, block 0 = qu, 1 = qv
__device__ __forceinline__ Col inv_col(bool winds, int blk, int c, int NC) {
  if (c >= NC) return Col{0, 0.0f};
  if (!winds) return col_of(0, c >> 1, c & 1, 1.0f, kRawK);
  const int p = c >> 2, q = c & 3;
  if (blk == 0) return col_of(q >> 1, p, (q & 1) ? 0 : 1, (q & 1) ? -1.0f : 1.0f, kRawK);
  return col_of((q >> 1) ? 0 : 1, p, q & 1, (q >> 1) ? -1.0f : 1.0f, kRawK);
}

// forward: fields 0 = u (plain: the input), 1 = v, 2 = ke (qtend: s)
__device__ __forceinline__ Col fwd_col(int kind, int blk, int c, int NC) {
  if (c >= NC) return Col{0, 0.0f};
  if (kind == LEG_PLAIN) return col_of(0, c >> 1, c & 1, 1.0f, kRawLat);
  if (kind == LEG_UV2DV) {
    const int p = c >> 2, q = c & 3;
    if (blk == 0) return col_of((q >> 1) ? 0 : 1, p, q & 1, (q >> 1) ? 1.0f : -1.0f, kRawLat);
    if (blk == 1) return col_of(q >> 1, p, (q & 1) ? 0 : 1, (q & 1) ? 1.0f : -1.0f, kRawLat);
    return q < 2 ? col_of(2, p, q, 1.0f, kRawLat) : Col{0, 0.0f};
  }
  const int p = c >> 1, q = c & 1;
  if (blk == 0) return col_of(1, p, q, 1.0f, kRawLat);
  if (blk == 1) return col_of(2, p, q, 1.0f, kRawLat);
  return col_of(0, p, q ? 0 : 1, q ? -1.0f : 1.0f, kRawLat);
}

// ------------------------------------------------------------- inverse

template <int P>
__device__ void leg_inv(const DevMember& M, const InvJob& job, int m, float* As, float* raw) {
  const int nlat = M.nlat, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc;
  const bool winds = job.kind == INV_DV2UV;
  const int NC = winds ? 4 * P : 2 * P;
  const int NT = (NC + 15) / 16, MT = nlat / 16;
  const int R = (K + 3) & ~3;
  const int as = a_stride(R);
  const int nf = winds ? 2 : 1;
  const float plavor = job.plavor ? M.k[CON_plavor] : 0.0f;
  // the job's coefficients of m (zeta without the planetary vorticity in
  // Re mode (0, 1)), zeros past K
  stage(
      nf * P, 2 * R,
      [&](int fp, int x) {
        const int f = fp / P, p = fp % P, k = x >> 1;
        if (k >= K) return 0.0f;
        const float v = M.arena[M.slot[f ? job.src1 : job.src0] + p * nrsp + 2 * w0 + x];
        return (winds && f == 1 && w0 + k == 1 && (x & 1) == 0) ? v - plavor : v;
      },
      [&](int fp, int x, float v) { raw[((fp / P) * kMaxLev + fp % P) * kRawK + x] = v; });
  f4 acc[kMaxTiles];
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) acc[t] = f4{0, 0, 0, 0};
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  const int blocks = winds ? 2 : 1;
  for (int blk = 0; blk < blocks; blk++) {
    const float* q = M.ltab + M.tabo[winds ? (blk == 0 ? (int)TAB_legendre_qu : (int)TAB_legendre_qv) : job.table];
    stage(
        nlat, R, [&](int lat, int k) { return k < K ? q[lat * ncsp + w0 + k] : 0.0f; },
        [&](int lat, int k, float v) { As[lat * as + k] = v; });
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) {
      const int tile = wave + 4 * t;
      if (tile < MT * NT) {
        const int mt = tile / NT, nt = tile % NT;
        const Col col = inv_col(winds, blk, nt * 16 + i, NC);
        const float* a = As + (mt * 16 + i) * as + kk;
        const float* b = raw + col.base + 2 * kk;
        f4 c = acc[t];
        for (int r = 0; r < R; r += 4)
          c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], col.sign * b[2 * r], c, 0, 0, 0);
        acc[t] = c;
      }
    }
    __syncthreads();
  }
  float* F0 = M.arena + M.slot[job.four0];
  float* F1 = winds ? M.arena + M.slot[job.four1] : nullptr;
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) {
    const int tile = wave + 4 * t;
    if (tile < MT * NT) {
      const int mt = tile / NT, nt = tile % NT;
      const int c = nt * 16 + i;
      if (c < NC) {
        float* base = winds ? ((c & 2) ? F1 : F0) : F0;
        const int p = winds ? c >> 2 : c >> 1, part = c & 1;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int lat = mt * 16 + 4 * kk + r;
          base[(p * nlat + lat) * nfc + 2 * m + part] = acc[t][r];
        }
      }
    }
  }
}

// ------------------------------------------------------------- forward

template <int P>
__device__ void leg_fwd(const DevMember& M, const LegJob& job, int m, float* Ah, float* raw) {
  const int nlat = M.nlat, half = nlat / 2, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc;
  const bool uv = job.kind == LEG_UV2DV;
  const int NC = uv ? 4 * P : 2 * P;
  const int NT = (NC + 15) / 16, MT = (K + 15) / 16;
  const int ash = a_stride(half);
  // blocks: table and whether its southern rows take the extra minus sign
  int tables[3], diff[3], blocks = 0;
  if (job.kind == LEG_PLAIN) {
    tables[blocks] = TAB_legendre_qc, diff[blocks++] = 0;
  } else if (uv) {
    tables[blocks] = TAB_legendre_qe, diff[blocks++] = 1;
    tables[blocks] = TAB_legendre_qm, diff[blocks++] = 0;
    if (job.ke >= 0) tables[blocks] = TAB_legendre_qq, diff[blocks++] = 0;
  } else {
    tables[blocks] = TAB_legendre_qe, diff[blocks++] = 1;
    tables[blocks] = TAB_legendre_qc, diff[blocks++] = 0;
    tables[blocks] = TAB_legendre_qm, diff[blocks++] = 0;
  }
  // the job's Fourier coefficients m of every row and level, once
  const int fields[3] = {job.kind == LEG_PLAIN ? job.src : job.u, job.v, uv ? job.ke : job.src};
  const int nf = job.kind == LEG_PLAIN ? 1 : (uv && job.ke < 0) ? 2 : 3;
  stage(
      nf * P, 2 * nlat,
      [&](int fp, int x) {
        const int f = fp / P, p = fp % P;
        return M.arena[M.slot[fields[f]] + (p * nlat + (x >> 1)) * nfc + 2 * m + (x & 1)];
      },
      [&](int fp, int x, float v) { raw[((fp / P) * kMaxLev + fp % P) * kRawLat + x] = v; });
  f4 acc[kMaxTiles];
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) acc[t] = f4{0, 0, 0, 0};
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  for (int blk = 0; blk < blocks; blk++) {
    const float* q = M.ltab + M.tabo[tables[blk]];
    // the northern half of A: Ah[k][l] = q(l, w0 + k)
    stage(
        half, MT * 16, [&](int l, int k) { return k < K ? q[l * ncsp + w0 + k] : 0.0f; },
        [&](int l, int k, float v) { Ah[k * ash + l] = v; });
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) {
      const int tile = wave + 4 * t;
      if (tile < MT * NT) {
        const int mt = tile / NT, nt = tile % NT;
        const int k = mt * 16 + i;
        // southern rows: (-1)^(n-m) q, negated again for qe*
        const float south = ((k & 1) ? -1.0f : 1.0f) * (diff[blk] ? -1.0f : 1.0f);
        const Col col = fwd_col(job.kind, blk, nt * 16 + i, NC);
        const float* a = Ah + k * ash;
        const float* b = raw + col.base + 2 * kk;
        f4 c = acc[t];
        for (int r = 0; r < half; r += 4)
          c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r + kk], col.sign * b[2 * r], c, 0, 0, 0);
        for (int r = half; r < nlat; r += 4)
          c = __builtin_amdgcn_mfma_f32_16x16x4f32(south * a[nlat - 1 - r - kk], col.sign * b[2 * r], c, 0, 0, 0);
        acc[t] = c;
      }
    }
    __syncthreads();
  }
  float* O0 = M.arena + M.slot[job.dst0];
  float* O1 = uv ? M.arena + M.slot[job.dst1] : nullptr;
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) {
    const int tile = wave + 4 * t;
    if (tile < MT * NT) {
      const int mt = tile / NT, nt = tile % NT;
      const int c = nt * 16 + i;
      if (c < NC) {
        float* base = uv ? ((c & 2) ? O1 : O0) : O0;
        const int p = uv ? c >> 2 : c >> 1, part = c & 1;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int k = mt * 16 + 4 * kk + r;
          if (k < K) base[p * nrsp + 2 * (w0 + k) + part] = acc[t][r];
        }
      }
    }
  }
}

}  // namespace

#define EXO_PLANES_DISPATCH(P, call) \
  do {                               \
    if ((P) == 10)                   \
      call(10);                      \
    else if ((P) == 5)               \
      call(5);                       \
    else                             \
      call(1);                       \
  } while (0)

extern "C" __global__ __launch_bounds__(256) void k_leg_inv(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                             InvJobs jobs) {
  __shared__ float As[kInvAs];
  __shared__ float raw[kInvRaw];
  int tile;
  const DevMember& M = mems[plan_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const InvJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) leg_inv<PP>(M, job, m, As, raw)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(256) void k_leg_fwd(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                             LegJobs jobs) {
  __shared__ float Ah[kFwdAs];
  __shared__ float raw[kFwdRaw];
  int tile;
  const DevMember& M = mems[plan_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const LegJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) leg_fwd<PP>(M, job, m, Ah, raw)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}
