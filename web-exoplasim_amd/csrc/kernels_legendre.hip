// Legendre transforms of the ExoPlaSim step on gfx950 matrix cores.
//
// For one member, zonal wavenumber m and job, a transform is a dense
// product:
//   inverse (sp2fc, sp2fcdmu, dv2uv; legendre/wgsl/inverse.wgsl:6-27,
//   dv2uv.wgsl:9-43):   F[lat][c] = sum_r Q[lat][r] B[r][c]
//   forward (fc2sp, uv2dv, mktend, qtend; forward.wgsl:7-38,
//   uv2dv.wgsl:11-63, qtend.wgsl:9-46):   S[k][c] = sum_r A[k][r] B[r][c]
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
// (F_N +- F_S, sign (-1)^(n-m)); here the forward tables are unfolded to
// every latitude while they are staged (southern row = (-1)^(n-m) times
// the northern value, negated again for the "difference" table qe*), so
// each block is a plain product over all NLAT rows.
//
// Each workgroup (4 waves) stages one block of A and B in LDS, then every
// wave runs v_mfma_f32_16x16x4_f32 over its 16x16 output tiles, keeping
// the accumulators in registers across blocks. LDS strides are chosen so
// the operand reads of one MFMA step are free of bank conflicts.
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

// Row strides: A rows are read 16 at a time with 4 columns per lane group
// (stride = 2 mod 32 spreads them over 32 banks); B rows are read as 16
// consecutive columns of 4 rows (stride = 16 mod 32).
__host__ __device__ constexpr int a_stride(int r) { return (r + 29) / 32 * 32 + 2; }
__host__ __device__ constexpr int b_stride(int c) { return (c + 15) / 32 * 32 + 16; }
// LDS tiles: inverse A is NLAT x K, B is K x 4L; forward A is K x NLAT,
// B is NLAT x 4L, plus the raw Fourier inputs (3 fields x L x NLAT x 2).
constexpr int kInvAs = kMaxLat * a_stride(kMaxK);
constexpr int kInvBs = kMaxK * b_stride(4 * kMaxLev);
constexpr int kFwdAs = kMaxK * a_stride(kMaxLat);
constexpr int kFwdBs = kMaxLat * b_stride(4 * kMaxLev);
constexpr int kFwdRaw = 3 * kMaxLev * 2 * kMaxLat;

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

// Two blocks with the same columns staged together, so the loads of both
// are in flight before either is stored.
template <typename LA, typename SA, typename LB, typename SB>
__device__ __forceinline__ void stage_pair(int rows_a, int rows_b, int cols, LA load_a, SA store_a, LB load_b,
                                           SB store_b) {
  const int r0 = threadIdx.x >> 6;
  const int rows = rows_a > rows_b ? rows_a : rows_b;
  for (int c = threadIdx.x & 63; c < cols; c += 64) {
    for (int base = r0; base < rows; base += 32) {
      float va[8], vb[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + 4 * j;
        va[j] = r < rows_a ? load_a(r, c) : 0.0f;
        vb[j] = r < rows_b ? load_b(r, c) : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + 4 * j;
        if (r < rows_a) store_a(r, c, va[j]);
        if (r < rows_b) store_b(r, c, vb[j]);
      }
    }
  }
}

// acc[t] += A[rows of tile][0..R) * B[0..R)[cols of tile] for the tiles
// this wave owns (tile = wave + 4 t of MT x NT).
__device__ __forceinline__ void mfma_block(const float* As, int as, const float* Bs, int bs, int R, int MT,
                                           int NT, f4* acc) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) {
    const int tile = wave + 4 * t;
    if (tile < MT * NT) {
      const int mt = tile / NT, nt = tile % NT;
      const float* a = As + (mt * 16 + i) * as + kk;
      const float* b = Bs + kk * bs + nt * 16 + i;
      f4 c = acc[t];
      for (int r = 0; r < R; r += 4) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], b[r * bs], c, 0, 0, 0);
      acc[t] = c;
    }
  }
}

// ------------------------------------------------------------- inverse

template <int P>
__device__ void leg_inv(const DevMember& M, const InvJob& job, int m, float* As, float* Bs) {
  const int nlat = M.nlat, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc;
  const bool winds = job.kind == INV_DV2UV;
  const int NC = winds ? 4 * P : 2 * P;
  const int NT = (NC + 15) / 16, MT = nlat / 16;
  const int R = (K + 3) & ~3;
  const int as = a_stride(R), bs = b_stride(NC);
  const float* S0 = M.arena + M.slot[job.src0];
  const float* S1 = winds ? M.arena + M.slot[job.src1] : nullptr;
  const float plavor = job.plavor ? M.k[CON_plavor] : 0.0f;
  f4 acc[kMaxTiles];
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) acc[t] = f4{0, 0, 0, 0};
  const int blocks = winds ? 2 : 1;
  for (int blk = 0; blk < blocks; blk++) {
    const float* q = M.ltab + M.tabo[winds ? (blk == 0 ? (int)TAB_legendre_qu : (int)TAB_legendre_qv) : job.table];
    // A[lat][k] and B[k][c], column c = (level, part) of the block's
    // arrangement; lanes walk the modes k, so a table row and a plane's
    // coefficients each load as one run
    stage_pair(
        nlat, NT * 16, R, [&](int lat, int k) { return k < K ? q[lat * ncsp + w0 + k] : 0.0f; },
        [&](int lat, int k, float v) { As[lat * as + k] = v; },
        [&](int c, int k) {
          float v = 0.0f;
          if (k < K && c < NC) {
            const int w = w0 + k;
            if (!winds) {
              v = S0[(c >> 1) * nrsp + 2 * w + (c & 1)];
            } else {
              const int p = c >> 2, part = c & 3;
              const float* d = S0 + p * nrsp + 2 * w;
              const float* z = S1 + p * nrsp + 2 * w;
              // zeta without the planetary vorticity in Re mode (0, 1)
              const float zre = z[0] - (w == 1 ? plavor : 0.0f), zim = z[1];
              if (blk == 0)
                v = part == 0 ? d[1] : part == 1 ? -d[0] : part == 2 ? zim : -zre;
              else
                v = part == 0 ? zre : part == 1 ? zim : part == 2 ? -d[0] : -d[1];
            }
          }
          return v;
        },
        [&](int c, int k, float v) { Bs[k * bs + c] = v; });
    __syncthreads();
    mfma_block(As, as, Bs, bs, R, MT, NT, acc);
    __syncthreads();
  }
  float* F0 = M.arena + M.slot[job.four0];
  float* F1 = winds ? M.arena + M.slot[job.four1] : nullptr;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) {
    const int tile = wave + 4 * t;
    if (tile < MT * NT) {
      const int mt = tile / NT, nt = tile % NT;
      const int c = nt * 16 + (lane & 15);
      if (c < NC) {
        float* base;
        int p, part;
        if (!winds) {
          base = F0;
          p = c >> 1;
          part = c & 1;
        } else {
          base = (c & 2) ? F1 : F0;
          p = c >> 2;
          part = c & 1;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int lat = mt * 16 + 4 * (lane >> 4) + r;
          base[(p * nlat + lat) * nfc + 2 * m + part] = acc[t][r];
        }
      }
    }
  }
}

// ------------------------------------------------------------- forward

template <int P>
__device__ void leg_fwd(const DevMember& M, const LegJob& job, int m, float* As, float* Bs, float* raw) {
  const int nlat = M.nlat, half = nlat / 2, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp;
  const bool uv = job.kind == LEG_UV2DV;
  const int NC = uv ? 4 * P : 2 * P;
  const int NT = (NC + 15) / 16, MT = (K + 15) / 16;
  const int R = nlat;
  const int as = a_stride(R), bs = b_stride(NC);
  f4 acc[kMaxTiles];
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) acc[t] = f4{0, 0, 0, 0};
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
  // the job's Fourier coefficients m of every row and level, once:
  // raw[(f * P + p) * nlat + lat][part] for fields f = (u, v, third)
  const int fields[3] = {job.kind == LEG_PLAIN ? job.src : job.u, job.v, uv ? job.ke : job.src};
  const int nf = job.kind == LEG_PLAIN ? 1 : (uv && job.ke < 0) ? 2 : 3;
  stage(
      nf * P, 2 * nlat,
      [&](int fp, int x) {
        const int f = fp / P, p = fp % P;
        return M.arena[M.slot[fields[f]] + (p * nlat + (x >> 1)) * M.nfc + 2 * m + (x & 1)];
      },
      [&](int fp, int x, float v) { raw[fp * 2 * nlat + x] = v; });
  auto F = [&](int f, int p, int lat, int part) { return raw[(f * P + p) * 2 * nlat + 2 * lat + part]; };
  for (int blk = 0; blk < blocks; blk++) {
    const float* q = M.ltab + M.tabo[tables[blk]];
    // A[k][lat]: the folded table's row l and its mirror nlat - 1 - l
    const int dsign = diff[blk];
    stage(
        half, MT * 16, [&](int l, int k) { return k < K ? q[l * ncsp + w0 + k] : 0.0f; },
        [&](int l, int k, float north) {
          const float s = (k & 1) ? -1.0f : 1.0f;
          As[k * as + l] = north;
          As[k * as + nlat - 1 - l] = dsign ? -s * north : s * north;
        });
    __syncthreads();
    stage(
        R, NT * 16,
        [&](int lat, int c) {
          float v = 0.0f;
          if (c < NC) {
            if (job.kind == LEG_PLAIN) {
              v = F(0, c >> 1, lat, c & 1);
            } else if (uv) {
              const int p = c >> 2, part = c & 3;
              if (blk == 0)
                v = part == 0 ? -F(1, p, lat, 0) : part == 1 ? -F(1, p, lat, 1) : part == 2 ? F(0, p, lat, 0)
                                                                                       : F(0, p, lat, 1);
              else if (blk == 1)
                v = part == 0 ? -F(0, p, lat, 1) : part == 1 ? F(0, p, lat, 0) : part == 2 ? -F(1, p, lat, 1)
                                                                                      : F(1, p, lat, 0);
              else
                v = part < 2 ? F(2, p, lat, part) : 0.0f;
            } else {
              const int p = c >> 1, part = c & 1;
              if (blk == 0)
                v = F(1, p, lat, part);
              else if (blk == 1)
                v = F(2, p, lat, part);
              else
                v = part == 0 ? F(0, p, lat, 1) : -F(0, p, lat, 0);
            }
          }
          return v;
        },
        [&](int lat, int c, float v) { Bs[lat * bs + c] = v; });
    __syncthreads();
    mfma_block(As, as, Bs, bs, R, MT, NT, acc);
    __syncthreads();
  }
  float* O0 = M.arena + M.slot[job.dst0];
  float* O1 = uv ? M.arena + M.slot[job.dst1] : nullptr;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < kMaxTiles; t++) {
    const int tile = wave + 4 * t;
    if (tile < MT * NT) {
      const int mt = tile / NT, nt = tile % NT;
      const int c = nt * 16 + (lane & 15);
      if (c < NC) {
        float* base;
        int p, part;
        if (!uv) {
          base = O0;
          p = c >> 1;
          part = c & 1;
        } else {
          base = (c & 2) ? O1 : O0;
          p = c >> 2;
          part = c & 1;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int k = mt * 16 + 4 * (lane >> 4) + r;
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
  __shared__ float Bs[kInvBs];
  int tile;
  const DevMember& M = mems[plan_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const InvJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) leg_inv<PP>(M, job, m, As, Bs)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(256) void k_leg_fwd(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                             LegJobs jobs) {
  __shared__ float As[kFwdAs];
  __shared__ float Bs[kFwdBs];
  __shared__ float raw[kFwdRaw];
  int tile;
  const DevMember& M = mems[plan_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const LegJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) leg_fwd<PP>(M, job, m, As, Bs, raw)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}
