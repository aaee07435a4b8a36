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
typedef float f2 __attribute__((ext_vector_type(2)));

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

// Staged inputs: inverse raw[row * kPairRawK + 2k + part] (below); forward
// raw[row * kRawLat + 2 lat + part], rows over (member, field, level).
// Row strides are 4 mod 32: the 16 product columns one MFMA step reads
// (8 rows x 2 parts) and its 2 row offsets then fall in 32 distinct banks.
//
// Fourier coefficients live coefficient-major: coefficient k = 2m + part of
// row (plane p, latitude lat) of a slot with R = planes x NLAT rows is
// element k R + p NLAT + lat, so one wavenumber's coefficients of every
// row and latitude are one contiguous block, which the per-wavenumber
// workgroups read and write whole.
constexpr int kRawLat = 2 * kMaxLat + 4;
constexpr int kFwdAs = kMaxK * a_stride(kMaxLat / 2);

// LDS staging of a rows x cols block: lane = column (in runs of 64), the W
// waves take rows r0, r0 + W, ...; each lane issues 8 loads before it
// stores any, so one memory latency covers 8 rows.
template <int W = 4, typename Load, typename Store>
__device__ __forceinline__ void stage(int rows, int cols, Load load, Store store) {
  const int r0 = threadIdx.x >> 6;
  for (int c = threadIdx.x & 63; c < cols; c += 64) {
    for (int base = r0; base < rows; base += 8 * W) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + W * j;
        v[j] = r < rows ? load(r, c) : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int r = base + W * j;
        if (r < rows) store(r, c, v[j]);
      }
    }
  }
}

// Where a product column's operand sits in the staged inputs (its offset
// for row 0; rows step by 2 floats) and its sign; sign 0 is a zero column.
// Staged input rows run over (member of the chunk, field, level):
// row = (g nf + f) P + p.
struct Col {
  int base;
  float sign;
};

__device__ __forceinline__ Col gcol(int g, int nf, int P, int field, int p, int part, float sign, int row_floats) {
  return Col{((g * nf + field) * P + p) * row_floats + part, sign};
}

// inverse: plain (S); dv2uv fields 0 = D, 1 = zeta, block 0 = qu, 1 = qv.
// Column c of the chunk belongs to member c / NCm; staged rows hold
// row_floats floats.
__device__ __forceinline__ Col inv_col(bool winds, int blk, int c, int NC, int NCm, int P, int row_floats) {
  if (c >= NC) return Col{0, 0.0f};
  const int g = c / NCm;
  c -= g * NCm;
  if (!winds) return gcol(g, 1, P, 0, c >> 1, c & 1, 1.0f, row_floats);
  const int p = c >> 2, q = c & 3;
  if (blk == 0) return gcol(g, 2, P, q >> 1, p, (q & 1) ? 0 : 1, (q & 1) ? -1.0f : 1.0f, row_floats);
  return gcol(g, 2, P, (q >> 1) ? 0 : 1, p, q & 1, (q >> 1) ? -1.0f : 1.0f, row_floats);
}

// forward: fields 0 = u (plain: the input), 1 = v, 2 = ke (qtend: s)
__device__ __forceinline__ Col fwd_col(int kind, int blk, int c, int NC, int NCm, int nf, int P) {
  if (c >= NC) return Col{0, 0.0f};
  const int g = c / NCm;
  c -= g * NCm;
  if (kind == LEG_PLAIN) return gcol(g, nf, P, 0, c >> 1, c & 1, 1.0f, kRawLat);
  if (kind == LEG_UV2DV) {
    const int p = c >> 2, q = c & 3;
    if (blk == 0) return gcol(g, nf, P, (q >> 1) ? 0 : 1, p, q & 1, (q >> 1) ? 1.0f : -1.0f, kRawLat);
    if (blk == 1) return gcol(g, nf, P, q >> 1, p, (q & 1) ? 0 : 1, (q & 1) ? 1.0f : -1.0f, kRawLat);
    return q < 2 ? gcol(g, nf, P, 2, p, q, 1.0f, kRawLat) : Col{0, 0.0f};
  }
  const int p = c >> 1, q = c & 1;
  if (blk == 0) return gcol(g, nf, P, 1, p, q, 1.0f, kRawLat);
  if (blk == 1) return gcol(g, nf, P, 2, p, q, 1.0f, kRawLat);
  return gcol(g, nf, P, 0, p, q ? 0 : 1, q ? -1.0f : 1.0f, kRawLat);
}

// ------------------------------------------------ member groups (FP32)
// Members of one truncation and level count share the transform tables, so
// one workgroup (8 waves) serves a group of them: it stages the table
// columns of m once and runs the product for the members in chunks whose
// inputs fit the staging buffer, the members' product columns side by
// side. Every output element is the same sum, in the same order, whatever
// the group, so a member computes the same alone and in a batch.

constexpr int kGroupThreads = 512;
constexpr int kGroupWaves = kGroupThreads / 64;
constexpr int kGTiles = 4;     // inverse output tiles per wave and chunk (<= 32 per chunk)
// forward: at most 4 x 4 tiles per chunk (K <= 64 rows; the 30 staged input
// rows give at most 60 product columns)
constexpr int kFwdGTiles = 2;
constexpr int kInvRows = 40;   // staged input rows: inverse
constexpr int kFwdRows = 30;   // staged input rows: forward
// input rows of a whole group: members x fields x levels
constexpr int kLegGroupRows = 256;
static_assert(kLegGroup * 3 * kMaxLev <= kLegGroupRows, "group input rows");

__device__ __forceinline__ int group_member(const GroupPlan& plan, int g, int i) {
  return plan.list[plan.start[g] + i];
}

// Member i of group g: its arena (global address space), from the plan.
__device__ __forceinline__ exo_gfloat* group_arena(const GroupPlan& plan, int g, int i) {
  return (exo_gfloat*)plan.arena[plan.start[g] + i];
}

// The planetary vorticity a zeta row of member i drops from Re mode (0, 1).
__device__ __forceinline__ float group_plavor(const DevMember* mems, const GroupPlan& plan, int g, int i) {
  return mems[group_member(plan, g, i)].k[CON_plavor];
}

// ------------------------------------------------------------- inverse
// A workgroup takes the wavenumber pair (m, T - m) of one job and member
// group: K(m) + K(T - m) = T + 2 modes whatever m, so every workgroup does
// the same work (m = T / 2 alone when T is even), and one staging round
// serves both. The pair's table columns and inputs sit side by side, each
// wavenumber's padded to a multiple of 4 with zeros: columns [0, Ra) for
// m, [Ra, Ra + Rb) for T - m.

constexpr int kPairK = 72;                 // Ra + Rb <= 72 at T63
constexpr int kPairRawK = 2 * kPairK + 4;  // input row: 4 x odd floats
constexpr int kInvGroupRows = kLegGroup * 2 * kMaxLev;  // members x fields x levels
// row strides of 2 mod 4 floats: the 16 rows x 4 columns one MFMA step
// reads fall in distinct banks
__host__ __device__ constexpr int pair_stride(int r) { return r + 2; }
constexpr int kPairAs = kMaxLat * pair_stride(kPairK);
// Staging of a rows x C block with a fixed column per thread: thread t
// takes column t % C and rows t / C + S j (S = 512 / C rows per pass), so
// its addresses step by a constant; all N loads are issued before any
// store. Table 96 x 72: 7 rows per pass, 14 each; inputs 40 x 144: 3 rows
// per pass, 14 each.
constexpr int kPairTabS = kGroupThreads / kPairK;
constexpr int kPairTabN = (kMaxLat + kPairTabS - 1) / kPairTabS;
constexpr int kPairRawS = kGroupThreads / (2 * kPairK);
constexpr int kPairRawN = (kInvRows + kPairRawS - 1) / kPairRawS;

// Every lane loads, unconditionally: rows past the end (and the lanes past
// S x C) load the last row, which store_cols then skips, so the loads
// issue back to back (no branch and wait each) and nothing consumes them
// before the stores. `load` must accept every row below `rows` (>= 1) and
// every column below C.
template <int N, int C, typename Load>
__device__ __forceinline__ void load_cols(int rows, Load load, float* v) {
  constexpr int S = kGroupThreads / C;
  const int c = threadIdx.x % C, r0 = threadIdx.x / C;
#pragma unroll
  for (int j = 0; j < N; j++) v[j] = load(min(r0 + S * j, rows - 1), c);
}

template <int N, int C, typename Store>
__device__ __forceinline__ void store_cols(int rows, Store store, const float* v) {
  constexpr int S = kGroupThreads / C;
  const int c = threadIdx.x % C, r0 = threadIdx.x / C;
#pragma unroll
  for (int j = 0; j < N; j++) {
    const int r = r0 + S * j;
    if (r0 < S && r < rows) store(r, c, v[j]);
  }
}

template <int P>
__device__ void leg_inv(const DevMember* mems, const GroupPlan& plan, int grp, const InvJob& job, int pair,
                        float* As, float* raw, exo_gcfloat** rowp, float* plav) {
  const DevMember& M = mems[group_member(plan, grp, 0)];
  const int count = plan.count[grp];
  const int nlat = M.nlat, ntp1 = M.ntru + 1;
  const int ncsp = M.ncsp, nrsp = M.nrsp;
  // the pair: m = pair (K = ntp1 - m modes) and T - m (K = m + 1)
  const int ma = pair, mb = ntp1 - 1 - pair, two = mb != ma;
  const int Ka = ntp1 - ma, Kb = two ? ntp1 - mb : 0;
  const int wa = mode_first(ma, ntp1), wb = two ? mode_first(mb, ntp1) : 0;
  const int Ra = (Ka + 3) & ~3, Rb = (Kb + 3) & ~3, RT = Ra + Rb;
  const int as = pair_stride(RT);
  const bool winds = job.kind == INV_DV2UV;
  const int nf = winds ? 2 : 1;
  const int NCm = winds ? 4 * P : 2 * P;  // product columns of one member
  const int chunk = max(1, min(count, kInvRows / (nf * P)));
  const int MT = nlat / 16;
  const int blocks = winds ? 2 : 1;
  const float* q0 = dev_ltab(M) + M.tabo[winds ? (int)TAB_legendre_qu : job.table];
  const float* q1 = dev_ltab(M) + M.tabo[TAB_legendre_qv];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  // staged column col < RT: mode w of the pair's wavenumber, or padding
  const auto mode_of = [=](int col) {
    if (col < Ra) return col < Ka ? wa + col : -1;
    const int k = col - Ra;
    return k < Kb ? wb + k : -1;
  };
  // the table columns of the pair (dv2uv: qu, qv) need nothing from the
  // members: their loads go first and stay in flight while the row
  // pointers are set up and published
  float va[2][kPairTabN];
#pragma unroll
  for (int blk = 0; blk < 2; blk++)
    if (blk < blocks) {
      const float* q = blk ? q1 : q0;
      load_cols<kPairTabN, kPairK>(nlat, [&](int lat, int col) {
        const int w = col < RT ? mode_of(col) : -1;
        return q[lat * ncsp + max(w, 0)];
      }, va[blk]);
    }
  // where each input row of the group starts (member g, field f, level p
  // at row (g nf + f) P + p)
  if (threadIdx.x < count * nf * P) {
    const int row = threadIdx.x, g = row / (nf * P), fp = row - g * nf * P, f = fp / P, p = fp - f * P;
    rowp[row] = group_arena(plan, grp, g) + M.slot[f ? job.src1 : job.src0] + p * nrsp;
  }
  // each member's planetary vorticity, for its zeta rows' Re mode (0, 1)
  if (job.plavor && threadIdx.x < count) plav[threadIdx.x] = group_plavor(mems, plan, grp, threadIdx.x);
  // the members' coefficients for the chunk from member c0 (the padding
  // loads mode 0; the store zeroes it)
  const auto raw_loader = [&](int c0) {
    const int r0 = c0 * nf * P;
    return [=](int row, int x) { return rowp[r0 + row][2 * max(mode_of(x >> 1), 0) + (x & 1)]; };
  };
  // staged as loaded, zeros in the padding, except Re mode (0, 1) (x = 2
  // of pair 0) of a zeta row (field 1), which loses the member's planetary
  // vorticity
  const bool plavor = job.plavor && ma == 0;
  const auto raw_storer = [&](int c0) {
    const int r0 = c0 * nf * P;
    return [=](int row, int x, float v) {
      if (mode_of(x >> 1) < 0) v = 0.0f;
      if (plavor && x == 2) {
        const int r = r0 + row, g = r / (2 * P);
        if (r - g * 2 * P >= P) v -= plav[g];
      }
      raw[row * kPairRawK + x] = v;
    };
  };
  // the first chunk's inputs follow; every load is issued before any store
  lds_barrier();  // rowp
  {
    float vr[kPairRawN];
    load_cols<kPairRawN, 2 * kPairK>(min(chunk, count) * nf * P, raw_loader(0), vr);
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      if (blk < blocks) {
        float* A = As + blk * kPairAs;
        store_cols<kPairTabN, kPairK>(nlat, [&](int lat, int col, float v) {
          if (col < RT) A[lat * as + col] = mode_of(col) >= 0 ? v : 0.0f;
        },
                                      va[blk]);
      }
    store_cols<kPairRawN, 2 * kPairK>(min(chunk, count) * nf * P, raw_storer(0), vr);
  }
  lds_barrier();
  for (int c0 = 0; c0 < count; c0 += chunk) {
    const int n = min(chunk, count - c0);
    const int NC = n * NCm, NT = (NC + 15) / 16;
    // the next chunk's inputs are in flight while this chunk's products run
    const int c1 = c0 + chunk, n1 = c1 < count ? min(chunk, count - c1) : 0;
    float vn[kPairRawN];
    if (n1 > 0) load_cols<kPairRawN, 2 * kPairK>(n1 * nf * P, raw_loader(c1), vn);
    for (int h = 0; h <= two; h++) {
      const int m = h ? mb : ma, col0 = h ? Ra : 0, R = h ? Rb : Ra;
      f4 acc[kGTiles];
#pragma unroll
      for (int t = 0; t < kGTiles; t++) acc[t] = f4{0, 0, 0, 0};
      for (int blk = 0; blk < blocks; blk++) {
        const float* A = As + blk * kPairAs;
#pragma unroll
        for (int t = 0; t < kGTiles; t++) {
          const int tile = wave + kGroupWaves * t;
          if (tile < MT * NT) {
            const int mt = tile / NT, nt = tile % NT;
            const Col col = inv_col(winds, blk, nt * 16 + i, NC, NCm, P, kPairRawK);
            const float* a = A + (mt * 16 + i) * as + col0 + kk;
            const float* b = raw + col.base + 2 * (col0 + kk);
            f4 c = acc[t];
            for (int r = 0; r < R; r += 4)
              c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], col.sign * b[2 * r], c, 0, 0, 0);
            acc[t] = c;
          }
        }
      }
#pragma unroll
      for (int t = 0; t < kGTiles; t++) {
        const int tile = wave + kGroupWaves * t;
        if (tile < MT * NT) {
          const int mt = tile / NT, nt = tile % NT;
          int c = nt * 16 + i;
          if (c < NC) {
            const int g = c / NCm;
            c -= g * NCm;
            exo_gfloat* base = group_arena(plan, grp, c0 + g) + M.slot[winds && (c & 2) ? job.four1 : job.four0];
            const int p = winds ? c >> 2 : c >> 1, part = c & 1;
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int lat = mt * 16 + 4 * kk + r;
              base[(2 * m + part) * (P * nlat) + p * nlat + lat] = acc[t][r];
            }
          }
        }
      }
    }
    if (n1 > 0) {
      lds_barrier();  // every wave is done with this chunk's inputs
      store_cols<kPairRawN, 2 * kPairK>(n1 * nf * P, raw_storer(c1), vn);
      lds_barrier();
    }
  }
}

// ------------------------------------------------------------- forward
// As fc2sp folds them (forward.wgsl:24-36), the sums run over the northern
// latitudes of the folded rows E = F(north) + F(south) and O = F(north) -
// F(south): with the table's parity across the equator (as for the
// inverse), mode k = n - m of a table of P takes E for even k and O for
// odd k, a table of dP/dmu (qe) the other. The staged table rows are the
// even modes of m (padded to a multiple of 16 rows), then its odd ones,
// so each 16-row output tile reads E or O per block.

// forward staging with a fixed column per thread: table half x 64 (8 rows
// per pass); inputs rows x (2 parts x 48 latitude pairs) (5 rows per pass),
// each thread loading a latitude and its mirror
constexpr int kFwdColTabN = (kMaxLat / 2 + kGroupThreads / 64 - 1) / (kGroupThreads / 64);
constexpr int kFwdFoldS = kGroupThreads / kMaxLat;
constexpr int kFwdFoldN = (kFwdRows + kFwdFoldS - 1) / kFwdFoldS;

// The chunk's folded inputs: thread t takes latitude pair column c = t %
// 96 (part c / half, northern latitude c % half) of rows t / 96 + 5 j.
template <typename Rows>
__device__ __forceinline__ void fold_load(int rows, int nlat, int plane_stride, Rows rowp, float (*v)[2]) {
  const int half = nlat >> 1;
  const int c = threadIdx.x % kMaxLat, r0 = threadIdx.x / kMaxLat;
  const int part = c >= half, l = c - part * half;
  const bool ok = r0 < kFwdFoldS && c < 2 * half;
  // every lane loads (past the rows or the pairs: the last row, latitude
  // pair (0, nlat - 1)), which fold_store then skips
  const int lc = ok ? l : 0, pc = ok ? part : 0;
#pragma unroll
  for (int j = 0; j < kFwdFoldN; j++) {
    const auto x = rowp(min(r0 + kFwdFoldS * j, rows - 1)) + pc * plane_stride;
    v[j][0] = x[lc];
    v[j][1] = x[nlat - 1 - lc];
  }
}

// E at raw[row][2 l + part], O at raw[row][kMaxLat + 2 l + part]
__device__ __forceinline__ void fold_store(int rows, int nlat, float* raw, const float (*v)[2]) {
  const int half = nlat >> 1;
  const int c = threadIdx.x % kMaxLat, r0 = threadIdx.x / kMaxLat;
  const int part = c >= half, l = c - part * half;
  if (r0 >= kFwdFoldS || c >= 2 * half) return;
#pragma unroll
  for (int j = 0; j < kFwdFoldN; j++) {
    const int r = r0 + kFwdFoldS * j;
    if (r < rows) {
      raw[r * kRawLat + 2 * l + part] = v[j][0] + v[j][1];
      raw[r * kRawLat + kMaxLat + 2 * l + part] = v[j][0] - v[j][1];
    }
  }
}

template <int P>
__device__ void leg_fwd(const DevMember* mems, const GroupPlan& plan, int grp, const LegJob& job, int m, float* Ah,
                        float* raw, exo_gcfloat** rowp) {
  const DevMember& M = mems[group_member(plan, grp, 0)];
  const int count = plan.count[grp];
  const int nlat = M.nlat, half = nlat / 2, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp;
  const bool uv = job.kind == LEG_UV2DV;
  const int NCm = uv ? 4 * P : 2 * P;
  // even modes in rows [0, re16), odd ones in [re16, re16 + ro16)
  const int re16 = ((K + 1) / 2 + 15) & ~15, ro16 = (K / 2 + 15) & ~15;
  const int MTe = re16 / 16, MT = (re16 + ro16) / 16;
  const int ash = a_stride(half);
  // blocks: table and whether it is a table of dP/dmu (qe)
  int tables[3], odd[3], blocks = 0;
  if (job.kind == LEG_PLAIN) {
    tables[blocks] = TAB_legendre_qc, odd[blocks++] = 0;
  } else if (uv) {
    tables[blocks] = TAB_legendre_qe, odd[blocks++] = 1;
    tables[blocks] = TAB_legendre_qm, odd[blocks++] = 0;
    if (job.ke >= 0) tables[blocks] = TAB_legendre_qq, odd[blocks++] = 0;
  } else {
    tables[blocks] = TAB_legendre_qe, odd[blocks++] = 1;
    tables[blocks] = TAB_legendre_qc, odd[blocks++] = 0;
    tables[blocks] = TAB_legendre_qm, odd[blocks++] = 0;
  }
  const int fields[3] = {job.kind == LEG_PLAIN ? job.src : job.u, job.v, uv ? job.ke : job.src};
  const int nf = job.kind == LEG_PLAIN ? 1 : (uv && job.ke < 0) ? 2 : 3;
  const int chunk = max(1, min(count, kFwdRows / (nf * P)));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  // staged table row r: mode k of m, or -1 for padding
  const auto row_mode = [=](int r) {
    const int k = r < re16 ? 2 * r : 2 * (r - re16) + 1;
    return k < K ? k : -1;
  };
  // one memory latency for the northern half of each table,
  // Ah[r][l] = q(l, w0 + k(r)), and the first chunk's inputs; the tables
  // need nothing from the members, so their loads go first and stay in
  // flight while the row pointers are set up and published
  float va[3][kFwdColTabN];
#pragma unroll
  for (int blk = 0; blk < 3; blk++)
    if (blk < blocks) {
      const float* q = dev_ltab(M) + M.tabo[tables[blk]];
      load_cols<kFwdColTabN, 64>(half, [&](int l, int r) { return q[l * ncsp + w0 + max(row_mode(r), 0)]; },
                                 va[blk]);
    }
  // where each input row of the group starts: Fourier coefficient m of
  // latitude 0 of member g, field f, level p at row (g nf + f) P + p
  if (threadIdx.x < count * nf * P) {
    const int row = threadIdx.x, g = row / (nf * P), fp = row - g * nf * P, f = fp / P, p = fp - f * P;
    rowp[row] = group_arena(plan, grp, g) + M.slot[fields[f]] + 2 * m * (P * nlat) + p * nlat;
  }
  lds_barrier();  // rowp
  {
    float vr[kFwdFoldN][2];
    fold_load(min(chunk, count) * nf * P, nlat, P * nlat, [&](int r) { return rowp[r]; }, vr);
#pragma unroll
    for (int blk = 0; blk < 3; blk++)
      if (blk < blocks) {
        float* A = Ah + blk * kFwdAs;
        store_cols<kFwdColTabN, 64>(half, [&](int l, int r, float v) {
          if (r < MT * 16) A[r * ash + l] = row_mode(r) >= 0 ? v : 0.0f;
        },
                                    va[blk]);
      }
    fold_store(min(chunk, count) * nf * P, nlat, raw, vr);
  }
  lds_barrier();
  for (int c0 = 0; c0 < count; c0 += chunk) {
    const int n = min(chunk, count - c0);
    const int NC = n * NCm, NT = (NC + 15) / 16;
    // the next chunk's inputs are in flight while this chunk's products run
    const int c1 = c0 + chunk, n1 = c1 < count ? min(chunk, count - c1) : 0;
    float vn[kFwdFoldN][2];
    if (n1 > 0) fold_load(n1 * nf * P, nlat, P * nlat, [&](int r) { return rowp[c1 * nf * P + r]; }, vn);
    f4 acc[kFwdGTiles];
#pragma unroll
    for (int t = 0; t < kFwdGTiles; t++) acc[t] = f4{0, 0, 0, 0};
    for (int blk = 0; blk < blocks; blk++) {
      const float* A = Ah + blk * kFwdAs;
#pragma unroll
      for (int t = 0; t < kFwdGTiles; t++) {
        const int tile = wave + kGroupWaves * t;
        if (tile < MT * NT) {
          const int mt = tile / NT, nt = tile % NT;
          // E for the symmetric pairs (even mode of a table of P, odd mode
          // of a table of dP/dmu), O for the others
          const int eo = ((mt >= MTe) + odd[blk]) & 1;
          const Col col = fwd_col(job.kind, blk, nt * 16 + i, NC, NCm, nf, P);
          const float* a = A + (mt * 16 + i) * ash + kk;
          const float* b = raw + col.base + eo * kMaxLat + 2 * kk;
          f4 c = acc[t];
          for (int r = 0; r < half; r += 4)
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], col.sign * b[2 * r], c, 0, 0, 0);
          acc[t] = c;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < kFwdGTiles; t++) {
      const int tile = wave + kGroupWaves * t;
      if (tile < MT * NT) {
        const int mt = tile / NT, nt = tile % NT;
        int c = nt * 16 + i;
        if (c < NC) {
          const int g = c / NCm;
          c -= g * NCm;
          exo_gfloat* base = group_arena(plan, grp, c0 + g) + M.slot[uv && (c & 2) ? job.dst1 : job.dst0];
          const int p = uv ? c >> 2 : c >> 1, part = c & 1;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int k = row_mode(mt * 16 + 4 * kk + r);
            if (k >= 0) base[p * nrsp + 2 * (w0 + k) + part] = acc[t][r];
          }
        }
      }
    }
    if (n1 > 0) {
      lds_barrier();  // every wave is done with this chunk's inputs
      fold_store(n1 * nf * P, nlat, raw, vn);
      lds_barrier();
    }
  }
}

// ---------------------------------------------- split-operand form (groups)
// With M.precision = 1 the products run on v_mfma_f32_16x16x16_f16 with
// every operand split into FP16 hi + lo parts, x s = hi + lo (hi the
// nearest FP16, lo the FP16 of the remainder), as hi.hi + hi.lo + lo.hi
// with FP32 accumulation: 22 significant bits per operand against FP32's
// 24. The scale s is a power of two per Legendre table (set once, from
// the whole table) and per staged input row (level and part of one field
// at wavenumber m), chosen so the largest magnitude lands in
// [2^12, 2^13), well inside FP16's range; the sums are unscaled exactly.
//
// The workgroups are those of the FP32 form (member groups sharing the
// tables; inverse per wavenumber pair, forward per wavenumber). Operands
// are split as they are staged: table values scaled and stored split,
// input rows loaded one wave per row so that the row's scale is one wave
// reduction, then stored split. Each group of 4 values is stored as 4 hi
// halves then 4 lo halves, so a lane's MFMA operands (4 consecutive
// reduction values, hi and lo) are one 16-byte LDS read. A wavenumber's
// reduction runs are padded with zeros to multiples of 16, the MFMA depth.

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// Magnitudes are reduced as bit patterns: for non-negative floats the
// integer order is the numeric order, and integer max needs no NaN
// canonicalisation.
__device__ __forceinline__ int abs_bits(float x) { return __float_as_int(x) & 0x7fffffff; }

// Power of two putting a largest magnitude (as bits) in [2^12, 2^13): with
// biased exponent E, x in [2^(E-127), 2^(E-126)) takes 2^(139-E), E held
// to >= 26 (tiny rows get 2^113); 1 for zero.
__device__ __forceinline__ float pow2_scale_bits(int b) {
  if (b == 0) return 1.0f;
  const int e = max(b >> 23, 26);
  return __int_as_float((266 - e) << 23);
}

// Largest of b over the 64 lanes, in every lane: DPP within quads, half
// rows and rows (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror), then the four rows' values read out (v_readlane). No LDS
// round trip, unlike a shuffle.
template <int Ctrl>
__device__ __forceinline__ int dpp_step_max(int b) {
  return max(b, __builtin_amdgcn_mov_dpp(b, Ctrl, 0xF, 0xF, true));
}
__device__ __forceinline__ int wave_max_bits(int b) {
  b = dpp_step_max<0xB1>(b);
  b = dpp_step_max<0x4E>(b);
  b = dpp_step_max<0x141>(b);
  b = dpp_step_max<0x140>(b);
  return max(max(__builtin_amdgcn_readlane(b, 0), __builtin_amdgcn_readlane(b, 16)),
             max(__builtin_amdgcn_readlane(b, 32), __builtin_amdgcn_readlane(b, 48)));
}

// Largest of b over each group of `lpr` lanes (16, 32 or 64, aligned), in
// every lane of the group: the DPP steps leave each 16-lane row's largest
// in its lanes, and the rows' values combine by v_readlane.
__device__ __forceinline__ int rows_max_bits(int b, int lpr) {
  b = dpp_step_max<0xB1>(b);
  b = dpp_step_max<0x4E>(b);
  b = dpp_step_max<0x141>(b);
  b = dpp_step_max<0x140>(b);
  if (lpr == 16) return b;
  const int r0 = __builtin_amdgcn_readlane(b, 0), r1 = __builtin_amdgcn_readlane(b, 16);
  const int r2 = __builtin_amdgcn_readlane(b, 32), r3 = __builtin_amdgcn_readlane(b, 48);
  if (lpr == 32) return (threadIdx.x & 63) < 32 ? max(r0, r1) : max(r2, r3);
  return max(max(r0, r1), max(r2, r3));
}

// x (already scaled) as hi and lo halves at value index v of a split row.
__device__ __forceinline__ void put_split(_Float16* row, int v, float x) {
  const _Float16 h = (_Float16)x;
  row[(v >> 2) * 8 + (v & 3)] = h;
  row[(v >> 2) * 8 + 4 + (v & 3)] = (_Float16)(x - (float)h);
}

__device__ __forceinline__ f4 mfma_split(const h8& a, const h8& b, f4 c) {
  const h4 ah = {a[0], a[1], a[2], a[3]}, al = {a[4], a[5], a[6], a[7]};
  const h4 bh = {b[0], b[1], b[2], b[3]}, bl = {b[4], b[5], b[6], b[7]};
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, c, 0, 0, 0);
}
__device__ __forceinline__ h8 load_h8(const _Float16* p) { return *reinterpret_cast<const h8*>(p); }

// Inverse: pair columns, each wavenumber's padded to 16: Ra + Rb <= 80 at
// T63 (K(m) + K(T - m) = T + 2). Staged FP32 rows: table [lat][col] with
// kSplitAs floats (20 mod 32: the 8 rows of one 16-byte read pass fall in
// distinct banks), inputs [row][2 col + part] with kSplitRaw floats, at
// most kSplitInvRows of them (one dv2uv member at L10), so two workgroups
// fit a CU's LDS; split in place, a table row is 2 kSplitAs halves (col
// groups of 8), an input row 2 kSplitRaw halves, part p's groups from
// p * 2 kSplitPairK.
constexpr int kSplitPairK = 80;
constexpr int kSplitAs = 84;
constexpr int kSplitRaw = 164;
constexpr int kSplitInvRows = 2 * kMaxLev;
constexpr int kSplitTabN = (kMaxLat + kGroupThreads / kSplitPairK - 1) / (kGroupThreads / kSplitPairK);
constexpr int kSplitRawN =
    (kSplitInvRows + kGroupThreads / (2 * kSplitPairK) - 1) / (kGroupThreads / (2 * kSplitPairK));
static_assert(2 * kSplitPairK <= kSplitRaw && kSplitPairK <= kSplitAs, "split rows");

// Split staging: the split is done at the LDS store, so no pass re-reads
// the staged FP32 values. Table values are scaled by the table's scale and
// stored split where they land. Input rows are loaded one wave per row
// (rows wave, wave + 8, ...): each lane holds columns lane and lane + 64
// of its row, both parts, so the row's largest magnitude per part is one
// wave reduction, and
// the values are stored split with that scale. Every load is unconditional
// (rows past the end load the last row, padding loads a valid element that
// the store zeroes), so a chunk's loads are all in flight before any use.
constexpr int kSplitInvRowsPerWave = (kSplitInvRows + kGroupWaves - 1) / kGroupWaves;  // 3
static_assert(kSplitPairK <= 128, "two columns a lane");
constexpr int kSplitFwdRowsPerWave = (kFwdRows + kGroupWaves - 1) / kGroupWaves;       // 4

// Lane values: columns col = lane + 64 j (j < 2), both parts, v[jr][2 j +
// part]; `load(row, col)` gives both parts of a column.
template <typename Load>
__device__ __forceinline__ void inv_split_load(int rows, Load load, float (*v)[4]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int jr = 0; jr < kSplitInvRowsPerWave; jr++) {
    const int r = min(wave + kGroupWaves * jr, rows - 1);
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const f2 x = load(r, lane + 64 * j);
      v[jr][2 * j] = x[0];
      v[jr][2 * j + 1] = x[1];
    }
  }
}

// Row r's columns col < width (`fix` giving the staged value of value x =
// 2 col + part) as split halves part * 2 kSplitPairK + group layout of col,
// and the per-part scale in scale[2 r + part].
template <typename Fix>
__device__ __forceinline__ void inv_split_store(int rows, int width, Fix fix, const float (*v)[4], float* raw,
                                                float* scale) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int jr = 0; jr < kSplitInvRowsPerWave; jr++) {
    const int r = wave + kGroupWaves * jr;
    if (r < rows) {
      float x[4];
      int mx[2] = {0, 0};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int c = lane + 64 * (k >> 1), part = k & 1;
        x[k] = c < width ? fix(r, 2 * c + part, v[jr][k]) : 0.0f;
        mx[part] = max(mx[part], abs_bits(x[k]));
      }
      const float sc[2] = {pow2_scale_bits(wave_max_bits(mx[0])), pow2_scale_bits(wave_max_bits(mx[1]))};
      if (lane < 2) scale[2 * r + lane] = lane ? sc[1] : sc[0];
      _Float16* h = reinterpret_cast<_Float16*>(raw + r * kSplitRaw);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int c = lane + 64 * (k >> 1), part = k & 1;
        if (c < width) put_split(h + part * (2 * kSplitPairK), c, x[k] * sc[part]);
      }
    }
  }
}

template <int P>
__device__ __forceinline__ void leg_inv_gsplit(const DevMember* mems, const GroupPlan& plan, int grp, const InvJob& job, int pair,
                               float* As, float* raw, exo_gcfloat** rowp, float* plav, float* bscale) {
  const DevMember& M = mems[group_member(plan, grp, 0)];
  const int count = plan.count[grp];
  const int nlat = M.nlat, ntp1 = M.ntru + 1;
  const int ncsp = M.ncsp, nrsp = M.nrsp;
  const int ma = pair, mb = ntp1 - 1 - pair, two = mb != ma;
  const int Ka = ntp1 - ma, Kb = two ? ntp1 - mb : 0;
  const int wa = mode_first(ma, ntp1), wb = two ? mode_first(mb, ntp1) : 0;
  const int Ra = (Ka + 15) & ~15, Rb = (Kb + 15) & ~15, RT = Ra + Rb;
  const bool winds = job.kind == INV_DV2UV;
  const int nf = winds ? 2 : 1;
  const int NCm = winds ? 4 * P : 2 * P;
  const int chunk = max(1, min(count, kSplitInvRows / (nf * P)));
  const int MT = nlat / 16;
  const int blocks = winds ? 2 : 1;
  const int tabs[2] = {winds ? (int)TAB_legendre_qu : job.table, (int)TAB_legendre_qv};
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  const auto mode_of = [=](int col) {
    if (col < Ra) return col < Ka ? wa + col : -1;
    const int k = col - Ra;
    return k < Kb ? wb + k : -1;
  };
  // the table columns of the pair need nothing from the members: their
  // loads go first and stay in flight while the row pointers are set up
  float va[2][kSplitTabN];
#pragma unroll
  for (int blk = 0; blk < 2; blk++)
    if (blk < blocks) {
      const float* q = dev_ltab(M) + M.tabo[tabs[blk]];
      load_cols<kSplitTabN, kSplitPairK>(nlat, [&](int lat, int col) {
        const int w = col < RT ? mode_of(col) : -1;
        return q[lat * ncsp + max(w, 0)];
      }, va[blk]);
    }
  if (threadIdx.x < count * nf * P) {
    const int row = threadIdx.x, g = row / (nf * P), fp = row - g * nf * P, f = fp / P, p = fp - f * P;
    rowp[row] = group_arena(plan, grp, g) + M.slot[f ? job.src1 : job.src0] + p * nrsp;
  }
  // each member's planetary vorticity, for its zeta rows' Re mode (0, 1)
  if (job.plavor && threadIdx.x < count) plav[threadIdx.x] = group_plavor(mems, plan, grp, threadIdx.x);
  // the members' coefficients for the chunk from member c0, value x = 2 col
  // + part of a row (columns past the pair load mode 0; the store zeroes it)
  const auto raw_loader = [&](int c0) {
    const int r0 = c0 * nf * P;
    return [=](int row, int col) {
      return *reinterpret_cast<const __attribute__((address_space(1))) f2*>(rowp[r0 + row] + 2 * max(mode_of(col), 0));
    };
  };
  // staged as loaded, zeros in the padding, except Re mode (0, 1) (x = 2
  // of pair 0) of a zeta row (field 1), which loses the member's planetary
  // vorticity
  const bool plavor = job.plavor && ma == 0;
  const auto raw_fix = [&](int c0) {
    const int r0 = c0 * nf * P;
    return [=](int row, int x, float v) {
      if (mode_of(x >> 1) < 0) v = 0.0f;
      if (plavor && x == 2) {
        const int r = r0 + row, g = r / (2 * P);
        if (r - g * 2 * P >= P) v -= plav[g];
      }
      return v;
    };
  };
  lds_barrier();  // rowp, plav
  {
    float vr[kSplitInvRowsPerWave][4];
    const int rows = min(chunk, count) * nf * P;
    inv_split_load(rows, raw_loader(0), vr);
#pragma unroll
    for (int blk = 0; blk < 2; blk++)
      if (blk < blocks) {
        _Float16* A = reinterpret_cast<_Float16*>(As + blk * kMaxLat * kSplitAs);
        const float sa = M.lscale[tabs[blk] - TAB_legendre_qc];
        store_cols<kSplitTabN, kSplitPairK>(nlat, [&](int lat, int col, float v) {
          put_split(A + lat * (2 * kSplitAs), col, (col < RT && mode_of(col) >= 0) ? v * sa : 0.0f);
        },
                                            va[blk]);
      }
    inv_split_store(rows, RT, raw_fix(0), vr, raw, bscale);
  }
  lds_barrier();
  for (int c0 = 0; c0 < count; c0 += chunk) {
    const int n = min(chunk, count - c0);
    const int NC = n * NCm, NT = (NC + 15) / 16;
    const int c1 = c0 + chunk, n1 = c1 < count ? min(chunk, count - c1) : 0;
    float vn[kSplitInvRowsPerWave][4];
    if (n1 > 0) inv_split_load(n1 * nf * P, raw_loader(c1), vn);
    const _Float16* rawh = reinterpret_cast<const _Float16*>(raw);
    for (int h = 0; h <= two; h++) {
      const int m = h ? mb : ma, col0 = h ? Ra : 0, R = h ? Rb : Ra;
      f4 acc[kGTiles];
#pragma unroll
      for (int t = 0; t < kGTiles; t++) acc[t] = f4{0, 0, 0, 0};
      for (int blk = 0; blk < blocks; blk++) {
        const _Float16* A = reinterpret_cast<const _Float16*>(As + blk * kMaxLat * kSplitAs);
        const float sa = M.lscale[tabs[blk] - TAB_legendre_qc];
#pragma unroll
        for (int t = 0; t < kGTiles; t++) {
          const int tile = wave + kGroupWaves * t;
          if (tile < MT * NT) {
            const int mt = tile / NT, nt = tile % NT;
            const Col col = inv_col(winds, blk, nt * 16 + i, NC, NCm, P, kSplitRaw);
            const int brow = col.base / kSplitRaw, part = col.base - brow * kSplitRaw;
            // 8-half groups: lane group kk's 4 values of each depth-16 step
            const _Float16* a = A + (mt * 16 + i) * (2 * kSplitAs) + 2 * (col0 + 4 * kk);
            const _Float16* b = rawh + brow * (2 * kSplitRaw) + part * (2 * kSplitPairK) + 2 * (col0 + 4 * kk);
            f4 c = f4{0, 0, 0, 0};
            for (int r = 0; r < R; r += 16) c = mfma_split(load_h8(a + 2 * r), load_h8(b + 2 * r), c);
            const float f = col.sign == 0.0f ? 0.0f : col.sign / (sa * bscale[2 * brow + part]);
            acc[t] += c * f;
          }
        }
      }
#pragma unroll
      for (int t = 0; t < kGTiles; t++) {
        const int tile = wave + kGroupWaves * t;
        if (tile < MT * NT) {
          const int mt = tile / NT, nt = tile % NT;
          int c = nt * 16 + i;
          if (c < NC) {
            const int g = c / NCm;
            c -= g * NCm;
            exo_gfloat* base = group_arena(plan, grp, c0 + g) + M.slot[winds && (c & 2) ? job.four1 : job.four0];
            const int p = winds ? c >> 2 : c >> 1, part = c & 1;
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int lat = mt * 16 + 4 * kk + r;
              base[(2 * m + part) * (P * nlat) + p * nlat + lat] = acc[t][r];
            }
          }
        }
      }
    }
    if (n1 > 0) {
      lds_barrier();  // every wave is done with this chunk's inputs
      inv_split_store(n1 * nf * P, RT, raw_fix(c1), vn, raw, bscale);
      lds_barrier();
    }
  }
}

// Forward: the plain form's fold, on split operands. A[r][l] over the
// northern latitudes, rows in the plain form's order (even modes of m,
// then odd ones, each padded to 16 rows), kSplitFwdAs floats a row split
// in place to 2 kSplitFwdAs halves. Each staged input row holds E at
// [0, 2 half) and O at [kMaxLat, kMaxLat + 2 half) floats, both parts
// interleaved; each (E or O, part) run of `half` values has its own scale
// and is split in place within its region, part p's groups from p kMaxLat
// halves. A 16-row output tile then takes one depth-half product per
// block against E or O, as in the plain form.
constexpr int kSplitFwdAs = 68;
static_assert(kMaxLat / 2 <= kSplitFwdAs && 2 * kMaxLat <= kRawLat, "split forward rows");

// A wave stages 64 / lpr rows a pass, lpr = 16, 32 or 64 lanes a row (the
// northern latitudes rounded up: T21, T42, T63): pass j of wave w takes rows
// w + 8 (j q + s), s = lane / lpr, q = 64 / lpr. Lane l = lane % lpr < half
// loads latitude l and its mirror nlat - 1 - l of both parts (part p at p
// plane_stride); every load is clamped (rows past the end load the last
// row, lanes past half latitude 0) and issued before any use.
__device__ __forceinline__ int fwd_lanes_per_row(int nlat) {
  return nlat <= 32 ? 16 : nlat <= 64 ? 32 : 64;
}

template <typename Rows>
__device__ __forceinline__ void fwd_split_load(int rows, int nlat, int plane_stride, Rows rowp, float (*v)[4]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lpr = fwd_lanes_per_row(nlat), q = 64 / lpr, sub = lane / lpr, l0 = lane & (lpr - 1);
  const int l = l0 < (nlat >> 1) ? l0 : 0;
#pragma unroll
  for (int jr = 0; jr < kSplitFwdRowsPerWave; jr++) {
    const auto x = rowp(min(wave + kGroupWaves * (jr * q + sub), rows - 1));
    v[jr][0] = x[l];
    v[jr][1] = x[nlat - 1 - l];
    v[jr][2] = x[plane_stride + l];
    v[jr][3] = x[plane_stride + nlat - 1 - l];
  }
}

// Each staged row's folded runs E = north + south and O = north - south of
// each part, each with its own scale scale[4 r + 2 eo + part], stored
// split: run (eo, part) at halves eo 2 kMaxLat + part kMaxLat of the row.
// The arithmetic runs in every lane (the reductions need them all); the
// stores are guarded.
__device__ __forceinline__ void fwd_split_store(int rows, int nlat, const float (*v)[4], float* raw, float* scale) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lpr = fwd_lanes_per_row(nlat), q = 64 / lpr, sub = lane / lpr, l = lane & (lpr - 1);
  const bool ok = l < (nlat >> 1);
#pragma unroll
  for (int jr = 0; jr < kSplitFwdRowsPerWave; jr++) {
    if (wave + kGroupWaves * jr * q < rows) {
      const int r = wave + kGroupWaves * (jr * q + sub);
      // k = 2 eo + part
      float x[4];
      x[0] = ok ? v[jr][0] + v[jr][1] : 0.0f;
      x[1] = ok ? v[jr][2] + v[jr][3] : 0.0f;
      x[2] = ok ? v[jr][0] - v[jr][1] : 0.0f;
      x[3] = ok ? v[jr][2] - v[jr][3] : 0.0f;
      float sc[4];
#pragma unroll
      for (int k = 0; k < 4; k++) sc[k] = pow2_scale_bits(rows_max_bits(abs_bits(x[k]), lpr));
      if (r < rows) {
        if (l < 4) scale[4 * r + l] = l == 0 ? sc[0] : l == 1 ? sc[1] : l == 2 ? sc[2] : sc[3];
        _Float16* h = reinterpret_cast<_Float16*>(raw + r * kRawLat);
        if (ok) {
#pragma unroll
          for (int k = 0; k < 4; k++) put_split(h + (k >> 1) * (2 * kMaxLat) + (k & 1) * kMaxLat, l, x[k] * sc[k]);
        }
      }
    }
  }
}

template <int P>
__device__ __forceinline__ void leg_fwd_gsplit(const DevMember* mems, const GroupPlan& plan, int grp, const LegJob& job, int m,
                               float* Ah, float* raw, exo_gcfloat** rowp, float* bscale) {
  const DevMember& M = mems[group_member(plan, grp, 0)];
  const int count = plan.count[grp];
  const int nlat = M.nlat, half = nlat / 2, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp;
  const bool uv = job.kind == LEG_UV2DV;
  const int NCm = uv ? 4 * P : 2 * P;
  const int re16 = ((K + 1) / 2 + 15) & ~15, ro16 = (K / 2 + 15) & ~15;
  const int MTe = re16 / 16, MT = (re16 + ro16) / 16;
  int tables[3], odd[3], blocks = 0;
  if (job.kind == LEG_PLAIN) {
    tables[blocks] = TAB_legendre_qc, odd[blocks++] = 0;
  } else if (uv) {
    tables[blocks] = TAB_legendre_qe, odd[blocks++] = 1;
    tables[blocks] = TAB_legendre_qm, odd[blocks++] = 0;
    if (job.ke >= 0) tables[blocks] = TAB_legendre_qq, odd[blocks++] = 0;
  } else {
    tables[blocks] = TAB_legendre_qe, odd[blocks++] = 1;
    tables[blocks] = TAB_legendre_qc, odd[blocks++] = 0;
    tables[blocks] = TAB_legendre_qm, odd[blocks++] = 0;
  }
  const int fields[3] = {job.kind == LEG_PLAIN ? job.src : job.u, job.v, uv ? job.ke : job.src};
  const int nf = job.kind == LEG_PLAIN ? 1 : (uv && job.ke < 0) ? 2 : 3;
  const int chunk = max(1, min(count, kFwdRows / (nf * P)));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kk = lane >> 4;
  const auto row_mode = [=](int r) {
    const int k = r < re16 ? 2 * r : 2 * (r - re16) + 1;
    return k < K ? k : -1;
  };
  // the tables need nothing from the members: their loads go first and stay
  // in flight while the row pointers are set up
  float va[3][kFwdColTabN];
#pragma unroll
  for (int blk = 0; blk < 3; blk++)
    if (blk < blocks) {
      const float* q = dev_ltab(M) + M.tabo[tables[blk]];
      load_cols<kFwdColTabN, 64>(half, [&](int l, int r) {
        const int k = row_mode(r);
        return q[l * ncsp + w0 + max(k, 0)];
      }, va[blk]);
    }
  if (threadIdx.x < count * nf * P) {
    const int row = threadIdx.x, g = row / (nf * P), fp = row - g * nf * P, f = fp / P, p = fp - f * P;
    rowp[row] = group_arena(plan, grp, g) + M.slot[fields[f]] + 2 * m * (P * nlat) + p * nlat;
  }
  lds_barrier();  // rowp
  {
    float vr[kSplitFwdRowsPerWave][4];
    const int rows = min(chunk, count) * nf * P;
    fwd_split_load(rows, nlat, P * nlat, [&](int r) { return rowp[r]; }, vr);
#pragma unroll
    for (int blk = 0; blk < 3; blk++)
      if (blk < blocks) {
        _Float16* A = reinterpret_cast<_Float16*>(Ah + blk * kMaxK * kSplitFwdAs);
        const float sa = M.lscale[tables[blk] - TAB_legendre_qc];
        store_cols<kFwdColTabN, 64>(half, [&](int l, int r, float v) {
          if (r < MT * 16) put_split(A + r * (2 * kSplitFwdAs), l, row_mode(r) >= 0 ? v * sa : 0.0f);
        },
                                    va[blk]);
      }
    fwd_split_store(rows, nlat, vr, raw, bscale);
  }
  lds_barrier();
  for (int c0 = 0; c0 < count; c0 += chunk) {
    const int n = min(chunk, count - c0);
    const int NC = n * NCm, NT = (NC + 15) / 16;
    const int c1 = c0 + chunk, n1 = c1 < count ? min(chunk, count - c1) : 0;
    float vn[kSplitFwdRowsPerWave][4];
    if (n1 > 0) fwd_split_load(n1 * nf * P, nlat, P * nlat, [&](int r) { return rowp[c1 * nf * P + r]; }, vn);
    const _Float16* rawh = reinterpret_cast<const _Float16*>(raw);
    f4 acc[kFwdGTiles];
#pragma unroll
    for (int t = 0; t < kFwdGTiles; t++) acc[t] = f4{0, 0, 0, 0};
    for (int blk = 0; blk < blocks; blk++) {
      const _Float16* A = reinterpret_cast<const _Float16*>(Ah + blk * kMaxK * kSplitFwdAs);
      const float sa = M.lscale[tables[blk] - TAB_legendre_qc];
#pragma unroll
      for (int t = 0; t < kFwdGTiles; t++) {
        const int tile = wave + kGroupWaves * t;
        if (tile < MT * NT) {
          const int mt = tile / NT, nt = tile % NT;
          const int eo = ((mt >= MTe) + odd[blk]) & 1;
          const Col col = fwd_col(job.kind, blk, nt * 16 + i, NC, NCm, nf, P);
          const int brow = col.base / kRawLat, part = col.base - brow * kRawLat;
          const _Float16* a = A + (mt * 16 + i) * (2 * kSplitFwdAs) + 2 * (4 * kk);
          const _Float16* b = rawh + brow * (2 * kRawLat) + eo * (2 * kMaxLat) + part * kMaxLat + 2 * (4 * kk);
          f4 c = f4{0, 0, 0, 0};
          for (int r = 0; r < half; r += 16) c = mfma_split(load_h8(a + 2 * r), load_h8(b + 2 * r), c);
          const float f = col.sign == 0.0f ? 0.0f : col.sign / (sa * bscale[4 * brow + 2 * eo + part]);
          acc[t] += f * c;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < kFwdGTiles; t++) {
      const int tile = wave + kGroupWaves * t;
      if (tile < MT * NT) {
        const int mt = tile / NT, nt = tile % NT;
        int c = nt * 16 + i;
        if (c < NC) {
          const int g = c / NCm;
          c -= g * NCm;
          exo_gfloat* base = group_arena(plan, grp, c0 + g) + M.slot[uv && (c & 2) ? job.dst1 : job.dst0];
          const int p = uv ? c >> 2 : c >> 1, part = c & 1;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int k = row_mode(mt * 16 + 4 * kk + r);
            if (k >= 0) base[p * nrsp + 2 * (w0 + k) + part] = acc[t][r];
          }
        }
      }
    }
    if (n1 > 0) {
      lds_barrier();  // every wave is done with this chunk's inputs
      fwd_split_store(n1 * nf * P, nlat, vn, raw, bscale);
      lds_barrier();
    }
  }
}

}  // namespace

// One kernel per operand form, so each is allocated only the registers and
// LDS of its own path; the step launches each over the members of its form.
#define EXO_P_DISPATCH(P, call) \
  do {                          \
    if ((P) == 10)              \
      call(10);                 \
    else if ((P) == 5)          \
      call(5);                  \
    else                        \
      call(1);                  \
  } while (0)

extern "C" __global__ __launch_bounds__(kGroupThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_leg_inv(const DevMember* __restrict__ mems,
                                                                       GroupPlan plan, InvJobs jobs) {
  __shared__ float As[2 * kPairAs];
  __shared__ float raw[kInvRows * kPairRawK];
  __shared__ exo_gcfloat* rowp[kInvGroupRows];
  __shared__ float plav[kLegGroup];
  int tile;
  const int grp = group_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile);
  const DevMember& M = mems[group_member(plan, grp, 0)];
  // job-major: neighbouring wavenumber pairs of one job, whose Fourier
  // rows share cache lines, go to one XCD (xcd_block)
  const int pairs = (M.ntru + 2) / 2;
  const InvJob job = jobs.j[job_of(jobs, tile / pairs)];
#define CALL(PP) leg_inv<PP>(mems, plan, grp, job, tile % pairs, As, raw, rowp, plav)
  EXO_P_DISPATCH(job.per_level ? M.nlev : 1, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(kGroupThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_leg_inv_split(const DevMember* __restrict__ mems,
                                                                             GroupPlan plan, InvJobs jobs) {
  __shared__ __align__(16) float As[2 * kMaxLat * kSplitAs];
  __shared__ __align__(16) float raw[kSplitInvRows * kSplitRaw];
  __shared__ exo_gcfloat* rowp[kInvGroupRows];
  __shared__ float plav[kLegGroup];
  __shared__ float bscale[2 * kSplitInvRows];
  int tile;
  const int grp = group_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile);
  const DevMember& M = mems[group_member(plan, grp, 0)];
  const int pairs = (M.ntru + 2) / 2;
  const InvJob job = jobs.j[job_of(jobs, tile / pairs)];
#define CALL(PP) leg_inv_gsplit<PP>(mems, plan, grp, job, tile % pairs, As, raw, rowp, plav, bscale)
  EXO_P_DISPATCH(job.per_level ? M.nlev : 1, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(kGroupThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_leg_fwd(const DevMember* __restrict__ mems,
                                                                       GroupPlan plan, LegJobs jobs) {
  __shared__ float Ah[3 * kFwdAs];
  __shared__ float raw[kFwdRows * kRawLat];
  __shared__ exo_gcfloat* rowp[kLegGroupRows];
  int tile;
  const int grp = group_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile);
  const DevMember& M = mems[group_member(plan, grp, 0)];
  // job-major: neighbouring wavenumbers of one job, whose Fourier rows
  // share cache lines, go to one XCD (xcd_block); m = 0 (longest) first
  const int ntp1 = M.ntru + 1;
  const LegJob job = jobs.j[job_of(jobs, tile / ntp1)];
#define CALL(PP) leg_fwd<PP>(mems, plan, grp, job, tile % ntp1, Ah, raw, rowp)
  EXO_P_DISPATCH(job.per_level ? M.nlev : 1, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(kGroupThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_leg_fwd_split(const DevMember* __restrict__ mems,
                                                                             GroupPlan plan, LegJobs jobs) {
  __shared__ __align__(16) float Ah[3 * kMaxK * kSplitFwdAs];
  __shared__ __align__(16) float raw[kFwdRows * kRawLat];
  __shared__ exo_gcfloat* rowp[kLegGroupRows];
  __shared__ float bscale[4 * kFwdRows];
  int tile;
  const int grp = group_find(plan, xcd_block(blockIdx.x, gridDim.x), &tile);
  const DevMember& M = mems[group_member(plan, grp, 0)];
  const int ntp1 = M.ntru + 1;
  const LegJob job = jobs.j[job_of(jobs, tile / ntp1)];
#define CALL(PP) leg_fwd_gsplit<PP>(mems, plan, grp, job, tile % ntp1, Ah, raw, rowp, bscale)
  EXO_P_DISPATCH(job.per_level ? M.nlev : 1, CALL);
#undef CALL
}
