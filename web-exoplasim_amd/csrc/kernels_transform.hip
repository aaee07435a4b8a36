// Spectral transforms of the ExoPlaSim step on gfx950, plain FP32 form.
//
// The reference runs one WGSL invocation per output value and rereads the
// Legendre and DFT tables for every plane (legendre/wgsl/*.wgsl,
// fourier/wgsl/*.wgsl). Here each workgroup owns one zonal wavenumber m
// (Legendre) or one latitude row (DFT) of one job and every level of it:
// the table block is staged in LDS once and reused for all levels, the
// spectral/Fourier operands come in as contiguous runs, and each lane
// keeps one accumulator per level in registers.
//
//  k_leg_inv  sp2fc / sp2fcdmu / dv2uv for wavenumber m: F(lat, p, m) =
//             sum_n q(lat, w(m, n)) S(p, w) over every latitude and level
//             (legendre/wgsl/inverse.wgsl:6-27, dv2uv.wgsl:9-43).
//  k_dft_inv  fc2gp (+ d/dlambda) of one latitude row, every level
//             (fourier/wgsl/inverse.wgsl:4-18, dlambda.wgsl:5-17).
//  k_dft_fwd  gp2fc of one latitude row, every level
//             (fourier/wgsl/forward.wgsl:4-18).
//  k_leg_fwd  fc2sp / uv2dv / mktend / qtend for wavenumber m over the
//             northern/southern row pairs (legendre/wgsl/forward.wgsl:7-38,
//             uv2dv.wgsl:11-63, qtend.wgsl:9-46).
//
// Sums run over n (inverse) and over row pairs (forward) in the reference's
// order. Every kernel walks a LaunchPlan so one launch serves a mixed batch.
#include "transforms.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxLat = 96, kMaxK = 64, kMaxLev = 10, kMaxFC = 128, kMaxLon = 192;
constexpr int kStride = kMaxK + 1;  // odd LDS row stride: no bank conflicts
constexpr int kLegFwdThreads = 256;  // k_leg_fwd: 4 waves, lane = mode (K <= kMaxK = 64)

__device__ __forceinline__ int mode_first(int m, int ntp1) { return m * (2 * ntp1 - m + 1) / 2; }

template <typename Jobs>
__device__ __forceinline__ int job_of(const Jobs& jobs, int index) {
  return index < jobs.n ? index : jobs.n - 1;
}

// Per-mode level vectors in LDS: for mode k, the (re, im) pairs of all P
// levels sit contiguously, padded to whole float4s, so one lane fetches a
// mode's every level with SV(P) / 4 128-bit LDS loads.
__host__ __device__ constexpr int SV(int P) { return (2 * P + 3) / 4 * 4; }

template <int P>
__device__ __forceinline__ void read_vec(const float* src, float* out) {
  const float4* v = reinterpret_cast<const float4*>(src);
#pragma unroll
  for (int i = 0; i < SV(P) / 4; i++) {
    const float4 t = v[i];
    out[4 * i] = t.x;
    out[4 * i + 1] = t.y;
    out[4 * i + 2] = t.z;
    out[4 * i + 3] = t.w;
  }
}

// sp2fc / dv2uv of wavenumber m. Lanes own latitudes; the kThreads / nlat
// lane groups (2 at T63, 4 at T42, 8 at T21) split the modes n
// (group g sums k = g, g + G, ...), then group 0 adds the other groups'
// partial sums in group order.
template <int P>
__device__ void leg_inv(const DevMember& M, const InvJob& job, int m, float* qab, float* sa, float* sb) {
  const int nlat = M.nlat, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc;
  const bool winds = job.kind == INV_DV2UV;
  float* qa = qab;
  float* qb = qab + kMaxLat * kStride;
  const float* q0 = M.tab + M.tabo[winds ? (int)TAB_legendre_qu : job.table];
  const float* q1 = M.tab + M.tabo[TAB_legendre_qv];
  for (int i = threadIdx.x; i < nlat * K; i += kThreads) {
    const int lat = i / K, k = i % K;
    qa[lat * kStride + k] = q0[lat * ncsp + w0 + k];
    if (winds) qb[lat * kStride + k] = q1[lat * ncsp + w0 + k];
  }
  const float* S0 = M.arena + M.slot[job.src0];
  const float* S1 = winds ? M.arena + M.slot[job.src1] : nullptr;
  const float plavor = job.plavor ? M.k[CON_plavor] : 0.0f;
  for (int i = threadIdx.x; i < P * 2 * K; i += kThreads) {
    const int p = i / (2 * K), r = i % (2 * K);
    const int at = (r >> 1) * SV(P) + 2 * p + (r & 1);
    sa[at] = S0[p * nrsp + 2 * w0 + r];
    // zeta without the planetary vorticity in Re mode (0, 1)
    if (winds) sb[at] = S1[p * nrsp + 2 * w0 + r] - ((w0 + (r >> 1) == 1 && (r & 1) == 0) ? plavor : 0.0f);
  }
  __syncthreads();
  const int G = kThreads / nlat;
  const int g = threadIdx.x / nlat, lat = threadIdx.x % nlat;
  const bool live = g < G;
  // plain: acc[2p + part]; winds: U in acc[2p + part], V in acc[2P + 2p + part]
  float acc[4 * P];
#pragma unroll
  for (int i = 0; i < 4 * P; i++) acc[i] = 0.0f;
  const int nacc = winds ? 4 * P : 2 * P;
  if (live && !winds) {
    for (int k = g; k < K; k += G) {
      const float qq = qa[lat * kStride + k];
      float v[SV(P)];
      read_vec<P>(sa + k * SV(P), v);
#pragma unroll
      for (int i = 0; i < 2 * P; i++) acc[i] += qq * v[i];
    }
  } else if (live) {
    for (int k = g; k < K; k += G) {
      const float u = qa[lat * kStride + k], v = qb[lat * kStride + k];
      float d[SV(P)], z[SV(P)];
      read_vec<P>(sa + k * SV(P), d);
      read_vec<P>(sb + k * SV(P), z);
#pragma unroll
      for (int p = 0; p < P; p++) {
        const float dre = d[2 * p], dim = d[2 * p + 1];
        const float zre = z[2 * p], zim = z[2 * p + 1];
        acc[2 * p] = acc[2 * p] + v * zre + u * dim;
        acc[2 * p + 1] = acc[2 * p + 1] + v * zim - u * dre;
        acc[2 * P + 2 * p] = acc[2 * P + 2 * p] + u * zim - v * dre;
        acc[2 * P + 2 * p + 1] = acc[2 * P + 2 * p + 1] - u * zre - v * dim;
      }
    }
  }
  // the table tiles are dead now: their LDS takes the partial sums
  constexpr int RS = 4 * P + 1;
  __syncthreads();
  if (live && g > 0) {
    float* r = qab + ((g - 1) * nlat + lat) * RS;
#pragma unroll
    for (int i = 0; i < 4 * P; i++)
      if (i < nacc) r[i] = acc[i];
  }
  __syncthreads();
  if (!live || g != 0) return;
  for (int j = 1; j < G; j++) {
    const float* r = qab + ((j - 1) * nlat + lat) * RS;
#pragma unroll
    for (int i = 0; i < 4 * P; i++)
      if (i < nacc) acc[i] += r[i];
  }
  float* F0 = M.arena + M.slot[job.four0];
#pragma unroll
  for (int p = 0; p < P; p++) {
    float* f = F0 + (p * nlat + lat) * nfc + 2 * m;
    f[0] = acc[2 * p];
    f[1] = acc[2 * p + 1];
  }
  if (winds) {
    float* F1 = M.arena + M.slot[job.four1];
#pragma unroll
    for (int p = 0; p < P; p++) {
      float* f = F1 + (p * nlat + lat) * nfc + 2 * m;
      f[0] = acc[2 * P + 2 * p];
      f[1] = acc[2 * P + 2 * p + 1];
    }
  }
}

// The zonal DFT matrices hold only NLON distinct cos/sin values:
// cos(m lambda_x) = cos(2 pi ((m x) mod NLON) / NLON). The twiddles are
// built in LDS from the FP32 inverse table (its m = 1 rows hold
// 2 cos(2 pi x / NLON) and -2 sin(2 pi x / NLON), exact doubles of the
// FP32 cos/sin), and every lane walks its (m x) mod NLON index.
__device__ void load_twiddles(const DevMember& M, float* ctw, float* stw) {
  const int nlon = M.nlon;
  const float* inv = M.tab + M.tabo[TAB_fourier_inverse];
  for (int x = threadIdx.x; x < nlon; x += kThreads) {
    ctw[x] = 0.5f * inv[2 * nlon + x];
    stw[x] = -0.5f * inv[3 * nlon + x];
  }
}

// Fourier rows of P levels into LDS as per-wavenumber level vectors:
// f[m * SV(P) + 2 p + part].
template <int P>
__device__ void load_rows_by_m(const float* F, int nlat, int nfc, float* f) {
  for (int i = threadIdx.x; i < P * nfc; i += kThreads) {
    const int p = i / nfc, k = i % nfc;
    f[(k >> 1) * SV(P) + 2 * p + (k & 1)] = F[p * nlat * nfc + k];
  }
}

template <int P>
__device__ void dft_inv(const DevMember& M, const InvJob& job, int lat, float* f0, float* f1, float* ctw,
                        float* stw) {
  const int nlat = M.nlat, nlon = M.nlon, nfc = M.nfc, nhor = M.nhor, ntp1 = M.ntru + 1;
  const bool two = job.kind != INV_PLAIN;
  load_rows_by_m<P>(M.arena + M.slot[job.four0] + lat * nfc, nlat, nfc, f0);
  if (job.kind == INV_DV2UV) load_rows_by_m<P>(M.arena + M.slot[job.four1] + lat * nfc, nlat, nfc, f1);
  load_twiddles(M, ctw, stw);
  __syncthreads();
  if (job.kind == INV_DLAMBDA) {
    for (int i = threadIdx.x; i < P * ntp1; i += kThreads) {
      const int p = i / ntp1, m = i % ntp1;
      const float re = f0[m * SV(P) + 2 * p], im = f0[m * SV(P) + 2 * p + 1];
      f1[m * SV(P) + 2 * p] = -im * (float)m;
      f1[m * SV(P) + 2 * p + 1] = re * (float)m;
    }
    __syncthreads();
  }
  float* G0 = M.arena + M.slot[job.dst0] + lat * nlon;
  float* G1 = two ? M.arena + M.slot[job.dst1] + lat * nlon : nullptr;
  for (int x = threadIdx.x; x < nlon; x += kThreads) {
    // fc2gp: g = Re c0 + sum_{m>0} 2 (Re c_m cos m lambda - Im c_m sin m lambda)
    float a[P], b[P];
    {
      float v0[SV(P)], v1[SV(P)];
      read_vec<P>(f0, v0);
      if (two) read_vec<P>(f1, v1);
#pragma unroll
      for (int p = 0; p < P; p++) {
        a[p] = v0[2 * p];
        b[p] = two ? v1[2 * p] : 0.0f;
      }
    }
    int idx = 0;
    for (int m = 1; m < ntp1; m++) {
      idx += x;
      if (idx >= nlon) idx -= nlon;
      const float c = 2.0f * ctw[idx], sn = -2.0f * stw[idx];
      float v0[SV(P)];
      read_vec<P>(f0 + m * SV(P), v0);
#pragma unroll
      for (int p = 0; p < P; p++) {
        a[p] += c * v0[2 * p];
        a[p] += sn * v0[2 * p + 1];
      }
      if (two) {
        float v1[SV(P)];
        read_vec<P>(f1 + m * SV(P), v1);
#pragma unroll
        for (int p = 0; p < P; p++) {
          b[p] += c * v1[2 * p];
          b[p] += sn * v1[2 * p + 1];
        }
      }
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
      G0[p * nhor + x] = a[p];
      if (two) G1[p * nhor + x] = b[p];
    }
  }
}

template <int P>
__device__ void dft_fwd(const DevMember& M, const FwdJob& job, int lat, float* row, float* ctw, float* stw) {
  // grid rows as per-longitude level vectors: row[x * PV + p]
  constexpr int PV = (P + 3) / 4 * 4;
  const int nlat = M.nlat, nlon = M.nlon, nfc = M.nfc, nhor = M.nhor;
  const float* G = M.arena + M.slot[job.src] + lat * nlon;
  for (int i = threadIdx.x; i < P * nlon; i += kThreads) {
    const int p = i / nlon, x = i % nlon;
    row[x * PV + p] = G[p * nhor + x];
  }
  load_twiddles(M, ctw, stw);
  __syncthreads();
  float* F = M.arena + M.slot[job.dst] + lat * nfc;
  const float scale = 1.0f / (float)nlon;
  // gp2fc: c_m = (1/N) sum_x g_x (cos m lambda_x - i sin m lambda_x); lane
  // k owns Re (even k) or Im (odd k) of wavenumber k / 2.
  for (int k = threadIdx.x; k < nfc; k += kThreads) {
    const int m = k >> 1;
    const float* tw = (k & 1) ? stw : ctw;
    const float sign = (k & 1) ? -1.0f : 1.0f;
    float a[PV];
#pragma unroll
    for (int p = 0; p < PV; p++) a[p] = 0.0f;
    int idx = 0;
    for (int x = 0; x < nlon; x++) {
      const float t = tw[idx];
      idx += m;
      if (idx >= nlon) idx -= nlon;
      const float4* v = reinterpret_cast<const float4*>(row + x * PV);
#pragma unroll
      for (int i = 0; i < PV / 4; i++) {
        const float4 g = v[i];
        a[4 * i] += t * g.x;
        a[4 * i + 1] += t * g.y;
        a[4 * i + 2] += t * g.z;
        a[4 * i + 3] += t * g.w;
      }
    }
#pragma unroll
    for (int p = 0; p < P; p++) F[p * nlat * nfc + k] = sign * a[p] * scale;
  }
}

// Fourier coefficient m of every row of `fields` Fourier slots into LDS as
// level vectors: fin[(f * nlat + lat) * SV(P) + 2 p + part].
template <int P>
__device__ void load_column(const DevMember& M, const int* slots, int fields, int m, float* fin) {
  const int nlat = M.nlat, nfc = M.nfc;
  const int per = P * nlat * 2;
  for (int i = threadIdx.x; i < fields * per; i += kLegFwdThreads) {
    const int f = i / per, r = i % per;
    const int p = r / (nlat * 2), lat = (r >> 1) % nlat, part = r & 1;
    fin[(f * nlat + lat) * SV(P) + 2 * p + part] =
        M.arena[M.slot[slots[f]] + (p * nlat + lat) * nfc + 2 * m + part];
  }
}

// Partial sums of waves 1..3 of k_leg_fwd, per lane: odd row stride.
constexpr int kRedStride = 4 * kMaxLev + 1;

template <int P>
__device__ void leg_fwd(const DevMember& M, const LegJob& job, int m, float* fin, float* red) {
  const int nlat = M.nlat, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1), nrsp = M.nrsp;
  const int ncsp = M.ncsp, half = nlat / 2;
  int slots[3];
  int fields = 0;
  if (job.kind == LEG_PLAIN) {
    slots[fields++] = job.src;
  } else if (job.kind == LEG_UV2DV) {
    slots[fields++] = job.u;
    slots[fields++] = job.v;
    if (job.ke >= 0) slots[fields++] = job.ke;
  } else {
    slots[fields++] = job.u;
    slots[fields++] = job.v;
    slots[fields++] = job.src;
  }
  load_column<P>(M, slots, fields, m, fin);
  __syncthreads();
  // lane k owns mode w0 + k for every level; wave wv sums latitude pairs
  // l = wv, wv + 4, ... Table columns come straight from memory
  // (consecutive lanes, consecutive modes).
  const int k = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool live = k < K;
  const float s = (k & 1) ? -1.0f : 1.0f;  // parity of m + n = 2m + k
  const int w = w0 + (live ? k : 0);
  auto row = [&](int f, int lat) { return fin + (f * nlat + lat) * SV(P); };
  // PLAIN / QTEND: acc[2p + part]; UV2DV: D in acc[2p + part], zeta in
  // acc[2P + 2p + part]
  float acc[4 * P];
#pragma unroll
  for (int i = 0; i < 4 * P; i++) acc[i] = 0.0f;
  const int nacc = job.kind == LEG_UV2DV ? 4 * P : 2 * P;
  if (live && job.kind == LEG_PLAIN) {
    const float* q = M.tab + M.tabo[TAB_legendre_qc] + w;
    for (int l = wv; l < half; l += 4) {
      const float qq = q[l * ncsp];
      float nv[SV(P)], sv[SV(P)];
      read_vec<P>(row(0, l), nv);
      read_vec<P>(row(0, nlat - 1 - l), sv);
#pragma unroll
      for (int i = 0; i < 2 * P; i++) acc[i] += qq * (nv[i] + s * sv[i]);
    }
  } else if (live && job.kind == LEG_UV2DV) {
    const bool kin = job.ke >= 0;
    const float* qe = M.tab + M.tabo[TAB_legendre_qe] + w;
    const float* qm = M.tab + M.tabo[TAB_legendre_qm] + w;
    const float* qq = M.tab + M.tabo[TAB_legendre_qq] + w;
    for (int l = wv; l < half; l += 4) {
      const int so = nlat - 1 - l;
      const float e = qe[l * ncsp], mm = qm[l * ncsp];
      float un[SV(P)], us[SV(P)], vn[SV(P)], vs[SV(P)];
      read_vec<P>(row(0, l), un);
      read_vec<P>(row(0, so), us);
      read_vec<P>(row(1, l), vn);
      read_vec<P>(row(1, so), vs);
      if (kin) {
        const float c = qq[l * ncsp];
        float kn[SV(P)], ks[SV(P)];
        read_vec<P>(row(2, l), kn);
        read_vec<P>(row(2, so), ks);
#pragma unroll
        for (int i = 0; i < 2 * P; i++) acc[i] = acc[i] + c * (kn[i] + s * ks[i]);
      }
#pragma unroll
      for (int p = 0; p < P; p++) {
        const float dure = un[2 * p] - s * us[2 * p];
        const float duim = un[2 * p + 1] - s * us[2 * p + 1];
        const float dvre = vn[2 * p] - s * vs[2 * p];
        const float dvim = vn[2 * p + 1] - s * vs[2 * p + 1];
        const float sure = un[2 * p] + s * us[2 * p];
        const float suim = un[2 * p + 1] + s * us[2 * p + 1];
        const float svre = vn[2 * p] + s * vs[2 * p];
        const float svim = vn[2 * p + 1] + s * vs[2 * p + 1];
        acc[2 * p] = acc[2 * p] - e * dvre - mm * suim;
        acc[2 * p + 1] = acc[2 * p + 1] - e * dvim + mm * sure;
        acc[2 * P + 2 * p] = acc[2 * P + 2 * p] + e * dure - mm * svim;
        acc[2 * P + 2 * p + 1] = acc[2 * P + 2 * p + 1] + e * duim + mm * svre;
      }
    }
  } else if (live) {
    const float* qe = M.tab + M.tabo[TAB_legendre_qe] + w;
    const float* qc = M.tab + M.tabo[TAB_legendre_qc] + w;
    const float* qm = M.tab + M.tabo[TAB_legendre_qm] + w;
    for (int l = wv; l < half; l += 4) {
      const int so = nlat - 1 - l;
      const float e = qe[l * ncsp], c = qc[l * ncsp], mm = qm[l * ncsp];
      float un[SV(P)], us[SV(P)], vn[SV(P)], vs[SV(P)], sn[SV(P)], ss[SV(P)];
      read_vec<P>(row(0, l), un);
      read_vec<P>(row(0, so), us);
      read_vec<P>(row(1, l), vn);
      read_vec<P>(row(1, so), vs);
      read_vec<P>(row(2, l), sn);
      read_vec<P>(row(2, so), ss);
#pragma unroll
      for (int p = 0; p < P; p++) {
        acc[2 * p] = acc[2 * p] + e * (vn[2 * p] - s * vs[2 * p]) + c * (sn[2 * p] + s * ss[2 * p]) +
                     mm * (un[2 * p + 1] + s * us[2 * p + 1]);
        acc[2 * p + 1] = acc[2 * p + 1] + e * (vn[2 * p + 1] - s * vs[2 * p + 1]) +
                         c * (sn[2 * p + 1] + s * ss[2 * p + 1]) - mm * (un[2 * p] + s * us[2 * p]);
      }
    }
  }
  // waves 1..3 hand their partial sums to wave 0, which adds them in wave
  // order (a fixed order: bit-reproducible)
  if (wv > 0 && live) {
    float* r = red + ((wv - 1) * 64 + k) * kRedStride;
#pragma unroll
    for (int i = 0; i < 4 * P; i++)
      if (i < nacc) r[i] = acc[i];
  }
  __syncthreads();
  if (wv != 0 || !live) return;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const float* r = red + (j * 64 + k) * kRedStride;
#pragma unroll
    for (int i = 0; i < 4 * P; i++)
      if (i < nacc) acc[i] += r[i];
  }
  float* o0 = M.arena + M.slot[job.dst0] + 2 * w;
#pragma unroll
  for (int p = 0; p < P; p++) {
    o0[p * nrsp] = acc[2 * p];
    o0[p * nrsp + 1] = acc[2 * p + 1];
  }
  if (job.kind == LEG_UV2DV) {
    float* o1 = M.arena + M.slot[job.dst1] + 2 * w;
#pragma unroll
    for (int p = 0; p < P; p++) {
      o1[p * nrsp] = acc[2 * P + 2 * p];
      o1[p * nrsp + 1] = acc[2 * P + 2 * p + 1];
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
  // qu and qv tiles; after the sums, the partial sums of lane groups 1..7
  // ((8 - 1) x 32 lanes x (4 x 10 + 1) floats at T21 fits)
  __shared__ float qab[2 * kMaxLat * kStride];
  __shared__ __align__(16) float sa[kMaxK * SV(kMaxLev)];
  __shared__ __align__(16) float sb[kMaxK * SV(kMaxLev)];
  static_assert(7 * 32 * (4 * kMaxLev + 1) <= 2 * kMaxLat * kStride, "partial sums exceed the table tiles");
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const InvJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) leg_inv<PP>(M, job, m, qab, sa, sb)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(256) void k_dft_inv(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                             InvJobs jobs) {
  __shared__ __align__(16) float f0[(kMaxFC / 2) * SV(kMaxLev)];
  __shared__ __align__(16) float f1[(kMaxFC / 2) * SV(kMaxLev)];
  __shared__ float ctw[kMaxLon], stw[kMaxLon];
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int lat = tile % M.nlat;
  const InvJob job = jobs.j[job_of(jobs, tile / M.nlat)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) dft_inv<PP>(M, job, lat, f0, f1, ctw, stw)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(256) void k_dft_fwd(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                             FwdJobs jobs) {
  __shared__ __align__(16) float row[kMaxLon * 12];
  __shared__ float ctw[kMaxLon], stw[kMaxLon];
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int lat = tile % M.nlat;
  const FwdJob job = jobs.j[job_of(jobs, tile / M.nlat)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) dft_fwd<PP>(M, job, lat, row, ctw, stw)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}

extern "C" __global__ __launch_bounds__(kLegFwdThreads) void k_leg_fwd(const DevMember* __restrict__ mems, LaunchPlan plan,
                                                            LegJobs jobs) {
  __shared__ __align__(16) float fin[3 * kMaxLat * SV(kMaxLev)];
  __shared__ float red[3 * 64 * kRedStride];
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const LegJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) leg_fwd<PP>(M, job, m, fin, red)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}
