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
constexpr int kLegFwdThreads = 64;  // k_leg_fwd: one wave, lane = mode (K <= kMaxK)

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

template <int P>
__device__ void leg_inv_plain(const DevMember& M, const InvJob& job, int m, float* qa, float* sa) {
  const int nlat = M.nlat, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc;
  const float* q = M.tab + M.tabo[job.table];
  const float* S = M.arena + M.slot[job.src0];
  for (int i = threadIdx.x; i < nlat * K; i += kThreads) {
    const int lat = i / K, k = i % K;
    qa[lat * kStride + k] = q[lat * ncsp + w0 + k];
  }
  for (int i = threadIdx.x; i < P * 2 * K; i += kThreads) {
    const int p = i / (2 * K), r = i % (2 * K);
    sa[(r >> 1) * SV(P) + 2 * p + (r & 1)] = S[p * nrsp + 2 * w0 + r];
  }
  __syncthreads();
  float* F = M.arena + M.slot[job.four0];
  for (int lat = threadIdx.x; lat < nlat; lat += kThreads) {
    float acc[SV(P)];
#pragma unroll
    for (int i = 0; i < SV(P); i++) acc[i] = 0.0f;
    for (int k = 0; k < K; k++) {
      const float qq = qa[lat * kStride + k];
      float s[SV(P)];
      read_vec<P>(sa + k * SV(P), s);
#pragma unroll
      for (int i = 0; i < 2 * P; i++) acc[i] += qq * s[i];
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
      float* f = F + (p * nlat + lat) * nfc + 2 * m;
      f[0] = acc[2 * p];
      f[1] = acc[2 * p + 1];
    }
  }
}

template <int P>
__device__ void leg_inv_winds(const DevMember& M, const InvJob& job, int m, float* qa, float* qb, float* sa,
                              float* sb) {
  const int nlat = M.nlat, ntp1 = M.ntru + 1, K = ntp1 - m, w0 = mode_first(m, ntp1);
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc;
  const float* qu = M.tab + M.tabo[TAB_legendre_qu];
  const float* qv = M.tab + M.tabo[TAB_legendre_qv];
  const float* D = M.arena + M.slot[job.src0];
  const float* Z = M.arena + M.slot[job.src1];
  const float plavor = job.plavor ? M.k[CON_plavor] : 0.0f;
  for (int i = threadIdx.x; i < nlat * K; i += kThreads) {
    const int lat = i / K, k = i % K;
    qa[lat * kStride + k] = qu[lat * ncsp + w0 + k];
    qb[lat * kStride + k] = qv[lat * ncsp + w0 + k];
  }
  for (int i = threadIdx.x; i < P * 2 * K; i += kThreads) {
    const int p = i / (2 * K), r = i % (2 * K);
    const int at = (r >> 1) * SV(P) + 2 * p + (r & 1);
    sa[at] = D[p * nrsp + 2 * w0 + r];
    // zeta without the planetary vorticity in Re mode (0, 1)
    sb[at] = Z[p * nrsp + 2 * w0 + r] - ((w0 + (r >> 1) == 1 && (r & 1) == 0) ? plavor : 0.0f);
  }
  __syncthreads();
  float* U = M.arena + M.slot[job.four0];
  float* V = M.arena + M.slot[job.four1];
  for (int lat = threadIdx.x; lat < nlat; lat += kThreads) {
    float ure[P], uim[P], vre[P], vim[P];
#pragma unroll
    for (int p = 0; p < P; p++) ure[p] = uim[p] = vre[p] = vim[p] = 0.0f;
    for (int k = 0; k < K; k++) {
      const float u = qa[lat * kStride + k], v = qb[lat * kStride + k];
      float d[SV(P)], z[SV(P)];
      read_vec<P>(sa + k * SV(P), d);
      read_vec<P>(sb + k * SV(P), z);
#pragma unroll
      for (int p = 0; p < P; p++) {
        const float dre = d[2 * p], dim = d[2 * p + 1];
        const float zre = z[2 * p], zim = z[2 * p + 1];
        ure[p] = ure[p] + v * zre + u * dim;
        uim[p] = uim[p] + v * zim - u * dre;
        vre[p] = vre[p] + u * zim - v * dre;
        vim[p] = vim[p] - u * zre - v * dim;
      }
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
      float* fu = U + (p * nlat + lat) * nfc + 2 * m;
      float* fv = V + (p * nlat + lat) * nfc + 2 * m;
      fu[0] = ure[p];
      fu[1] = uim[p];
      fv[0] = vre[p];
      fv[1] = vim[p];
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

template <int P>
__device__ void leg_fwd(const DevMember& M, const LegJob& job, int m, float* fin) {
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
  // lane k owns mode w0 + k for every level; its table column is read
  // straight from memory (consecutive lanes, consecutive modes)
  const int k = threadIdx.x;
  if (k >= K) return;
  const float s = (k & 1) ? -1.0f : 1.0f;  // parity of m + n = 2m + k
  const int w = w0 + k;
  auto row = [&](int f, int lat) { return fin + (f * nlat + lat) * SV(P); };
  if (job.kind == LEG_PLAIN) {
    const float* q = M.tab + M.tabo[TAB_legendre_qc] + w;
    float acc[SV(P)];
#pragma unroll
    for (int i = 0; i < SV(P); i++) acc[i] = 0.0f;
    for (int l = 0; l < half; l++) {
      const float qq = q[l * ncsp];
      float nv[SV(P)], sv[SV(P)];
      read_vec<P>(row(0, l), nv);
      read_vec<P>(row(0, nlat - 1 - l), sv);
#pragma unroll
      for (int i = 0; i < 2 * P; i++) acc[i] += qq * (nv[i] + s * sv[i]);
    }
    float* out = M.arena + M.slot[job.dst0] + 2 * w;
#pragma unroll
    for (int p = 0; p < P; p++) {
      out[p * nrsp] = acc[2 * p];
      out[p * nrsp + 1] = acc[2 * p + 1];
    }
  } else if (job.kind == LEG_UV2DV) {
    const bool kin = job.ke >= 0;
    const float* qe = M.tab + M.tabo[TAB_legendre_qe] + w;
    const float* qm = M.tab + M.tabo[TAB_legendre_qm] + w;
    const float* qq = M.tab + M.tabo[TAB_legendre_qq] + w;
    float dre[P], dim[P], zre[P], zim[P];
#pragma unroll
    for (int p = 0; p < P; p++) dre[p] = dim[p] = zre[p] = zim[p] = 0.0f;
    for (int l = 0; l < half; l++) {
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
        for (int p = 0; p < P; p++) {
          dre[p] = dre[p] + c * (kn[2 * p] + s * ks[2 * p]);
          dim[p] = dim[p] + c * (kn[2 * p + 1] + s * ks[2 * p + 1]);
        }
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
        dre[p] = dre[p] - e * dvre - mm * suim;
        dim[p] = dim[p] - e * dvim + mm * sure;
        zre[p] = zre[p] + e * dure - mm * svim;
        zim[p] = zim[p] + e * duim + mm * svre;
      }
    }
    float* o0 = M.arena + M.slot[job.dst0] + 2 * w;
    float* o1 = M.arena + M.slot[job.dst1] + 2 * w;
#pragma unroll
    for (int p = 0; p < P; p++) {
      o0[p * nrsp] = dre[p];
      o0[p * nrsp + 1] = dim[p];
      o1[p * nrsp] = zre[p];
      o1[p * nrsp + 1] = zim[p];
    }
  } else {
    const float* qe = M.tab + M.tabo[TAB_legendre_qe] + w;
    const float* qc = M.tab + M.tabo[TAB_legendre_qc] + w;
    const float* qm = M.tab + M.tabo[TAB_legendre_qm] + w;
    float re[P], im[P];
#pragma unroll
    for (int p = 0; p < P; p++) re[p] = im[p] = 0.0f;
    for (int l = 0; l < half; l++) {
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
        re[p] = re[p] + e * (vn[2 * p] - s * vs[2 * p]) + c * (sn[2 * p] + s * ss[2 * p]) +
                mm * (un[2 * p + 1] + s * us[2 * p + 1]);
        im[p] = im[p] + e * (vn[2 * p + 1] - s * vs[2 * p + 1]) + c * (sn[2 * p + 1] + s * ss[2 * p + 1]) -
                mm * (un[2 * p] + s * us[2 * p]);
      }
    }
    float* out = M.arena + M.slot[job.dst0] + 2 * w;
#pragma unroll
    for (int p = 0; p < P; p++) {
      out[p * nrsp] = re[p];
      out[p * nrsp + 1] = im[p];
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
  __shared__ float qa[kMaxLat * kStride];
  __shared__ float qb[kMaxLat * kStride];
  __shared__ __align__(16) float sa[kMaxK * SV(kMaxLev)];
  __shared__ __align__(16) float sb[kMaxK * SV(kMaxLev)];
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const InvJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP)                                                    \
  if (job.kind == INV_DV2UV)                                        \
    leg_inv_winds<PP>(M, job, m, qa, qb, sa, sb);                   \
  else                                                              \
    leg_inv_plain<PP>(M, job, m, qa, sa)
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
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int ntp1 = M.ntru + 1;
  const int m = tile % ntp1;
  const LegJob job = jobs.j[job_of(jobs, tile / ntp1)];
  const int P = job.per_level ? M.nlev : 1;
#define CALL(PP) leg_fwd<PP>(M, job, m, fin)
  EXO_PLANES_DISPATCH(P, CALL);
#undef CALL
}
