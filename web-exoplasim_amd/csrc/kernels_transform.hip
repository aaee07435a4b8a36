// Zonal DFTs of the ExoPlaSim step on gfx950, plain FP32 form (the
// Legendre half of each transform is in kernels_legendre.hip).
//
// The reference runs one WGSL invocation per output value and rereads the
// DFT tables for every plane (fourier/wgsl/*.wgsl). Here each workgroup
// owns one latitude row of one job and every level of it: the Fourier
// coefficients come in as per-wavenumber level vectors in LDS, the
// twiddles are built once per row, and each lane keeps one accumulator
// per level in registers.
//
//  k_dft_inv  fc2gp (+ d/dlambda) of one latitude row, every level
//             (fourier/wgsl/inverse.wgsl:4-18, dlambda.wgsl:5-17).
//  k_dft_fwd  gp2fc of one latitude row, every level
//             (fourier/wgsl/forward.wgsl:4-18).
//
// Every kernel walks a LaunchPlan so one launch serves a mixed batch.
#include "transforms.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxLev = 10, kMaxFC = 128, kMaxLon = 192;

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

// The zonal DFT matrices hold only NLON distinct cos/sin values:
// cos(m lambda_x) = cos(2 pi ((m x) mod NLON) / NLON). The twiddles are
// built in LDS from the FP32 inverse table (its m = 1 rows hold
// 2 cos(2 pi x / NLON) and -2 sin(2 pi x / NLON), exact doubles of the
// FP32 cos/sin), and every lane walks its (m x) mod NLON index.
__device__ void load_twiddles(const DevMember& M, float* ctw, float* stw) {
  const int nlon = M.nlon;
  const float* inv = M.ltab + M.tabo[TAB_fourier_inverse];
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

