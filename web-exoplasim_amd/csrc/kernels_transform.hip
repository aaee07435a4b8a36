// Spectral transforms of the ExoPlaSim step on gfx950, plain FP32 form.
//
//  k_sp2gp  sp2fc/sp2fcdmu/dv2uv fused with the zonal synthesis fc2gp: one
//           64-lane wavefront per (member, latitude row, job plane). Lane m
//           sums the row's Fourier coefficient m over n (the reference's
//           legendre/wgsl/inverse.wgsl:6-27 and dv2uv.wgsl:9-43), the
//           coefficients stay in LDS, and the lanes then synthesise the
//           NLON grid points of the row (fourier/wgsl/inverse.wgsl:4-18),
//           so the Fourier intermediate never goes to HBM.
//  k_gp2fc  zonal analysis gp2fc of one grid row into its Fourier row
//           (fourier/wgsl/forward.wgsl:4-18), the row staged in LDS.
//  k_fc2sp  fc2sp/uv2dv/mktend/qtend: lane = spectral mode w, summing the
//           northern/southern row pairs (legendre/wgsl/forward.wgsl:7-38,
//           uv2dv.wgsl:11-63, qtend.wgsl:9-46).
//
// Every kernel walks a LaunchPlan so one launch serves a mixed batch.
#include "transforms.h"

namespace {

__device__ __forceinline__ int mode_first(int m, int ntp1) { return m * (2 * ntp1 - m + 1) / 2; }

// (m, n) of complex mode w in m-major order.
__device__ __forceinline__ void mode_of(int w, int ntp1, int* m_out, int* n_out) {
  int m = 0, start = 0;
  while (w >= start + ntp1 - m) {
    start += ntp1 - m;
    m++;
  }
  *m_out = m;
  *n_out = m + w - start;
}

}  // namespace

extern "C" __global__ __launch_bounds__(64) void k_sp2gp(const DevMember* __restrict__ mems,
                                                          LaunchPlan plan, InvJobs jobs) {
  int tile;
  const int mi = plan_find(plan, blockIdx.x, &tile);
  const DevMember& M = mems[mi];
  const int nlat = M.nlat, nlon = M.nlon, L = M.nlev, ntp1 = M.ntru + 1;
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc, nhor = M.nhor;
  const int lat = tile % nlat;
  int jp = tile / nlat;
  int j = 0;
  for (; j < jobs.n; j++) {
    const int np = jobs.j[j].per_level ? L : 1;
    if (jp < np) break;
    jp -= np;
  }
  const InvJob job = jobs.j[j];
  const int plane = jp;
  __shared__ float f0[128];
  __shared__ float f1[128];
  const int m = threadIdx.x;
  float* arena = M.arena;
  if (job.kind == INV_DV2UV) {
    const float* d = arena + M.slot[job.src0] + (size_t)plane * nrsp;
    const float* z = arena + M.slot[job.src1] + (size_t)plane * nrsp;
    const float* qu = M.tab + M.tabo[TAB_legendre_qu] + (size_t)lat * ncsp;
    const float* qv = M.tab + M.tabo[TAB_legendre_qv] + (size_t)lat * ncsp;
    const float plavor = job.plavor ? M.k[CON_plavor] : 0.0f;
    if (m < ntp1) {
      const int first = mode_first(m, ntp1);
      float ure = 0, uim = 0, vre = 0, vim = 0;
      for (int w = first; w < first + ntp1 - m; w++) {
        const float u = qu[w], v = qv[w];
        const float zre = w == 1 ? z[2 * w] - plavor : z[2 * w];
        const float zim = z[2 * w + 1];
        const float dre = d[2 * w], dim = d[2 * w + 1];
        ure = ure + v * zre + u * dim;
        uim = uim + v * zim - u * dre;
        vre = vre + u * zim - v * dre;
        vim = vim - u * zre - v * dim;
      }
      f0[2 * m] = ure;
      f0[2 * m + 1] = uim;
      f1[2 * m] = vre;
      f1[2 * m + 1] = vim;
    }
  } else {
    const float* s = arena + M.slot[job.src0] + (size_t)plane * nrsp;
    const float* q = M.tab + M.tabo[job.table] + (size_t)lat * ncsp;
    if (m < ntp1) {
      const int first = mode_first(m, ntp1);
      float re = 0, im = 0;
      for (int w = first; w < first + ntp1 - m; w++) {
        const float qq = q[w];
        re += qq * s[2 * w];
        im += qq * s[2 * w + 1];
      }
      f0[2 * m] = re;
      f0[2 * m + 1] = im;
      if (job.kind == INV_DLAMBDA) {
        // fourier/wgsl/dlambda.wgsl:5-17: (re, im) * i m
        f1[2 * m] = -im * (float)m;
        f1[2 * m + 1] = re * (float)m;
      }
    }
  }
  __syncthreads();
  const float* inv = M.tab + M.tabo[TAB_fourier_inverse];
  const bool two = job.kind != INV_PLAIN;
  float* g0 = arena + M.slot[job.dst0] + (size_t)plane * nhor + (size_t)lat * nlon;
  float* g1 = two ? arena + M.slot[job.dst1] + (size_t)plane * nhor + (size_t)lat * nlon : nullptr;
  for (int x = threadIdx.x; x < nlon; x += 64) {
    float a = 0, b = 0;
    for (int k = 0; k < nfc; k++) {
      const float t = inv[(size_t)k * nlon + x];
      a += t * f0[k];
      if (two) b += t * f1[k];
    }
    g0[x] = a;
    if (two) g1[x] = b;
  }
}

extern "C" __global__ __launch_bounds__(64) void k_gp2fc(const DevMember* __restrict__ mems,
                                                          LaunchPlan plan, FwdJobs jobs) {
  int tile;
  const int mi = plan_find(plan, blockIdx.x, &tile);
  const DevMember& M = mems[mi];
  const int nlat = M.nlat, nlon = M.nlon, L = M.nlev, nfc = M.nfc, nhor = M.nhor;
  const int lat = tile % nlat;
  int jp = tile / nlat;
  int j = 0;
  for (; j < jobs.n; j++) {
    const int np = jobs.j[j].per_level ? L : 1;
    if (jp < np) break;
    jp -= np;
  }
  const FwdJob job = jobs.j[j];
  const int plane = jp;
  __shared__ float row[192];
  const float* g = M.arena + M.slot[job.src] + (size_t)plane * nhor + (size_t)lat * nlon;
  for (int x = threadIdx.x; x < nlon; x += 64) row[x] = g[x];
  __syncthreads();
  const float* fwd = M.tab + M.tabo[TAB_fourier_forward];
  float* f = M.arena + M.slot[job.dst] + (size_t)plane * nlat * nfc + (size_t)lat * nfc;
  for (int k = threadIdx.x; k < nfc; k += 64) {
    float sum = 0;
    for (int x = 0; x < nlon; x++) sum += fwd[(size_t)x * nfc + k] * row[x];
    f[k] = sum;
  }
}

extern "C" __global__ __launch_bounds__(64) void k_fc2sp(const DevMember* __restrict__ mems,
                                                          LaunchPlan plan, LegJobs jobs) {
  int tile;
  const int mi = plan_find(plan, blockIdx.x, &tile);
  const DevMember& M = mems[mi];
  const int nlat = M.nlat, L = M.nlev, ntp1 = M.ntru + 1;
  const int ncsp = M.ncsp, nrsp = M.nrsp, nfc = M.nfc;
  const int wtiles = (ncsp + 63) / 64;
  const int wt = tile % wtiles;
  int jp = tile / wtiles;
  int j = 0;
  for (; j < jobs.n; j++) {
    const int np = jobs.j[j].per_level ? L : 1;
    if (jp < np) break;
    jp -= np;
  }
  const LegJob job = jobs.j[j];
  const int plane = jp;
  const int w = wt * 64 + threadIdx.x;
  if (w >= ncsp) return;
  int m, n;
  mode_of(w, ntp1, &m, &n);
  const float s = ((m + n) % 2 == 0) ? 1.0f : -1.0f;
  const size_t fplane = (size_t)plane * nlat * nfc + 2 * m;
  float* arena = M.arena;
  const size_t out = (size_t)plane * nrsp + 2 * w;
  if (job.kind == LEG_PLAIN) {
    const float* f = arena + M.slot[job.src] + fplane;
    const float* q = M.tab + M.tabo[TAB_legendre_qc] + w;
    float re = 0, im = 0;
    for (int l = 0; l < nlat / 2; l++) {
      const float qq = q[(size_t)l * ncsp];
      const float* nn = f + (size_t)l * nfc;
      const float* ss = f + (size_t)(nlat - 1 - l) * nfc;
      re += qq * (nn[0] + s * ss[0]);
      im += qq * (nn[1] + s * ss[1]);
    }
    arena[M.slot[job.dst0] + out] = re;
    arena[M.slot[job.dst0] + out + 1] = im;
  } else if (job.kind == LEG_UV2DV) {
    const float* U = arena + M.slot[job.u] + fplane;
    const float* V = arena + M.slot[job.v] + fplane;
    const bool kin = job.ke >= 0;
    const float* K = kin ? arena + M.slot[job.ke] + fplane : nullptr;
    const float* qe = M.tab + M.tabo[TAB_legendre_qe] + w;
    const float* qm = M.tab + M.tabo[TAB_legendre_qm] + w;
    const float* qq = M.tab + M.tabo[TAB_legendre_qq] + w;
    float dre = 0, dim = 0, zre = 0, zim = 0;
    for (int l = 0; l < nlat / 2; l++) {
      const size_t no = (size_t)l * nfc, so = (size_t)(nlat - 1 - l) * nfc;
      const float e = qe[(size_t)l * ncsp], mm = qm[(size_t)l * ncsp];
      const float dure = U[no] - s * U[so];
      const float duim = U[no + 1] - s * U[so + 1];
      const float dvre = V[no] - s * V[so];
      const float dvim = V[no + 1] - s * V[so + 1];
      const float sure = U[no] + s * U[so];
      const float suim = U[no + 1] + s * U[so + 1];
      const float svre = V[no] + s * V[so];
      const float svim = V[no + 1] + s * V[so + 1];
      if (kin) {
        const float c = qq[(size_t)l * ncsp];
        dre = dre + c * (K[no] + s * K[so]);
        dim = dim + c * (K[no + 1] + s * K[so + 1]);
      }
      dre = dre - e * dvre - mm * suim;
      dim = dim - e * dvim + mm * sure;
      zre = zre + e * dure - mm * svim;
      zim = zim + e * duim + mm * svre;
    }
    arena[M.slot[job.dst0] + out] = dre;
    arena[M.slot[job.dst0] + out + 1] = dim;
    arena[M.slot[job.dst1] + out] = zre;
    arena[M.slot[job.dst1] + out + 1] = zim;
  } else {
    const float* U = arena + M.slot[job.u] + fplane;
    const float* V = arena + M.slot[job.v] + fplane;
    const float* S = arena + M.slot[job.src] + fplane;
    const float* qe = M.tab + M.tabo[TAB_legendre_qe] + w;
    const float* qc = M.tab + M.tabo[TAB_legendre_qc] + w;
    const float* qm = M.tab + M.tabo[TAB_legendre_qm] + w;
    float re = 0, im = 0;
    for (int l = 0; l < nlat / 2; l++) {
      const size_t no = (size_t)l * nfc, so = (size_t)(nlat - 1 - l) * nfc;
      const float e = qe[(size_t)l * ncsp], c = qc[(size_t)l * ncsp], mm = qm[(size_t)l * ncsp];
      re = re + e * (V[no] - s * V[so]) + c * (S[no] + s * S[so]) + mm * (U[no + 1] + s * U[so + 1]);
      im = im + e * (V[no + 1] - s * V[so + 1]) + c * (S[no + 1] + s * S[so + 1]) - mm * (U[no] + s * U[so]);
    }
    arena[M.slot[job.dst0] + out] = re;
    arena[M.slot[job.dst0] + out + 1] = im;
  }
}
