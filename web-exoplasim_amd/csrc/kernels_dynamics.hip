// Dynamical core of the ExoPlaSim step on gfx950: gridpointa's column
// work, the semi-implicit spectrala, spectrald's implicit diffusion and
// time filter, mkdheat's grid-point stages, and gridpointd's physics frame.
// spectrala, whose levels couple, runs one thread per coefficient and level
// and passes level vectors through LDS; calcgp and mkdheat's stages one
// thread per column or point; the rest one thread per point or coefficient
// and level. The level count is a template parameter.
#include "device.h"

namespace {

// Waves of a level-coupled workgroup: 64 coefficients of an L10 member,
// 128 of an L5 one (wave w: level w % NL of tile w / NL).
constexpr int kLevelWaves = 10;

typedef float LevelRows[kLevelWaves][64];

template <int NL>
__device__ void calcgp_column(const DevMember& M, int h) {
  // dynamics/wgsl/calcgp.wgsl:14-147
  const int NHOR = M.nhor;
  float* A = dev_arena(M);
  const float* T = dev_tab(M);
  auto lev = [&](int slot, int l) -> float& { return A[M.slot[slot] + l * NHOR + h]; };
  auto tab = [&](int t, int i) { return T[M.tabo[t] + i]; };
  const float PSURF = M.k[CON_psurf], WW = M.k[CON_ww], CV = M.k[CON_cv];
  const float AKAP = M.k[CON_akap], RDBRV = M.k[CON_rdbrv], ADV = M.k[CON_adv];
  const float rcsq = tab(TAB_rcsq, h);
  const float ps = expf(A[M.slot[SLOT_gp] + h]);
  A[M.slot[SLOT_gp] + h] = ps;
  const float gpm = A[M.slot[SLOT_gpmt] + h];
  const float gpj = A[M.slot[SLOT_gpj] + h];
  float gu[NL], gv[NL], gt[NL], gd[NL], gz[NL], gq[NL], zvgpg[NL], ztv1[NL], ztv2[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    gu[l] = lev(SLOT_gu, l);
    gv[l] = lev(SLOT_gv, l);
    gt[l] = lev(SLOT_gt, l);
    gd[l] = lev(SLOT_gd, l);
    gz[l] = lev(SLOT_gz, l);
    gq[l] = lev(SLOT_gq, l);
    zvgpg[l] = rcsq * (gu[l] * gpm + gv[l] * gpj);
    const float zq = fmaxf(gq[l] / ps, 0.0f);
    const float t0 = tab(TAB_t0, l);
    const float tvirt = (gt[l] + t0) * (1.0f + (1.0f / RDBRV - 1.0f) * zq);
    ztv1[l] = tvirt - t0;
    ztv2[l] = tvirt / (1.0f + ADV * zq) - t0;
  }
  float zsdotp[NL];
  float zsumd = tab(TAB_dsigma, 0) * gd[0];
  float gvp = tab(TAB_dsigma, 0) * zvgpg[0];
  zsdotp[0] = zsumd + gvp;
#pragma unroll
  for (int l = 1; l < NL - 1; l++) {
    zsumd = zsumd + tab(TAB_dsigma, l) * gd[l];
    gvp = gvp + tab(TAB_dsigma, l) * zvgpg[l];
    zsdotp[l] = zsumd + gvp;
  }
  zsumd = zsumd + tab(TAB_dsigma, NL - 1) * gd[NL - 1];
  gvp = gvp + tab(TAB_dsigma, NL - 1) * zvgpg[NL - 1];
  float gtd[NL], gqm[NL], gud[NL], gvd[NL];
#pragma unroll
  for (int l = 0; l < NL - 1; l++) {
    zsdotp[l] = tab(TAB_sigmah, l) * (zsumd + gvp) - zsdotp[l];
    gtd[l] = zsdotp[l] * (gt[l + 1] - gt[l]);
    gqm[l] = zsdotp[l] * (gq[l + 1] + gq[l]);
    gud[l] = zsdotp[l] * (gu[l + 1] - gu[l]);
    gvd[l] = zsdotp[l] * (gv[l + 1] - gv[l]);
  }
  float gtn[NL], gqn[NL], guz[NL], gvz[NL], dw[NL];
  zsumd = zvgpg[0] * tab(TAB_dsigma, 0);
  const float rdsig0 = tab(TAB_rdsig, 0);
  gtn[0] = gt[0] * gd[0] - AKAP * ztv2[0] * gd[0] -
           rdsig0 * (gtd[0] + tab(TAB_t01s2, 0) * (tab(TAB_sigmah, 0) * gvp - zsumd));
  gqn[0] = -rdsig0 * gqm[0];
  guz[0] = -gu[0] * gz[0] - gpj * ztv1[0] - rdsig0 * gvd[0];
  gvz[0] = gv[0] * gz[0] - gpm * ztv1[0] - rdsig0 * gud[0];
  dw[0] = gd[0] * ps * PSURF * WW;
#pragma unroll
  for (int l = 1; l < NL; l++) {
    const int column = l * NL;
    float ztpta = tab(TAB_c, column) * zvgpg[0];
    float ztptb = tab(TAB_c, column) * (zvgpg[0] + gd[0]);
    for (int j = 1; j <= l; j++) {
      ztpta = ztpta + tab(TAB_c, column + j) * zvgpg[j];
      ztptb = ztptb + tab(TAB_c, column + j) * (zvgpg[j] + gd[j]);
    }
    const float rdsig = tab(TAB_rdsig, l);
    const float t01s2 = tab(TAB_t01s2, l), t01s2m = tab(TAB_t01s2, l - 1);
    const float sigmah = tab(TAB_sigmah, l), sigmahm = tab(TAB_sigmah, l - 1);
    const float tkp = tab(TAB_tkp, l);
    if (l < NL - 1) {
      zsumd = zsumd + zvgpg[l] * tab(TAB_dsigma, l);
      gtn[l] = gt[l] * gd[l] + AKAP * ztv2[l] * (zvgpg[l] - ztptb) + tkp * (zvgpg[l] - ztpta) -
               rdsig * (gtd[l] + gtd[l - 1] + gvp * (t01s2 * sigmah + t01s2m * sigmahm) -
                        zsumd * (t01s2m + t01s2) + zvgpg[l] * tab(TAB_dsigma, l) * t01s2m);
      gqn[l] = -rdsig * (gqm[l] - gqm[l - 1]);
      guz[l] = -gu[l] * gz[l] - gpj * ztv1[l] - rdsig * (gvd[l] + gvd[l - 1]);
      gvz[l] = gv[l] * gz[l] - gpm * ztv1[l] - rdsig * (gud[l] + gud[l - 1]);
    } else {
      gtn[l] = gt[l] * gd[l] + AKAP * ztv2[l] * (zvgpg[l] - ztptb) + tkp * (zvgpg[l] - ztpta) -
               rdsig * (gtd[NL - 2] + t01s2m * (sigmahm * gvp - zsumd));
      gqn[l] = rdsig * gqm[NL - 2];
      guz[l] = -gu[l] * gz[l] - gpj * ztv1[l] - rdsig * gvd[NL - 2];
      gvz[l] = gv[l] * gz[l] - gpm * ztv1[l] - rdsig * gud[NL - 2];
    }
    dw[l] = (zvgpg[l] - ztptb) * ps * PSURF * WW;
  }
  const float scale = sqrtf(rcsq);
#pragma unroll
  for (int l = 0; l < NL; l++) {
    float gsum = 0;
    for (int k = 0; k < NL; k++) gsum = gsum + tab(TAB_g, l * NL + k) * (ztv1[k] - gt[k]);
    const float gphi = gsum * 2.0f / rcsq;
    lev(SLOT_gtn, l) = gtn[l];
    lev(SLOT_gqn, l) = gqn[l];
    lev(SLOT_guz, l) = guz[l];
    lev(SLOT_gvz, l) = gvz[l];
    lev(SLOT_dw, l) = dw[l];
    lev(SLOT_gut, l) = gu[l] * gt[l];
    lev(SLOT_gvt, l) = gv[l] * gt[l];
    lev(SLOT_gke, l) = gu[l] * gu[l] + gv[l] * gv[l] + gphi;
    lev(SLOT_guq, l) = gu[l] * gq[l];
    lev(SLOT_gvq, l) = gv[l] * gq[l];
    lev(SLOT_du0, l) = CV * gu[l] * scale;
    lev(SLOT_dv0, l) = CV * gv[l] * scale;
  }
  A[M.slot[SLOT_gvpp] + h] = gvp;
  A[M.slot[SLOT_dp0] + h] = PSURF * ps;
}

// dynamics/wgsl/spectrala.wgsl:15-109, one thread per coefficient and
// level. The semi-implicit solve couples a coefficient's levels three
// times (the G products, the bm1 solve, the tau products); the level
// vectors pass through LDS between them. Every sum runs over the levels in
// the same order as a one-thread-per-coefficient pass.

template <int NL>
__device__ void spectrala_levels(const DevMember& M, int tile, LevelRows* lds) {
  constexpr int kSub = kLevelWaves / NL;
  static_assert(kSub * NL == kLevelWaves, "levels must divide the workgroup's waves");
  LevelRows &sh_stt = lds[0], &sh_atm = lds[1], &sh_zz = lds[2], &sh_sdt = lds[3];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int l = wave % NL, row0 = wave - l;
  const int s = (tile * kSub + wave / NL) * 64 + lane;
  const int NRSP = M.nrsp;
  const bool live = s < NRSP;
  const int sc = live ? s : 0;
  float* A = dev_arena(M);
  const float* T = dev_tab(M);
  auto lev = [&](int slot) -> float& { return A[M.slot[slot] + l * NRSP + sc]; };
  auto tab = [&](int t, int i) { return T[M.tabo[t] + i]; };
  const float DELT = M.k[CON_delt], DELT2 = M.k[CON_delt2];
  const float PNU = M.k[CON_pnu], PNU21 = M.k[CON_pnu21];
  const bool FILTER = M.k[CON_filter] != 0.0f;
  const float adm = lev(SLOT_sdm), azm = lev(SLOT_szm), atm = lev(SLOT_stm), aqm = lev(SLOT_sqm);
  const float adt = lev(SLOT_sdt), stt0 = lev(SLOT_stt);
  const float apm = A[M.slot[SLOT_spm] + sc];
  const float spt0 = A[M.slot[SLOT_spt] + sc];
  const float spp = A[M.slot[SLOT_spp] + sc];
  const float sop = tab(TAB_sop, sc);
  const int jn = (int)tab(TAB_nindex, sc);
  sh_stt[wave][lane] = stt0;
  sh_atm[wave][lane] = atm;
  __syncthreads();
  float gt = 0, gm = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) {
    gt = gt + tab(TAB_g, l * NL + k) * sh_stt[row0 + k][lane];
    gm = gm + tab(TAB_g, l * NL + k) * sh_atm[row0 + k][lane];
  }
  // this level's term of the bm1 solve
  const float zq = 1.0f / (float)(jn * (jn + 1));
  const float z0 = tab(TAB_t0, l);
  const float zt = gt - z0 * spt0;
  const float zm = gm + z0 * apm;
  const float za = adt * zq + sop;
  const float zb = adm * zq;
  sh_zz[wave][lane] = zb + DELT * (zm + za + DELT * zt);
  __syncthreads();
  float zsum = 0;
  if (jn > 0) {
    const float* bm1 = T + M.tabo[TAB_bm1] + ((jn - 1) * NL + l) * NL;
#pragma unroll
    for (int l2 = 0; l2 < NL; l2++) zsum = zsum + sh_zz[row0 + l2][lane] * bm1[l2];
  }
  const float sdt = s < 2 ? 0.0f : zsum;
  sh_sdt[wave][lane] = sdt;
  __syncthreads();
  float tau = 0;
#pragma unroll
  for (int k = 0; k < NL; k++) tau = tau + tab(TAB_tau, l * NL + k) * sh_sdt[row0 + k][lane];
  const float stt = stt0 - tau;
  if (!live) return;
  if (l == 0) {
    float spt = spt0;
#pragma unroll
    for (int k = 0; k < NL; k++) spt = spt + tab(TAB_dsigma, k) * sh_sdt[row0 + k][lane];
    if (FILTER) A[M.slot[SLOT_spm] + s] = PNU21 * spp + PNU * apm;
    const float sppNew = s < 2 ? 0.0f : apm - DELT2 * spt;
    A[M.slot[SLOT_spp] + s] = sppNew;
    A[M.slot[SLOT_sp] + s] = sppNew;
    A[M.slot[SLOT_spt] + s] = spt;
  }
  if (FILTER) {
    lev(SLOT_sdm) = PNU21 * lev(SLOT_sdp) + PNU * adm;
    lev(SLOT_szm) = PNU21 * lev(SLOT_szp) + PNU * azm;
    lev(SLOT_stm) = PNU21 * lev(SLOT_stp) + PNU * atm;
    lev(SLOT_sqm) = PNU21 * lev(SLOT_sqp) + PNU * aqm;
  }
  const float sdp = 2.0f * sdt - adm;
  const float stp = DELT2 * stt + atm;
  const float szp = DELT2 * lev(SLOT_szt) + azm;
  const float sqp = DELT2 * lev(SLOT_sqt) + aqm;
  lev(SLOT_sdt) = sdt;
  lev(SLOT_stt) = stt;
  lev(SLOT_sdp) = sdp;
  lev(SLOT_stp) = stp;
  lev(SLOT_szp) = szp;
  lev(SLOT_sqp) = sqp;
  lev(SLOT_sd) = sdp;
  lev(SLOT_st) = stp;
  lev(SLOT_sz) = szp;
  lev(SLOT_sq) = sqp;
}

// One thread per coefficient and level: the diffusion and the update
// touch no other level.
__device__ void diffusion_coeff(const DevMember& M, int s, int l) {
  // dynamics/wgsl/spectrald-diffusion.wgsl:12-60
  const int NRSP = M.nrsp;
  float* A = dev_arena(M);
  const float* T = dev_tab(M);
  auto lev = [&](int slot, int l) -> float& { return A[M.slot[slot] + l * NRSP + s]; };
  const float DELT2 = M.k[CON_delt2], PLAVOR = M.k[CON_plavor];
  if (s < 2 && l == 0) A[M.slot[SLOT_spp] + s] = 0.0f;
  {
    const float szp = lev(SLOT_szp, l) + DELT2 * lev(SLOT_szt, l);
    const float stp = lev(SLOT_stp, l) + DELT2 * lev(SLOT_stt, l);
    const float sdp = lev(SLOT_sdp, l) + DELT2 * lev(SLOT_sdt, l);
    const float sqp = lev(SLOT_sqp, l) + DELT2 * lev(SLOT_sqt, l);
    const float sakpp = T[M.tabo[TAB_sakpp] + l * NRSP + s];
    const float damp = T[M.tabo[TAB_damp] + l];
    const float tdisst = T[M.tabo[TAB_tdisst] + l];
    const float tdissq = T[M.tabo[TAB_tdissq] + l];
    const float tdissd = T[M.tabo[TAB_tdissd] + l];
    const float tdissz = T[M.tabo[TAB_tdissz] + l];
    const float tfrc = T[M.tabo[TAB_tfrc] + l];
    const float srp = T[M.tabo[TAB_srp] + l * NRSP + s];
    const float stt = ((-damp - tdisst * sakpp) * stp + damp * srp) / (1.0f + DELT2 * (damp + tdisst * sakpp));
    const float sqt = -tdissq * sakpp * sqp / (1.0f + DELT2 * tdissq * sakpp);
    const float zsdt1 = -tfrc * sdp / (1.0f + DELT2 * tfrc);
    const float zsdt2 = -tdissd * sakpp * sdp / (1.0f + DELT2 * tdissd * sakpp);
    float zszt1 = -tfrc * szp / (1.0f + DELT2 * tfrc);
    float zszt2 = -tdissz * sakpp * szp / (1.0f + DELT2 * tdissz * sakpp);
    float szt = zszt1 + zszt2;
    if (s == 2) {
      zszt1 = zszt1 + PLAVOR * tfrc / (1.0f + DELT2 * tfrc);
      zszt2 = zszt2 + PLAVOR * tdissz * sakpp / (1.0f + DELT2 * tdissz * sakpp);
      szt = szt + PLAVOR * (tfrc / (1.0f + DELT2 * tfrc) + tdissz * sakpp / (1.0f + DELT2 * tdissz * sakpp));
    }
    lev(SLOT_szp, l) = szp;
    lev(SLOT_stp, l) = stp;
    lev(SLOT_sdp, l) = sdp;
    lev(SLOT_sqp, l) = sqp;
    lev(SLOT_stt, l) = stt;
    lev(SLOT_sqt, l) = sqt;
    lev(SLOT_sdt, l) = zsdt1 + zsdt2;
    lev(SLOT_szt, l) = szt;
    lev(SLOT_zdsdp1, l) = zsdt1 * DELT2;
    lev(SLOT_zdszp1, l) = zszt1 * DELT2;
    lev(SLOT_zdsdp2, l) = zsdt2 * DELT2;
    lev(SLOT_zdszp2, l) = zszt2 * DELT2;
  }
}

__device__ void update_coeff(const DevMember& M, int s, int l) {
  // dynamics/wgsl/spectrald-update.wgsl:12-45
  const int NRSP = M.nrsp;
  float* A = dev_arena(M);
  auto lev = [&](int slot, int l) -> float& { return A[M.slot[slot] + l * NRSP + s]; };
  const float DELT2 = M.k[CON_delt2], PNU = M.k[CON_pnu];
  const bool FILTER = M.k[CON_filter] != 0.0f;
  if (l == 0) {
    const float spp = A[M.slot[SLOT_spp] + s];
    if (FILTER) A[M.slot[SLOT_spm] + s] = A[M.slot[SLOT_spm] + s] + PNU * spp;
    A[M.slot[SLOT_sp] + s] = spp;
  }
  {
    const float stt = lev(SLOT_stt, l) + lev(SLOT_zstt1, l) + lev(SLOT_zstt2, l);
    lev(SLOT_stt, l) = stt;
    const float szp = lev(SLOT_szp, l) + DELT2 * lev(SLOT_szt, l);
    const float stp = lev(SLOT_stp, l) + DELT2 * stt;
    const float sdp = lev(SLOT_sdp, l) + DELT2 * lev(SLOT_sdt, l);
    const float sqp = lev(SLOT_sqp, l) + DELT2 * lev(SLOT_sqt, l);
    lev(SLOT_szp, l) = szp;
    lev(SLOT_stp, l) = stp;
    lev(SLOT_sdp, l) = sdp;
    lev(SLOT_sqp, l) = sqp;
    if (FILTER) {
      lev(SLOT_szm, l) = lev(SLOT_szm, l) + PNU * szp;
      lev(SLOT_stm, l) = lev(SLOT_stm, l) + PNU * stp;
      lev(SLOT_sdm, l) = lev(SLOT_sdm, l) + PNU * sdp;
      lev(SLOT_sqm, l) = lev(SLOT_sqm, l) + PNU * sqp;
    }
    lev(SLOT_sz, l) = szp;
    lev(SLOT_st, l) = stp;
    lev(SLOT_sd, l) = sdp;
    lev(SLOT_sq, l) = sqp;
  }
}

// dynamics/wgsl/physics-in.wgsl:8-34 at one point and level (levels
// 0..NL: the tendency rows have NL + 1 planes); each thread takes dp from
// gp itself, the level-0 thread stores it.
__device__ void physics_in_point(const DevMember& M, int h, int l) {
  const int NHOR = M.nhor, NL = M.nlev;
  float* A = dev_arena(M);
  const float PSURF = M.k[CON_psurf], CV = M.k[CON_cv], CT = M.k[CON_ct];
  const int at = l * NHOR + h;
  A[M.slot[SLOT_dudt] + at] = 0.0f;
  A[M.slot[SLOT_dvdt] + at] = 0.0f;
  A[M.slot[SLOT_dtdt] + at] = 0.0f;
  A[M.slot[SLOT_dqdt] + at] = 0.0f;
  if (l >= NL) return;
  const float dp = PSURF * expf(A[M.slot[SLOT_gp] + h]);
  if (l == 0) A[M.slot[SLOT_dp] + h] = dp;
  const float scale = sqrtf(dev_tab(M)[M.tabo[TAB_rcsq] + h]);
  A[M.slot[SLOT_gudt] + at] = 0.0f;
  A[M.slot[SLOT_gvdt] + at] = 0.0f;
  A[M.slot[SLOT_gtdt] + at] = 0.0f;
  A[M.slot[SLOT_gqdt] + at] = 0.0f;
  A[M.slot[SLOT_du] + at] = CV * A[M.slot[SLOT_gu] + at] * scale;
  A[M.slot[SLOT_dv] + at] = CV * A[M.slot[SLOT_gv] + at] * scale;
  A[M.slot[SLOT_dt] + at] = CT * (A[M.slot[SLOT_gt] + at] + dev_tab(M)[M.tabo[TAB_t0] + l]);
  A[M.slot[SLOT_dq] + at] = A[M.slot[SLOT_gq] + at] * PSURF / dp;
}

}  // namespace

#define EXO_LEVEL_DISPATCH(fn, M, i) \
  do {                               \
    if ((M).nlev == 10)              \
      fn<10>(M, i);                  \
    else                             \
      fn<5>(M, i);                   \
  } while (0)

extern "C" __global__ __launch_bounds__(64) void k_calcgp(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int h = tile * 64 + threadIdx.x;
  if (h >= M.nhor) return;
  EXO_LEVEL_DISPATCH(calcgp_column, M, h);
}

extern "C" __global__ __launch_bounds__(640) void k_spectrala(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  __shared__ LevelRows lds[4];
  if (M.nlev == 10)
    spectrala_levels<10>(M, tile, lds);
  else
    spectrala_levels<5>(M, tile, lds);
}

extern "C" __global__ __launch_bounds__(256) void k_spectrald_diffusion(const DevMember* __restrict__ mems,
                                                                         LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int i = tile * 256 + threadIdx.x;
  if (i >= M.nlev * M.nrsp) return;
  const int l = i / M.nrsp;
  diffusion_coeff(M, i - l * M.nrsp, l);
}

extern "C" __global__ __launch_bounds__(256) void k_spectrald_update(const DevMember* __restrict__ mems,
                                                                      LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int i = tile * 256 + threadIdx.x;
  if (i >= M.nlev * M.nrsp) return;
  const int l = i / M.nrsp;
  update_coeff(M, i - l * M.nrsp, l);
}

extern "C" __global__ __launch_bounds__(256) void k_physics_in(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int i = tile * 256 + threadIdx.x;
  if (i >= (M.nlev + 1) * M.nhor) return;
  const int l = i / M.nhor;
  physics_in_point(M, i - l * M.nhor, l);
}

// dynamics/wgsl/mkdheat.wgsl:17-44, stages 1 and 2 in one pass per level
// and point. The reference runs them one after the other because both
// use the increment winds' rows (zdu, zdv); here the first increments'
// winds come in the frame's wind rows (gu, gv), free by then, and stage
// 1's heating goes out through the frame's gtn row, so both stages read
// the same winds (zu, zv) together: stage 1 heats from the first
// increments, stage 2 takes the kinetic-energy defect of the second.
extern "C" __global__ __launch_bounds__(64) void k_mkdheat(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int i = tile * 64 + threadIdx.x;
  if (i >= M.nlev * M.nhor) return;
  const int l = i / M.nhor, h = i % M.nhor;
  float* A = dev_arena(M);
  const float CV = M.k[CON_cv], CT = M.k[CON_ct], WW = M.k[CON_ww], GA = M.k[CON_ga];
  const float ACPD = M.k[CON_acpd], ADV = M.k[CON_adv], DELTSEC2 = M.k[CON_deltsec2];
  const float heat = 1.0f + ADV * A[M.slot[SLOT_dq] + i];
  const float dsig = dev_tab(M)[M.tabo[TAB_dsigma] + l];
  const float scale = CV * sqrtf(dev_tab(M)[M.tabo[TAB_rcsq] + h]);
  const float u = A[M.slot[SLOT_zu] + i] * scale;
  const float v = A[M.slot[SLOT_zv] + i] * scale;
  {
    const float du = A[M.slot[SLOT_gu] + i] * scale;
    const float dv = A[M.slot[SLOT_gv] + i] * scale;
    const float change = du * (2.0f * u + du) + dv * (2.0f * v + dv);
    const float dtdt = -change / DELTSEC2 * 0.5f / ACPD / heat;
    A[M.slot[SLOT_gtn] + i] = dtdt / CT / WW;
  }
  {
    const float du = A[M.slot[SLOT_zdu] + i] * scale;
    const float dv = A[M.slot[SLOT_zdv] + i] * scale;
    const float change = du * (2.0f * u + du) + dv * (2.0f * v + dv);
    A[M.slot[SLOT_zdekin] + i] = change / DELTSEC2 * A[M.slot[SLOT_dp] + h] / GA * dsig;
  }
}


// mkdheat's dissipation of the mean kinetic-energy defect, stages 2b-3
// (dynamics/index.ts:285-293): the defect zdekin (grid) is transformed to
// spectral space (gp2fc then fc2sp with qc), cut to its global-mean mode
// (mean-mode.wgsl:4-13), transformed back (sp2fc with qi then fc2gp), and
// turned into heating (mkdheat.wgsl, stage 3). Only mode (0, 0) survives,
// so the chain is, per level l:
//   S = sum_lat,x qc(lat, 0) zdekin(lat, x) fwd(x, 0)
//   zdekin(lat, x) = (qi(lat, 0) S) inv(0, x)
// with the same tables; one workgroup per member and level computes it
// (every other spectral coefficient of zsdef set to 0, as mean-mode
// leaves it) and the heating.
// Rows per wave and longitudes per lane of the largest grid plane (T63:
// 96 rows of 192 points over 16 waves of 64 lanes).
constexpr int kDekinRows = 6, kDekinCols = 3;

extern "C" __global__ __launch_bounds__(1024) void k_mean_dekin(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int l;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &l)];
  const int NH = M.nhor, nlat = M.nlat, nlon = M.nlon, ncsp = M.ncsp, nrsp = M.nrsp;
  float* A = dev_arena(M);
  float* zdekin = A + M.slot[SLOT_zdekin] + l * NH;
  const float* qc = dev_ltab(M) + M.tabo[TAB_legendre_qc];
  const float* qi = dev_ltab(M) + M.tabo[TAB_legendre_qi];
  // The m = 0 column of the forward DFT table is cos(0) / NLON and the m = 0
  // row of the inverse table cos(0) = 1 at every longitude: one value each.
  const float fwd0 = dev_ltab(M)[M.tabo[TAB_fourier_forward]];
  const float inv0 = dev_ltab(M)[M.tabo[TAB_fourier_inverse]];
  const int half = nlat / 2;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // Wave w takes rows w, w + 16, ...; lane x takes longitudes x, x + 64, ...
  // Points past the plane load its last row or longitude, unconditionally,
  // and are neither summed nor stored.
  const auto lat_of = [&](int r) { return min(wave + 16 * r, nlat - 1); };
  const auto lon_of = [&](int c) { return min(lane + 64 * c, nlon - 1); };
  const auto inside = [&](int r, int c) { return wave + 16 * r < nlat && lane + 64 * c < nlon; };
  // Fourier mode 0 of each point's row, weighted by its latitude's qc
  // (mode (0, 0) is even about the equator: a southern row takes its
  // northern twin's), every load issued before the sums. The planet's
  // dissipated energy is summed in FP64 (thread, then wave butterfly, then
  // the 16 waves in order) and rounded to FP32 once.
  float prod[kDekinRows][kDekinCols];
#pragma unroll
  for (int r = 0; r < kDekinRows; r++) {
    const int lat = lat_of(r);
    const float w = qc[(lat < half ? lat : nlat - 1 - lat) * ncsp];
#pragma unroll
    for (int c = 0; c < kDekinCols; c++) prod[r][c] = w * (zdekin[lat * nlon + lon_of(c)] * fwd0);
  }
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < kDekinRows; r++)
#pragma unroll
    for (int c = 0; c < kDekinCols; c++)
      if (inside(r, c)) s += (double)prod[r][c];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  __shared__ double part[16];
  __shared__ float total;
  if (lane == 0) part[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; w++) t += part[w];
    total = (float)t;
  }
  __syncthreads();
  const float S = total;
  // zsdef keeps only Re mode (0, 0)
  float* zsdef = A + M.slot[SLOT_zsdef] + l * nrsp;
  for (int c = threadIdx.x; c < nrsp; c += 1024) zsdef[c] = c == 0 ? S : 0.0f;
  // back to the grid, then the heating (mkdheat stage 3); the inputs of
  // every point first, then the stores
  const float CT = M.k[CON_ct], WW = M.k[CON_ww], GA = M.k[CON_ga], ACPD = M.k[CON_acpd], ADV = M.k[CON_adv];
  const float dsig = dev_tab(M)[M.tabo[TAB_dsigma] + l];
  const float* dq = A + M.slot[SLOT_dq] + l * NH;
  const float* dp = A + M.slot[SLOT_dp];
  float* zdtdt = A + M.slot[SLOT_zdtdt] + l * NH;
  float f[kDekinRows], dqv[kDekinRows][kDekinCols], dpv[kDekinRows][kDekinCols];
#pragma unroll
  for (int r = 0; r < kDekinRows; r++) {
    const int lat = lat_of(r);
    f[r] = qi[lat * ncsp] * S;
#pragma unroll
    for (int c = 0; c < kDekinCols; c++) {
      dqv[r][c] = dq[lat * nlon + lon_of(c)];
      dpv[r][c] = dp[lat * nlon + lon_of(c)];
    }
  }
#pragma unroll
  for (int r = 0; r < kDekinRows; r++)
#pragma unroll
    for (int c = 0; c < kDekinCols; c++)
      if (inside(r, c)) {
        const int h = (wave + 16 * r) * nlon + lane + 64 * c;
        const float e = f[r] * inv0;
        zdekin[h] = e;
        const float heat = 1.0f + ADV * dqv[r][c];
        const float dtdt = -e * 0.5f / ACPD / heat / dpv[r][c] * GA / dsig;
        zdtdt[h] = dtdt / CT / WW;
      }
}
