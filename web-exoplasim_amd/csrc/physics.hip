// Column physics of the ExoPlaSim step (gridpointd's parameterizations) on
// gfx950, fused one lane per grid column. The reference dispatches 23
// kernels between the physics frame in and out (fixer 4, fluxstep 2,
// radstep 5, rainstep 6, surfstep 6); here the per-column routines run
// back to back inside a few kernels whose boundaries are where a routine
// needs every column of the planet first, or where splitting a pass lets
// each part run at a higher occupancy than their union:
//
//   k_fixer_column   fixer's column pass (physics/fixer/wgsl/column.wgsl)
//   k_fixer_rows     fixer's latitude rows, one wave per row
//                    (fixer/wgsl/row.wgsl)
//   k_fixer_sum      then its global sum, one workgroup per member
//                    (reduction/wgsl/sum.wgsl)
//   k_surface_fluxes fixer apply, surflx, vdiff
//   k_shortwave      solang, mko3, swr
//   k_longwave       lwr, radstep
//   k_convection     kuo + mkshallow, mkdca's column pass
//   k_dca_limit      mkdca's global give-up flag (mkdca-limit.wgsl)
//   k_rain           rain column: mkdca apply, mklsp, mkrain
//   k_clouds         mkclouds
//   k_soil           soil and snow (land/wgsl/soil.wgsl)
//   k_mkradv         river routing stencil (land/wgsl/mkradv.wgsl)
//   k_physics_c      land surface and albedo, sea ice up to getiflx
//   k_ice_sums       getiflx's two global sums
//   k_physics_d      ice finish + slab ocean + sea surface, carbonstep,
//                    and the physics frame out (physics-out.wgsl)
//   k_storm          hurricanestep's storm diagnostics, members with
//                    them on (storm/wgsl/storm.wgsl)
//
// Each routine's statement order follows the reference WGSL cited at it.
#include "engine.h"

namespace {

#define F(s) (dev_arena(M) + M.slot[SLOT_##s])
#define T(s) (dev_tab(M) + M.tabo[TAB_##s])
#define K(s) (M.k[CON_##s])
#define LV(s, l, h) (F(s)[(l) * NH + (h)])

__device__ __forceinline__ float powi(float x, int n) {
  float r = 1.0f;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

// Logarithms and powers of positive, normal arguments on the hardware's
// base-2 units (v_log_f32, v_exp_f32; about 1 ulp each) instead of the
// library's correctly rounded routines, whose argument reduction (with a
// full-precision division) costs several times as much. Used where the
// routine's arguments are positive and normal (the longwave absorptivities'
// column amounts + offsets, the shortwave's ozone and vapour paths and
// cloud optical depths, the surface layer's roughness ratio and stability
// terms); the results differ from the library's by a few ulp. Exponentials
// of large arguments keep the library's expf, whose argument reduction the
// plain exp2(x log2 e) would lose.
__device__ __forceinline__ float log2_pos(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float pow_pos(float a, float b) { return __builtin_amdgcn_exp2f(b * log2_pos(a)); }
__device__ __forceinline__ float log10_pos(float x) { return log2_pos(x) * 0.30102999566398120f; }
__device__ __forceinline__ float log_pos(float x) { return log2_pos(x) * 0.69314718055994531f; }
// e^x for x <= 0 down to the normal range: exp2(x log2 e)
__device__ __forceinline__ float exp_neg(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// --------------------------------------------------------------- fixer

template <int NL>
__device__ __forceinline__ void fixer_column(const DevMember& M, int h) {
  // physics/fixer/wgsl/column.wgsl:4-36
  const int NH = M.nhor;
  float neg = 0, pos = 0;
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float dq = LV(dq, l, h);
    const float q = dq * T(dsigma)[l];
    if (dq < 0) neg += q;
    if (dq > 0) pos += q;
  }
  float zneg = -neg * F(dp)[h] * T(zgw)[h];
  float zpos = pos * F(dp)[h] * T(zgw)[h];
  float factor = 1;
  const bool fixed = zneg > 0 && zpos >= zneg;
  if (fixed) {
    factor = (zpos - zneg) / zpos;
    zneg = 0;
    zpos = zpos * factor;
  }
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float q = LV(dq, l, h);
    LV(fixq, l, h) = fixed ? fmaxf(q * factor, 0.0f) : q;
  }
  F(fixneg)[h] = zneg;
  F(fixpos)[h] = zpos;
}

template <int NL>
__device__ __forceinline__ void fixer_apply(const DevMember& M, int h) {
  // physics/fixer/wgsl/apply.wgsl:4-23
  const int NH = M.nhor;
  const float zzneg = F(fixsum)[0], zzpos = F(fixsum)[1];
  const float dp = F(dp)[h];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    float q = LV(fixq, l, h);
    if (zzneg > 0) q = fmaxf(q * ((zzpos - zzneg) / zzpos), 0.0f);
    LV(gqdt, l, h) = LV(gqdt, l, h) + dp * (q - LV(dq, l, h)) / K(delt2) / K(psurf);
    LV(dq, l, h) = q;
  }
}

// -------------------------------------------------------- surface flux

template <int NL>
__device__ __forceinline__ void surflx(const DevMember& M, int h) {
  // physics/surface-flux/wgsl/surflx.wgsl:30-128
  const int NH = M.nhor, low = NL - 1;
  const float VK = 0.4f, GA = K(ga), GASCON = K(gascon), DELTSEC2 = K(deltsec2);
  const float moisture = 1.0f / K(rdbrv) - 1.0f;
  const float u = LV(du, low, h), v = LV(dv, low, h), t = LV(dt, low, h), q = LV(dq, low, h);
  const float ts = LV(dt, NL, h), qs = LV(dq, NL, h);
  const float dp = F(dp)[h], dls = F(dls)[h], dsigma = T(dsigma)[low];
  const float zfac = T(flux_sigfac)[0], zlnsig = T(flux_lnsig)[0];
  float dtsa = t * zfac * (1.0f + moisture * q);
  const float zabsu2 = fmaxf(K(zumin), u * u + v * v);
  const float znl = -GASCON * 0.5f * (t + dtsa) * zlnsig / GA;
  const float zbz0 = znl / F(dz0)[h];
  const float zri = GA * znl / (zabsu2 * dtsa) * (dtsa - ts * (1.0f + moisture * qs));
  const float zkblnz = VK / log_pos(zbz0 + 1.0f);
  const float zkblnz2 = zkblnz * zkblnz;
  float zrifm, zrifh;
  if (zri <= 0) {
    const float zdenom = 1.0f + 3.0f * K(vdiff_c) * K(vdiff_b) * sqrtf(-1.0f * zri * (zbz0 + 1.0f)) * zkblnz2;
    zrifm = 1.0f - 2.0f * K(vdiff_b) * zri / zdenom;
    if (dls < 1.0f) {
      const float zdth = -(dtsa - ts * (1.0f + moisture * qs));
      // zdth >= 0 here; pow_pos(0, 1/3) = exp2(-inf) = 0
      zrifh = pow_pos(1.0f + pow_pos(0.0016f * pow_pos(zdth, 1.0f / 3.0f) / sqrtf(zabsu2) / zkblnz2, 1.25f), 0.8f);
    } else {
      zrifh = 1.0f - 3.0f * K(vdiff_b) * zri / zdenom;
    }
  } else {
    const float zdenom = sqrtf(1.0f + K(vdiff_d) * zri);
    zrifh = 1.0f / (1.0f + 3.0f * K(vdiff_b) * zri * zdenom);
    zrifm = 1.0f / (1.0f + 2.0f * K(vdiff_b) * zri / zdenom);
  }
  const float ztrans = sqrtf(zabsu2) * zkblnz2;
  const float transh = ztrans * zrifh, transm = ztrans * zrifm;
  F(dtransh)[h] = transh;
  F(dtransm)[h] = transm;
  const float zkonst1 = GA * DELTSEC2 / (GASCON * dsigma);
  const float zkonst2 = dsigma / DELTSEC2 / GA;
  const float zkdiffm = zkonst1 * transm / ts;
  const float zun = u / (1.0f + zkdiffm), zvn = v / (1.0f + zkdiffm);
  LV(dudt, low, h) = LV(dudt, low, h) + (zun - u) / DELTSEC2;
  LV(dvdt, low, h) = LV(dvdt, low, h) + (zvn - v) / DELTSEC2;
  float dtdt = LV(dtdt, low, h);
  if (K(ndheat) != 0)
    dtdt = dtdt + -(zun * zun - u * u + zvn * zvn - v * v) / DELTSEC2 * 0.5f / K(acpd) / (1.0f + K(adv) * q);
  const float taux = dp * zkonst2 * zkdiffm * zun, tauy = dp * zkonst2 * zkdiffm * zvn;
  F(dtaux)[h] = taux;
  F(dtauy)[h] = tauy;
  const float ustar = sqrtf(sqrtf(taux * taux + tauy * tauy) * GASCON * ts / dp);
  F(dust3)[h] = ustar * ustar * ustar;
  const float zkdiffh = zkonst1 * transh / ts;
  const float ztn = (t + zkdiffh * ts) / (1.0f + zkdiffh * zfac);
  dtdt = dtdt + (ztn - t) / DELTSEC2;
  LV(dtdt, low, h) = dtdt;
  const float moist = fmaxf(1.0f, 1.0f + K(adv) * q);
  F(dshfl)[h] = zkonst2 * zkdiffh * dp * K(acpd) * moist * (ztn * zfac - ts);
  F(dshdt)[h] = -1.0f * zkonst2 * zkdiffh * dp * K(acpd) * moist;
  dtsa = ztn * zfac;
  F(dtsa)[h] = dtsa;
  const float zkdiffq = F(drhs)[h] * zkonst1 * transh / ts;
  float zqn = fmaxf(q, (q + zkdiffq * qs) / (1.0f + zkdiffq));
  if (dls > 0) zqn = fminf(zqn, F(dwatc)[h] / K(deltsec) * 1000.0f / dp / zkonst2 + q);
  LV(dqdt, low, h) = LV(dqdt, low, h) + (zqn - q) / DELTSEC2;
  const float devap = -dp * zkonst2 / 1000.0f * (zqn - q);
  F(devap)[h] = devap;
  const float latent = (ts > K(tmelt) || dls < 0.5f) ? K(alv) : K(als);
  const float dlhfl = devap * latent * 1000.0f;
  F(dlhfl)[h] = dlhfl;
  const float curvature = (ts - K(ra4)) * (ts - K(ra4));
  const float dlhdt = -1.0f * latent * zkdiffq * zkonst2 * dp * K(ra2) * (K(tmelt) - K(ra4)) * qs / curvature;
  F(dlhdt)[h] = dlhfl == 0.0f ? 0.0f : dlhdt;
}

// The elimination divides each row by its pivot once (one reciprocal, then
// products) where the reference divides twice; the results agree to an ulp.
template <int NL>
__device__ __forceinline__ void diffuse(const DevMember& M, float* x, const float* k) {
  // physics/diffusion/wgsl/implicit.wgsl:4-26
  float ebs[NL];
  const float* ds = T(dsigma);
  float r = 1.0f / (ds[0] + k[0]);
  ebs[0] = k[0] * r;
  x[0] = ds[0] * x[0] * r;
#pragma unroll
  for (int l = 1; l < NL - 1; l++) {
    r = 1.0f / (ds[l] + k[l] + k[l - 1] * (1.0f - ebs[l - 1]));
    ebs[l] = k[l] * r;
    x[l] = (x[l] * ds[l] + k[l - 1] * x[l - 1]) * r;
  }
  x[NL - 1] = (x[NL - 1] * ds[NL - 1] + k[NL - 2] * x[NL - 2]) / (ds[NL - 1] + k[NL - 2] * (1.0f - ebs[NL - 2]));
#pragma unroll
  for (int l = NL - 2; l >= 0; l--) x[l] = x[l] + ebs[l] * x[l + 1];
}

template <int NL>
__device__ __forceinline__ void diffuse_potential(const DevMember& M, float* x, const float* k, const float* skap) {
  // physics/diffusion/wgsl/implicit.wgsl:28-47
  float ebs[NL], rs[NL];
  const float* ds = T(dsigma);
#pragma unroll
  for (int l = 0; l < NL; l++) rs[l] = 1.0f / skap[l];
  float r = 1.0f / (ds[0] + k[0] * rs[0]);
  ebs[0] = k[0] * r;
  x[0] = ds[0] * x[0] * r;
#pragma unroll
  for (int l = 1; l < NL - 1; l++) {
    r = 1.0f / (ds[l] + (k[l] + k[l - 1] * (1.0f - ebs[l - 1] * rs[l - 1])) * rs[l]);
    ebs[l] = k[l] * r;
    x[l] = (x[l] * ds[l] + k[l - 1] * rs[l - 1] * x[l - 1]) * r;
  }
  const int last = NL - 1;
  x[last] = (x[last] * ds[last] + k[NL - 2] * x[NL - 2] * rs[NL - 2]) /
            (ds[last] + k[NL - 2] * rs[last] * (1.0f - ebs[NL - 2] * rs[NL - 2]));
#pragma unroll
  for (int l = NL - 2; l >= 0; l--) x[l] = x[l] + ebs[l] * x[l + 1] * rs[l + 1];
}

template <int NL>
__device__ __forceinline__ void vdiff(const DevMember& M, int h) {
  // physics/surface-flux/wgsl/vdiff.wgsl:22-120
  // Scheduling barriers between the solves keep the compiler from hoisting
  // every level's loads and products of one solve into the previous: the
  // routine then fits 207 VGPRs with no spill (256 and 84 bytes of scratch
  // per lane without them); the arithmetic is unchanged.
  const int NH = M.nhor;
  const float DELTSEC2 = K(deltsec2), GA = K(ga), GASCON = K(gascon);
  const float moisture = 1.0f / K(rdbrv) - 1.0f;
  float zt[NL], zq[NL], zu[NL], zv[NL], zabsu[NL], zskap[NL], zskaph[NL], kh[NL], km[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    zt[l] = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
    zq[l] = LV(dq, l, h) + LV(dqdt, l, h) * DELTSEC2;
    zu[l] = LV(du, l, h) + LV(dudt, l, h) * DELTSEC2;
    zv[l] = LV(dv, l, h) + LV(dvdt, l, h) * DELTSEC2;
    zabsu[l] = sqrtf(zu[l] * zu[l] + zv[l] * zv[l]);
    zskap[l] = T(vdiff_skap)[l];
    zskaph[l] = T(vdiff_skaph)[l];
  }
  const float zkonst1 = GA * DELTSEC2 / GASCON;
  const float zkonst2 = zkonst1 * GA / GASCON;
  const float zkonst3 = zkonst2 * GA / GASCON;
  kh[NL - 1] = km[NL - 1] = 0;
#pragma unroll
  for (int l = 0; l < NL - 1; l++) {
    const int lp = l + 1;
    const float sigmah = T(sigmah)[l], dsigma = T(dsigma)[l], dsigmap = T(dsigma)[lp];
    const float zmixh2 = T(vdiff_mixh2)[l], zmixm2 = T(vdiff_mixm2)[l], zrdsig = T(vdiff_rdsig)[l];
    const float zfac = zkonst3 * T(vdiff_sigmah3)[l];
    const float zzkfac = T(vdiff_kfac)[l];
    const float zrthl = (dsigmap + dsigma) / (zt[l] * dsigmap + zt[lp] * dsigma);
    const float zdvds = fmaxf(K(zumin), fabsf(zabsu[l] - zabsu[lp])) * zrdsig;
    const float zdthds =
        (zt[l] * (1.0f + moisture * zq[l]) / zskap[l] - zt[lp] * (1.0f + moisture * zq[lp]) / zskap[lp]) * zrdsig;
    const float zri = zskaph[l] * GASCON * zdthds / (sigmah * zdvds * zdvds);
    float zrifh, zrifm;
    if (zri <= 0) {
      const float zdh = 1.0f + 3.0f * K(vdiff_c) * K(vdiff_b) * sqrtf(-1.0f * zri) * zzkfac * zmixh2;
      const float zdm = 1.0f + 3.0f * K(vdiff_c) * K(vdiff_b) * sqrtf(-1.0f * zri) * zzkfac * zmixm2;
      zrifh = 1.0f - 3.0f * K(vdiff_b) * zri / zdh;
      zrifm = 1.0f - 2.0f * K(vdiff_b) * zri / zdm;
    } else {
      const float zdenom = sqrtf(1.0f + K(vdiff_d) * zri);
      zrifh = 1.0f / (1.0f + 3.0f * K(vdiff_b) * zri * zdenom);
      zrifm = 1.0f / (1.0f + 2.0f * K(vdiff_b) * zri / zdenom);
    }
    kh[l] = zfac * zmixh2 * zrifh * zdvds * zrthl * zrthl * zrthl;
    km[l] = zfac * zmixm2 * zrifm * zdvds * zrthl * zrthl * zrthl;
  }
  __builtin_amdgcn_sched_barrier(0);
  float zun[NL], zvn[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    zun[l] = zu[l];
    zvn[l] = zv[l];
  }
  diffuse<NL>(M, zun, km);
  diffuse<NL>(M, zvn, km);
#pragma unroll
  for (int l = 0; l < NL; l++) {
    LV(dudt, l, h) = LV(dudt, l, h) + (zun[l] - zu[l]) / DELTSEC2;
    LV(dvdt, l, h) = LV(dvdt, l, h) + (zvn[l] - zv[l]) / DELTSEC2;
  }
  __builtin_amdgcn_sched_barrier(0);
  if (K(ndheat) != 0) {
    float zke[NL], zken[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) zke[l] = zken[l] = 0.5f * zabsu[l] * zabsu[l];
    diffuse<NL>(M, zken, km);
#pragma unroll
    for (int l = 0; l < NL; l++) {
      const float zdtdt = -((zun[l] * zun[l] - zu[l] * zu[l] + zvn[l] * zvn[l] - zv[l] * zv[l]) / DELTSEC2 -
                            (zken[l] - zke[l]) / DELTSEC2) *
                          0.5f / K(acpd) / (1.0f + K(adv) * LV(dq, l, h));
      LV(dtdt, l, h) = LV(dtdt, l, h) + zdtdt;
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  float zqn[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) zqn[l] = zq[l];
  diffuse<NL>(M, zqn, kh);
#pragma unroll
  for (int l = 0; l < NL; l++) LV(dqdt, l, h) = LV(dqdt, l, h) + (zqn[l] - zq[l]) / DELTSEC2;
#pragma unroll
  for (int l = 0; l < NL - 1; l++) kh[l] = kh[l] * zskaph[l];
  __builtin_amdgcn_sched_barrier(0);
  float ztn[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) ztn[l] = zt[l];
  diffuse_potential<NL>(M, ztn, kh, zskap);
#pragma unroll
  for (int l = 0; l < NL; l++) LV(dtdt, l, h) = LV(dtdt, l, h) + (ztn[l] - zt[l]) / DELTSEC2;
}

// ------------------------------------------------------- storm diagnostics
// hurricanemod's storm diagnostics (physics/storm/wgsl/storm.wgsl:1-208,
// hurricanemod.f90 at ExoPlaSim 5f8dff4 with its humidity convention):
// CAPE of the lowest-level parcel, Emanuel's maximum potential intensity
// (a fixed-point iteration on the storm's central pressure), the entropy
// deficit, the ventilation index and the genesis potential index. Every
// level loop is unrolled so the sounding and the buoyancy stay in
// registers; indices found at run time pick their level by comparison.

namespace storm {

constexpr float CPD = 1005.7f, CL = 2500.0f, RVP = 461.5f, ALV0 = 2.501e6f, CPVMCL = -630.0f;

struct Cape {
  float energy, temperature, pressure;
};

// saturation vapour pressure over water, hPa, of a Celsius temperature
__device__ __forceinline__ float es(float tc) { return 6.112f * expf(17.67f * tc / (243.5f + tc)); }
// vapour pressure of mixing ratio r at pressure p; eps = RD / RVP
__device__ __forceinline__ float vapor(float r, float p, float eps) { return r * p / (eps + r); }
// mixing ratio of vapour pressure e at pressure p
__device__ __forceinline__ float mixing(float e, float p, float eps) { return eps * e / (p - e); }
// density temperature of temperature t, total water rt, vapour r
__device__ __forceinline__ float density(float t, float rt, float r, float eps) {
  return t * (1.0f + r / eps) / (1.0f + rt);
}
__device__ __forceinline__ float entropy(float t, float r, float p, float RD, float eps) {
  const float e = vapor(r, p, eps);
  const float rh = fminf(e / es(t - 273.15f), 1.0f);
  const float lv = ALV0 + CPVMCL * (t - 273.15f);
  const float dry = (CPD + r * CL) * logf(t) - RD * logf(p - e);
  if (r == 0.0f) return dry;
  return dry + lv * r / t - r * RVP * logf(rh);
}

// storm.wgsl:56-131: CAPE of parcel (tp, rp, pp) in the sounding (t, r, p)
template <int NL>
__device__ __forceinline__ Cape cape(const float* t, const float* r, const float* p, float tp, float rp, float pp,
                                     float RD) {
  const float eps = RD / RVP;
  if (rp < 1e-6f || tp < 200.0f || pp <= 0.0f) return Cape{0.0f, -1.0f, -1.0f};
  const float rh = fminf(vapor(rp, pp, eps) / es(tp - 273.15f), 1.0f);
  const float parcel_entropy = entropy(tp, rp, pp, RD, eps);
  const float plcl = pp * powf(rh, tp / (1669.0f - 122.0f * rh - tp));
  float buoyancy[NL];
  // the topmost level first (storm.wgsl:67-68); a level whose saturated
  // ascent fails ends the sounding (status 1: at that level's
  // temperature, 2: at the bottom level's temperature and pressure)
  int status = 0;
  float fail_t = 0.0f;
#pragma unroll
  for (int l = NL - 1; l >= 0; l--) {
    buoyancy[l] = 0.0f;
    if (status == 0) {
      float tg, rg = rp;
      if (p[l] >= plcl) {
        tg = tp * powf(p[l] / pp, RD / CPD);
      } else {
        float next = t[l];
        tg = 0.0f;
        unsigned count = 0;
        for (;;) {
          if (fabsf(next - tg) <= 0.001f) break;
          tg = next;
          const float tc = tg - 273.15f;
          const float enew = es(tc);
          rg = mixing(enew, p[l], eps);
          count++;
          const float lv = ALV0 + CPVMCL * tc;
          const float sl = (CPD + rp * CL + lv * lv * rg / (RVP * tg * tg)) / tg;
          const float em = vapor(rg, p[l], eps);
          if (p[l] - em <= 0.0f || tg <= 0.0f) {
            status = 1;
            fail_t = t[l];
            break;
          }
          const float sg = (CPD + rp * CL) * logf(tg) - RD * logf(p[l] - em) + lv * rg / tg;
          next = tg + (count < 3u ? 0.3f : 1.0f) * (parcel_entropy - sg) / sl;
          if (count > 500u || enew > p[l] - 1.0f || !(next > 0.0f)) {
            status = 2;
            break;
          }
        }
      }
      buoyancy[l] = density(tg, rp, rg, eps) - density(t[l], r[l], r[l], eps);
    }
  }
  if (status == 1) return Cape{0.0f, fail_t, -1.0f};
  if (status == 2) return Cape{0.0f, t[NL - 1], p[NL - 1]};
  // inb: the first (topmost) positively buoyant level, with its values and
  // those of the level above, picked while searching
  int inb = NL - 1;
  float bi = buoyancy[NL - 1], ti = t[NL - 1], pi = p[NL - 1], bm = 0.0f, tm = 0.0f, pm = 0.0f;
#pragma unroll
  for (int l = NL - 1; l >= 0; l--)
    if (buoyancy[l] > 0.0f) {
      inb = l;
      bi = buoyancy[l];
      ti = t[l];
      pi = p[l];
      bm = l > 0 ? buoyancy[l > 0 ? l - 1 : 0] : 0.0f;
      tm = l > 0 ? t[l > 0 ? l - 1 : 0] : 0.0f;
      pm = l > 0 ? p[l > 0 ? l - 1 : 0] : 0.0f;
    }
  if (inb == NL - 1) return Cape{0.0f, t[NL - 1], 0.0f};
  float area = 0.0f;
#pragma unroll
  for (int j = NL - 2; j >= 0; j--)
    if (j >= inb) area += RD * (buoyancy[j] + buoyancy[j + 1]) * (p[j + 1] - p[j]) / (p[j] + p[j + 1]);
  area += RD * (pp - p[NL - 1]) / (pp + p[NL - 1]) * buoyancy[NL - 1];
  float tob = ti, lnb = pi;
  if (inb > 0) {
    lnb = (pm * bi - pi * bm) / (bi - bm);
    area += RD * bi * (pi - lnb) / (pi + lnb);
    tob = (ti * (lnb - pm) + tm * (pi - lnb)) / (pi - pm);
  }
  return Cape{fmaxf(area, 0.0f), tob, lnb};
}

// storm.wgsl:133-138: the real part of upstream's principal complex cube
// roots; zero at vi = 0 and |vi| >= vithresh
__device__ __forceinline__ float reduced(float vi, float vithresh) {
  if (vi == 0.0f || fabsf(vi) >= vithresh) return 0.0f;
  return 2.0f / sqrtf(3.0f) * cosf(acosf(-vi / vithresh) / 3.0f);
}

}  // namespace storm

template <int NL>
__device__ __forceinline__ void storm_diagnostics(const DevMember& M, int h) {
  // storm.wgsl:140-208
  using namespace storm;
  const int NH = M.nhor;
  const float RD = K(gascon), OMEGA = K(ww), CKCD = K(ckcd), eps = RD / RVP;
  const float* sigma = T(sigma);
  float t[NL], r[NL], p[NL];
  const float msl = F(dp)[h] * 0.01f;
  const float sst = LV(dt, NL, h);
  int i600 = 0, i850 = 0, i200 = 0;
  bool cold = false;
#pragma unroll
  for (int l = 0; l < NL; l++) {
    t[l] = LV(dt, l, h);
    r[l] = LV(dq, l, h);
    p[l] = msl * sigma[l];
    cold = cold || t[l] <= 100.0f;
    if (fabsf(sigma[l] - 0.6f) < fabsf(sigma[i600] - 0.6f)) i600 = l;
    if (fabsf(sigma[l] - 0.85f) < fabsf(sigma[i850] - 0.85f)) i850 = l;
    if (fabsf(sigma[l] - 0.2f) < fabsf(sigma[i200] - 0.2f)) i200 = l;
  }
  const int bottom = NL - 1;
  const float u850 = LV(du, i850, h), v850 = LV(dv, i850, h), u200 = LV(du, i200, h), v200 = LV(dv, i200, h);
  const float lowwind = sqrtf(u850 * u850 + v850 * v850), highwind = sqrtf(u200 * u200 + v200 * v200);
  const float shear = fabsf(highwind - lowwind);
  const float vort = LV(gz, i850, h) * OMEGA;
  // the level nearest sigma 0.6 is the member's (sigma is a table)
  const float t600 = LV(dt, i600, h), r600 = LV(dq, i600, h), p600 = msl * sigma[i600];
  float deficit = 1.0f;
  if (r600 > 0.0f && r[bottom] > 0.0f) {
    const float smr = mixing(es(t600 - 273.15f), p600, eps);
    const float sms = entropy(t600, smr, p600, RD, eps);
    const float sme = entropy(t600, fminf(r600, smr), p600, RD, eps);
    const float smb = entropy(sst, LV(dq, NL, h), msl, RD, eps);
    const float smo = entropy(sst, mixing(es(sst - 273.15f), msl, eps), msl, RD, eps);
    if (smo != smb) deficit = (sms - sme) / (smo - smb);
  }
  Cape environmental{-1.0f, -1.0f, -1.0f};
  float vmax = -1.0f;
  if (sst > 278.15f && !cold && msl > 0.0f) {
    environmental = cape<NL>(t, r, p, t[bottom], r[bottom], p[bottom], RD);
    float pm = 970.0f, old = pm, next = 0.0f, capem = 0.0f, capems = 0.0f, ratio = 1.0f;
    bool failed = false;
    unsigned count = 0;
    for (;;) {
      if (fabsf(next - old) <= 0.5f) break;
      const float pp = fminf(pm, 1000.0f);
      const float rp = eps * r[bottom] * msl / (pp * (eps + r[bottom]) - r[bottom] * msl);
      const float rs = mixing(es(sst - 273.15f), pp, eps);
      // the two parcels of one iteration through one inlined CAPE
      float saturated_t = 0.0f;
#pragma unroll 1
      for (int k = 0; k < 2; k++) {
        const Cape c = cape<NL>(t, r, p, k ? sst : t[bottom], k ? rs : rp, pp, RD);
        if (k == 0) {
          capem = c.energy;
        } else {
          capems = c.energy;
          saturated_t = c.temperature;
        }
      }
      ratio = sst / saturated_t;
      const float tvav = 0.5f * (density(t[bottom], r[bottom], r[bottom], eps) + density(sst, rs, rs, eps));
      const float cat = fmaxf((capem - environmental.energy) + 0.5f * CKCD * ratio * (capems - capem), 0.0f);
      next = msl * expf(-cat / (RD * tvav));
      old = pm;
      pm = next;
      count++;
      if (count > 200u || pm < 400.0f || !(pm > 0.0f)) {
        failed = true;
        break;
      }
    }
    if (!failed) vmax = sqrtf(fmaxf(0.0f, CKCD * ratio * fmaxf(0.0f, capems - capem)));
  }
  float vi = 1.0f;
  if (vmax > 0.0f) vi = shear * deficit / vmax;
  const float rh = vapor(r600, p600, eps) / es(t600 - 273.15f);
  const float v70 = vmax / 70.0f, vent = 1.0f + 0.1f * shear;
  const float gpi = powf(fabsf(vort * 1e5f), 1.5f) * powf(rh * 2.0f, 3.0f) * v70 * v70 * v70 / (vent * vent);
  F(capen)[h] = environmental.energy;
  F(lnb)[h] = environmental.pressure;
  F(mpoti)[h] = vmax;
  F(chim)[h] = deficit;
  F(laav)[h] = vort;
  F(venti)[h] = vi;
  F(vrmpi)[h] = reduced(vi, K(vithresh));
  F(gpi)[h] = gpi;
  F(shear)[h] = shear;
}

// ----------------------------------------------------------- radiation

__device__ __forceinline__ void solang(const DevMember& M, int h) {
  // physics/radiation/wgsl/solang.wgsl:10-32
  const float PI = 3.14159265359f, TWOPI = PI + PI;
  const float ZRLON = TWOPI / M.nlon;
  const float* rs = F(radiation_step);
  const int lat = h / M.nlon, lon = h % M.nlon;
  float hangle = rs[2] + lon * ZRLON - PI;
  if (hangle < -PI) hangle = hangle + TWOPI;
  if (hangle > PI) hangle = hangle - TWOPI;
  if (K(rad_synchronous) != 0) hangle = hangle + PI;
  const float muz = rs[0] * T(sid)[lat] + T(cola)[lat] * rs[1] * cosf(hangle);
  F(gmu0)[h] = muz > K(zdawn) ? muz : 0.0f;
}

template <int NL>
__device__ __forceinline__ void mko3(const DevMember& M, int h) {
  // physics/radiation/wgsl/mko3.wgsl:15-37
  const int NH = M.nhor;
  const float ZFO3 = 100.0f / 2.14f, GA = K(ga), GASCON = K(gascon);
  const float sid = T(sid)[h / M.nlon];
  const float za = K(a0o3) + K(a1o3) * fabsf(sid) + K(aco3) * sid * F(radiation_step)[4];
  const float dp = F(dp)[h];
  float o3t = za, zh = 0;
#pragma unroll
  for (int l = NL - 1; l >= 1; l--) {
    zh = zh - LV(dt, l, h) * GASCON / GA * T(rad_lnsig)[l];
    const float o3 = -(za + za * K(zconst_o3)) / (1.0f + expf((zh - K(bo3)) / K(co3))) + o3t;
    LV(dqo3, l, h) = o3 * GA / (ZFO3 * T(dsigma)[l] * dp);
    o3t = o3t - o3;
  }
  LV(dqo3, 0, h) = o3t * GA / (ZFO3 * T(dsigma)[0] * dp);
}

__device__ __forceinline__ float ozone_absorption(const DevMember& M, float o3) {
  return (0.02118f * o3 / (1.0f + 0.042f * o3 + 0.000323f * (o3 * o3)) + 1.082f * o3 / pow_pos(1.0f + 138.6f * o3, 0.805f) +
          0.0658f * o3 / (1.0f + (103.6f * o3) * (103.6f * o3) * (103.6f * o3))) /
         K(zsolar1);
}

__device__ __forceinline__ float vapour_absorption(const DevMember& M, float wv) {
  return 2.9f * wv / (pow_pos(1.0f + 141.5f * wv, 0.635f) + 5.925f * wv) / K(zsolar2);
}

template <int NL>
__device__ __forceinline__ void swr(const DevMember& M, int h) {
  // physics/radiation/wgsl/swr.wgsl:45-276
  const int NH = M.nhor;
  const float ZERO = 1.0e-6f, ZBETTA = 1.66f, ZMBAR = 1.9f, ZFO3 = 100.0f / 2.14f, ZMU00 = 0.5f;
  const float GA = K(ga);
  const float mu0 = F(gmu0)[h];
  const float gdist2 = F(radiation_step)[3];
  const float ftop1 = K(zsolar1) * K(gsol0) * gdist2 * mu0;
  const float ftop2 = K(zsolar2) * K(gsol0) * gdist2 * mu0;
  if (!(ftop1 + ftop2 > ZERO)) {
#pragma unroll
    for (int l = 0; l <= NL; l++) {
      LV(dfu, l, h) = 0;
      LV(dfd, l, h) = 0;
      LV(dswfl, l, h) = 0;
    }
    return;
  }
  const float dp = F(dp)[h];
  float rcl1[NL], rcl1s[NL], rcl2[NL], rcl2s[NL], tcl2[NL], tcl2s[NL], cc[NL];
  float cs = 1.0f;
#pragma unroll
  for (int l = 0; l < NL; l++) {
    cc[l] = LV(dcc, l, h);
    rcl1[l] = rcl1s[l] = rcl2[l] = rcl2s[l] = 0;
    tcl2[l] = tcl2s[l] = 1;
    if (cc[l] > 0) {
      const float lwp = fminf(1000.0f, 1000.0f * LV(dql, l, h) * dp / GA * T(dsigma)[l]);
      const float tau = 2.0f * pow_pos(log10_pos(lwp + 1.5f), 3.9f);
      const float zlog = log_pos(1000.0f / tau);
      float b2 = K(zb4) / log_pos(3.0f + 0.1f * tau);
      float om0 = fminf(0.9999f, 1.0f - K(zb5) * zlog);
      float un = 1.0f - om0;
      float uz = un + 2.0f * b2 * om0;
      float u = sqrtf(uz / un);
      float ex = expf(fminf(25.0f, tau * sqrtf(uz * un) / ZMU00));
      float r = (u + 1.0f) * (u + 1.0f) * ex - (u - 1.0f) * (u - 1.0f) / ex;
      rcl1s[l] = 1.0f - 1.0f / (1.0f + K(zb3) * tau);
      tcl2s[l] = 4.0f * u / r;
      rcl2s[l] = (u * u - 1.0f) / r * (ex - 1.0f / ex);
      const float b1 = K(tswr1) * sqrtf(mu0);
      b2 = K(tswr2) * sqrtf(mu0) / log_pos(3.0f + 0.1f * tau);
      om0 = fminf(0.9999f, 1.0f - K(tswr3) * mu0 * mu0 * zlog);
      un = 1.0f - om0;
      uz = un + 2.0f * b2 * om0;
      u = sqrtf(uz / un);
      ex = expf(fminf(25.0f, tau * sqrtf(uz * un) / mu0));
      r = (u + 1.0f) * (u + 1.0f) * ex - (u - 1.0f) * (u - 1.0f) / ex;
      const float direct1 = 1.0f - 1.0f / (1.0f + b1 * tau / mu0);
      const float directT = 4.0f * u / r;
      const float direct2 = (u * u - 1.0f) / r * (ex - 1.0f / ex);
      rcl1[l] = cs * direct1 + (1.0f - cs) * rcl1s[l];
      tcl2[l] = cs * directT + (1.0f - cs) * tcl2s[l];
      rcl2[l] = cs * direct2 + (1.0f - cs) * rcl2s[l];
    }
    cs = cs * (1.0f - cc[l]);
  }
  const float zm = 35.0f / sqrtf(1.0f + 1224.0f * mu0 * mu0);
  float o3l[NL], xo3l[NL], wvl[NL], ywvl[NL];
  float o3t = 0, xo3t = 0, wvt = 0, ywvt = 0;
  cs = 1.0f;
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float dsg = T(dsigma)[l];
    const float o3 = ZFO3 * dsg * dp * LV(dqo3, l, h) / GA;
    o3t = o3t + o3;
    xo3t = cs * (xo3t + zm * o3) + (1.0f - cs) * (xo3t + ZMBAR * o3);
    o3l[l] = o3t;
    xo3l[l] = xo3t;
    const float wv = 0.1f * dsg * LV(dq, l, h) * dp / GA * sqrtf(273.0f / LV(dt, l, h)) * T(sigma)[l] * dp / 100000.0f;
    wvt = wvt + wv;
    ywvt = cs * (ywvt + zm * wv) + (1.0f - cs) * (ywvt + ZBETTA * wv);
    wvl[l] = wvt;
    ywvl[l] = ywvt;
    cs = cs * (1.0f - cc[l]);
  }
  float ta1 = 1, ta1s = 1, ra1 = 0, ra1s = 0, ta2 = 1, ta2s = 1, ra2 = 0, ra2s = 0;
  float to3t = 1.0f;
  float to3tu = 1.0f - ozone_absorption(M, xo3t + ZMBAR * o3t);
  float twvt = 1.0f;
  float twvtu = 1.0f - vapour_absorption(M, ywvt + ZBETTA * wvt);
  const float ccn = cc[NL - 1];
  const float scf = K(rcoeff) * dp / 101100.0f * 9.80665f / GA;
  const float diffuseRayleigh = 1.0f - expf(scf * K(ln_diffuse_rayleigh));
  const float rcsuSurface = diffuseRayleigh * (1.0f - ccn);
  const float rcsSurface =
      (1.0f - expf(scf * logf(1.0f - 0.219f / (1.0f + 0.816f * mu0)))) * cs + diffuseRayleigh * (1.0f - cs - ccn);
  float t1[NL + 1], t2[NL + 1], r1s[NL + 1], r2s[NL + 1];
  float tb1[NL], tb1u[NL], rb1[NL], rb1s[NL], tb2[NL], tb2u[NL], rb2[NL], rb2s[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float c = cc[l];
    t1[l] = ta1;
    t2[l] = ta2;
    r1s[l] = ra1s;
    r2s[l] = ra2s;
    const bool surface = l == NL - 1;
    rb1[l] = (surface ? rcsSurface : 0.0f) + rcl1[l] * c;
    rb1s[l] = (surface ? rcsuSurface : 0.0f) + rcl1s[l] * c;
    float o3 = xo3l[l];
    const float to3 = (1.0f - ozone_absorption(M, o3)) / to3t;
    to3t = to3t * to3;
    o3 = xo3t + ZMBAR * (o3t - o3l[l]);
    const float to3u = to3tu / (1.0f - ozone_absorption(M, o3));
    to3tu = to3tu / to3u;
    tb1[l] = 1.0f - (1.0f - to3) * (1.0f - c) - rb1[l];
    tb1u[l] = 1.0f - (1.0f - to3u) * (1.0f - c) - rb1s[l];
    const float m1 = 1.0f / (1.0f - ra1s * rb1s[l]);
    ra1 = ra1 + ta1 * rb1[l] * ta1s * m1;
    ta1 = ta1 * tb1[l] * m1;
    ra1s = rb1s[l] + tb1u[l] * ra1s * tb1[l] * m1;
    ta1s = tb1u[l] * ta1s * m1;
    rb2[l] = rcl2[l] * c;
    rb2s[l] = rcl2s[l] * c;
    float wv = ywvl[l];
    const float twv = (1.0f - vapour_absorption(M, wv)) / twvt;
    twvt = twvt * twv;
    wv = ywvt + ZBETTA * (wvt - wvl[l]);
    const float twvu = twvtu / (1.0f - vapour_absorption(M, wv));
    twvtu = twvtu / twvu;
    tb2[l] = 1.0f - (1.0f - twv) * (1.0f - c) - (1.0f - tcl2[l]) * c;
    tb2u[l] = 1.0f - (1.0f - twvu) * (1.0f - c) - (1.0f - tcl2s[l]) * c;
    const float m2 = 1.0f / (1.0f - ra2s * rb2s[l]);
    ra2 = ra2 + ta2 * rb2[l] * ta2s * m2;
    ta2 = ta2 * tb2[l] * m2;
    ra2s = rb2s[l] + tb2u[l] * ra2s * tb2[l] * m2;
    ta2s = tb2u[l] * ta2s * m2;
  }
  t1[NL] = ta1;
  t2[NL] = ta2;
  r1s[NL] = ra1s;
  r2s[NL] = ra2s;
  const float dalb = F(dalb)[h], salb1 = F(dsalb)[h], salb2 = F(dsalb)[NH + h];
  const bool nst = K(nstartemp) != 0;
  ra1s = nst ? salb1 : dalb;
  ra2s = nst ? salb2 : dalb;
  const float dls = F(dls)[h], dicec = F(dicec)[h];
  const float water = (1.0f - dls) * (1.0f - dicec) * fminf(0.05f / (mu0 + 0.15f), 0.15f);
  const float nalb1 = dls * salb1 + (1.0f - dls) * dicec * salb1 + water;
  const float nalb2 = dls * salb2 + (1.0f - dls) * dicec * salb2 + water;
  F(dsalb)[h] = nalb1;
  F(dsalb)[NH + h] = nalb2;
  const float albedo = nst ? K(zsolars1) * nalb1 + K(zsolars2) * nalb2 : dls * dalb + (1.0f - dls) * dicec * dalb + water;
  F(dalb)[h] = albedo;
  ra1 = nst ? nalb1 : albedo;
  ra2 = nst ? nalb2 : albedo;
  float rl1[NL + 1], rl2[NL + 1], rl1s[NL + 1], rl2s[NL + 1];
#pragma unroll
  for (int l = NL - 1; l >= 0; l--) {
    rl1[l + 1] = ra1;
    rl2[l + 1] = ra2;
    rl1s[l + 1] = ra1s;
    rl2s[l + 1] = ra2s;
    const float d1 = 1.0f - ra1s * rb1s[l];
    ra1 = rb1[l] + tb1[l] * ra1 * tb1u[l] / d1;
    ra1s = rb1s[l] + tb1u[l] * ra1s * tb1u[l] / d1;
    const float d2 = 1.0f - ra2s * rb2s[l];
    ra2 = rb2[l] + tb2[l] * ra2 * tb2u[l] / d2;
    ra2s = rb2s[l] + tb2u[l] * ra2s * tb2u[l] / d2;
  }
  rl1[0] = ra1;
  rl2[0] = ra2;
  rl1s[0] = ra1s;
  rl2s[0] = ra2s;
#pragma unroll
  for (int l = 0; l <= NL; l++) {
    const float m1 = 1.0f / (1.0f - r1s[l] * rl1s[l]);
    const float m2 = 1.0f / (1.0f - r2s[l] * rl2s[l]);
    const float fu = -t1[l] * rl1[l] * m1 * ftop1 + -t2[l] * rl2[l] * m2 * ftop2;
    const float fd = t1[l] * m1 * ftop1 + t2[l] * m2 * ftop2;
    LV(dfu, l, h) = fu;
    LV(dfd, l, h) = fd;
    LV(dswfl, l, h) = fu + fd;
  }
}

template <int NL>
__device__ __forceinline__ void lwr(const DevMember& M, int h) {
  // physics/radiation/wgsl/lwr.wgsl:36-179
  const int NH = M.nhor;
  const float ZERO = 1.0e-6f, ZFH2O = 0.1f, ZFO3 = 100.0f / 2.14f;
  const float* sg = T(sigma);
  const float* sh = T(sigmah);
  const float dp = F(dp)[h], ps2 = dp * dp;
  float st4[NL + 1], st4h[NL + 1], bu[NL + 2], bd[NL + 2];
#pragma unroll
  for (int l = 0; l <= NL; l++) {
    const float t = LV(dt, l, h), t2 = t * t;
    st4[l] = K(sbk) * (t2 * t2);
  }
  st4h[0] = st4[0] - (st4[0] - st4[1]) * sg[0] / (sg[0] - sg[1]);
#pragma unroll
  for (int l = 1; l < NL; l++)
    st4h[l] = (st4[l] * (sg[l - 1] - sh[l - 1]) + st4[l - 1] * (sh[l - 1] - sg[l])) / (sg[l - 1] - sg[l]);
  const int n = NL - 1;
  if ((st4[n] - st4h[n]) * (st4[n] - st4[NL]) > 0)
    st4h[NL] = st4[NL];
  else
    st4h[NL] = st4[n] + (st4[n] - st4[n - 1]) * (1.0f - sg[n]) / (sg[n] - sg[n - 1]);
  bd[0] = 0;
  const float dls = F(dls)[h];
  const float eps = dls + 0.98f * (1.0f - dls);
  bu[NL + 1] = eps * st4[NL];
  float q[NL], qo3[NL], qco2[NL], taucc0[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float sfac = T(lwr_zzf1)[l] * ps2;
    q[l] = ZFH2O * sfac * LV(dq, l, h);
    qo3[l] = ZFO3 * sfac * LV(dqo3, l, h);
    qco2[l] = K(zfco2_zzf2) * sfac * LV(dqco2, l, h);
    taucc0[l] = 1.0f - LV(dcc, l, h) * (1.0f - expf(T(lwr_zzf3)[l] * LV(dql, l, h) * dp));
  }
  float dftu[NL + 1], dftd[NL + 1], taucs[NL], tau[NL], tausf[NL];
#pragma unroll
  for (int l = 0; l <= NL; l++) dftu[l] = dftd[l] = 0;
  for (int j = 0; j < NL; j++) {
    float taucc = 1.0f, sumwv = 0, sumo3 = 0, sumco2 = 0;
    for (int k = j; k < NL; k++) {
      sumwv = sumwv + q[k];
      sumo3 = sumo3 + qo3[k];
      sumco2 = sumco2 + qco2[k];
      // the column amounts are >= 0, so every argument below is >= 3.59e-5
      float ah2o;
      if (sumwv <= 0.01f)
        ah2o = 0.846f * pow_pos(sumwv + 3.59e-5f, 0.243f) - K(zh2o0a);
      else
        ah2o = 0.24f * log10_pos(sumwv + 0.01f) + K(zah2oc);
      if (K(th2oc) > 0) ah2o = fminf(ah2o + (1.0f - expf(-K(th2oc) * sumwv)), 1.0f);
      float aco2;
      if (sumco2 <= 1.0f)
        aco2 = 0.0676f * pow_pos(sumco2 + 0.01022f, 0.421f) - K(zco20);
      else
        aco2 = 0.0546f * log10_pos(sumco2) + K(zaco2c);
      float th2o;
      if (sumwv <= 2.0f)
        th2o = 1.0f - (0.832f * pow_pos(sumwv + 0.0286f, 0.26f) - K(zh2o0));
      else
        th2o = fmaxf(0.0f, K(zth2oc) - 0.1196f * log_pos(sumwv - 0.6931f));
      float ao3;
      if (sumo3 <= 0.01f)
        ao3 = 0.209f * pow_pos(sumo3 + 7.0e-5f, 0.436f) - K(zao30);
      else
        ao3 = 0.0212f * log10_pos(sumo3) + K(zao3c);
      const float cs = fminf(1.0f - ZERO, fmaxf(ZERO, 1.0f - ah2o - ao3 - aco2 * th2o));
      taucs[k] = cs;
      taucc = taucc * taucc0[k];
      tau[k] = cs * taucc;
    }
    if (j == 0) {
      float tau0 = 1.0f - ZERO;
      for (int k = 0; k < NL; k++) {
        tau0 = taucs[k] / tau0 * taucc0[k] * K(tpofmt);
        tau0 = fminf(1.0f - ZERO, fmaxf(ZERO, tau0));
        if ((st4[k] - st4h[k]) * (st4[k] - st4h[k + 1]) > 0) {
          bd[k + 1] = 0.5f * st4[k] + 0.25f * (st4h[k] + st4h[k + 1]);
          bu[k + 1] = bd[k + 1];
        } else {
          bd[k + 1] = (st4h[k + 1] - tau0 * st4h[k]) / (1.0f - tau0) - (st4h[k] - st4h[k + 1]) / logf(tau0);
          bu[k + 1] = st4h[k] + st4h[k + 1] - bd[k + 1];
        }
        tau0 = taucs[k];
      }
    }
    dftu[j] = dftu[j] - bu[j + 1];
    dftd[j] = dftd[j] + bd[j];
    const float bdl = bd[j] - bd[j + 1];
    for (int k = j; k < NL; k++) {
      dftu[j] = dftu[j] - (bu[k + 2] - bu[k + 1]) * tau[k];
      dftd[k + 1] = dftd[k + 1] + bdl * tau[k];
    }
    tausf[j] = tau[NL - 1];
  }
  dftu[NL] = dftu[NL] - bu[NL + 1];
  dftd[NL] = dftd[NL] + bd[NL];
  const float reflected = (1.0f - eps) * dftd[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) dftu[l] = dftu[l] - tausf[l] * reflected;
#pragma unroll
  for (int l = 0; l <= NL; l++) {
    LV(dftu, l, h) = dftu[l];
    LV(dftd, l, h) = dftd[l];
    LV(dlwfl, l, h) = dftu[l] + dftd[l];
  }
}

template <int NL>
__device__ __forceinline__ void radstep(const DevMember& M, int h) {
  // physics/radiation/wgsl/radstep.wgsl:11-32
  const int NH = M.nhor;
  const float GA = K(ga), GASCON = K(gascon);
  float flux[NL + 1];
#pragma unroll
  for (int l = 0; l <= NL; l++) {
    flux[l] = LV(dlwfl, l, h) + LV(dswfl, l, h);
    LV(dflux, l, h) = flux[l];
  }
  const float dp = F(dp)[h];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float zdh = -LV(dt, l, h) * GASCON / GA * T(rad_lnsig)[l];
    const float den = T(dsigma)[l] * dp * K(acpd) * (1.0f + K(adv) * LV(dq, l, h));
    LV(dtdt, l, h) = LV(dtdt, l, h) - GA * (flux[l + 1] - flux[l]) / den;
    LV(dtdtswr, l, h) = -GA * (LV(dswfl, l + 1, h) - LV(dswfl, l, h)) / den;
    LV(dtdtlwr, l, h) = -GA * (LV(dlwfl, l + 1, h) - LV(dlwfl, l, h)) / den;
    LV(dconv, l, h) = (flux[l] - flux[l + 1]) / (zdh + 1.0e-9f);
  }
}

// ------------------------------------------------------- precipitation

__device__ __forceinline__ float sat_capped(const DevMember& M, float t, float p) {
  // physics/precipitation/wgsl/saturation.wgsl:15-24
  return fminf(K(rdbrv) * K(ra1) * expf(K(ra2) * (t - K(tmelt)) / (t - K(ra4))) / p, K(rdbrv));
}
__device__ __forceinline__ float sat_divided(const DevMember& M, float t, float p) {
  const float q = sat_capped(M, t, p);
  return q / (1.0f - (1.0f / K(rdbrv) - 1.0f) * q);
}

struct Parcel {
  float t, q, cor;
  bool saturated;
};

__device__ __forceinline__ Parcel condense(const DevMember& M, float t0, float q0, float p) {
  // physics/precipitation/wgsl/kuo.wgsl:22-47
  const float MOIST = 1.0f / K(rdbrv) - 1.0f;
  float s = sat_capped(M, t0, p);
  float cor = 1.0f / (1.0f - MOIST * s);
  s = s * cor;
  if (!(s < q0)) return Parcel{t0, q0, cor, false};
  float t = t0, q = q0;
  const float lcp = (t0 < K(tmelt) ? K(als) : K(alv)) / (K(acpd) * (1.0f + K(adv) * q0));
  for (int i = 0; i < 2; i++) {
    if (i == 1) {
      s = sat_capped(M, t, p);
      cor = 1.0f / (1.0f - MOIST * s);
      s = s * cor;
    }
    const float dqsdt = K(ra2) * (K(tmelt) - K(ra4)) * s * cor / ((t - K(ra4)) * (t - K(ra4)));
    const float rain = (q - s) / (1.0f + lcp * dqsdt);
    t = t + lcp * rain;
    q = q - rain;
  }
  return Parcel{t, q, cor, true};
}

template <int NL>
__device__ __forceinline__ void kuo(const DevMember& M, int h) {
  // physics/precipitation/wgsl/kuo.wgsl:53-298
  const int NH = M.nhor, NLEM = NL - 1;
  const float DELTSEC2 = K(deltsec2), ALV = K(alv), ACPD = K(acpd), ADV = K(adv);
  const float MOIST = 1.0f / K(rdbrv) - 1.0f;
  const float *ds = T(dsigma), *sg = T(sigma), *skap = T(vdiff_skap), *zf = T(kuo_zfac);
  const float dp = F(dp)[h];
  float zdqt[NL], zqe[NL], zte[NL], zdqtot[NL], zalcpe[NL], zacpe[NL], zdqdt[NL], zdtdt[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    zdqt[l] = LV(dqt, l, h) * K(psurf) * K(ww) / dp;
    LV(dqt, l, h) = zdqt[l];
    const float dq = LV(dq, l, h), dqdt = LV(dqdt, l, h);
    zdqtot[l] = (zdqt[l] + dqdt) * DELTSEC2;
    zqe[l] = dq + dqdt * DELTSEC2;
    zte[l] = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
    zacpe[l] = ACPD * (1.0f + ADV * dq);
    zalcpe[l] = ALV / zacpe[l];
    zdqdt[l] = 0;
    zdtdt[l] = 0;
  }
  int ilift = -1, itop = -1;
  float zi = 0, zpt = 0, zpq = 0, zbeta = 0, zpbeta = 0, ztp = 0, zqp = 0;
  const int n = NL - 1;
  const float dsn = ds[n];
  const float ztsurf = LV(dt, NL, h) * skap[n];
  if (zdqtot[n] > 0 && ztsurf > zte[n]) {
    const float p = dp * sg[n];
    const float zqsnl = sat_divided(M, zte[n], p);
    const float zqnew = LV(dq, NL, h) * fminf(1.0f, zqe[n] / zqsnl);
    const Parcel parcel = condense(M, ztsurf, zqnew, p);
    if (parcel.saturated) {
      ilift = NL + 1;
      itop = NL;
      ztp = parcel.t;
      zqp = parcel.q;
      const float zqsa = sat_divided(M, zte[n], p);
      zdqdt[n] = fmaxf(zqsa - zqe[n], 0.0f);
      zdtdt[n] = ztp - zte[n];
      zpt = zdtdt[n] * dsn * zacpe[n];
      zpq = zdqdt[n] * dsn * ALV;
      zbeta = fminf(1.0f, zqe[n] / zqsa) * dsn;
      zpbeta = dsn;
      zi = zdqtot[n] * dsn * ALV;
    }
  }
  for (int l = NLEM - 1; l >= 0; l--) {
    const int lp = l + 1;
    if (ilift == -1 && zdqtot[lp] > 0) {
      const float p = dp * sg[l];
      const Parcel parcel = condense(M, zte[lp] * zf[l], zqe[lp], p);
      if (parcel.saturated && parcel.t > zte[l]) {
        const float d = ds[l], dspp = ds[lp];
        ilift = l + 2;
        itop = l + 1;
        ztp = parcel.t;
        zqp = parcel.q;
        const float zqsa = sat_divided(M, zte[l], p);
        zdqdt[l] = fmaxf(zqsa - zqe[l], 0.0f);
        zdtdt[l] = ztp - zte[l];
        zpt = zdtdt[l] * d * zacpe[l];
        zpq = zdqdt[l] * d * ALV;
        zbeta = fminf(1.0f, zqe[l] / zqsa) * d;
        zpbeta = d;
        const float pp = dp * sg[lp];
        float zqsnl = sat_divided(M, zte[lp], pp);
        if (zqe[lp] > zqsnl) {
          float t = zte[lp], q = zqe[lp];
          const float lcp = (t < K(tmelt) ? K(als) : ALV) / (ACPD * (1.0f + ADV * q));
          float cor = parcel.cor;
          for (int i = 0; i < 2; i++) {
            if (i == 1) {
              zqsnl = sat_capped(M, t, pp);
              cor = 1.0f / (1.0f - MOIST * zqsnl);
              zqsnl = zqsnl * cor;
            }
            const float dqsdt = K(ra2) * (K(tmelt) - K(ra4)) * zqsnl * cor / ((t - K(ra4)) * (t - K(ra4)));
            const float rain = (q - zqsnl) / (1.0f + lcp * dqsdt);
            t = t + lcp * rain;
            q = q - rain;
          }
          zdqdt[lp] = fmaxf(q - zqe[lp], 0.0f);
          zdtdt[lp] = t - zte[lp];
          zpt = zdtdt[lp] * dspp * zacpe[lp] + zpt;
          zpq = zdqdt[lp] * dspp * ALV + zpq;
          zbeta = fminf(1.0f, zqe[lp] / zqsnl) * dspp + zbeta;
          zpbeta = dspp + zpbeta;
        }
        zi = zdqtot[l] * d * ALV + zdqtot[lp] * dspp * ALV;
      }
    }
  }
  for (int l = NLEM - 1; l >= 0; l--) {
    if (itop == l + 2) {
      const float p = dp * sg[l];
      const Parcel parcel = condense(M, ztp * zf[l], zqp, p);
      if (parcel.saturated && parcel.t > zte[l]) {
        const float d = ds[l];
        itop = l + 1;
        ztp = parcel.t;
        zqp = parcel.q;
        const float zqsa = sat_divided(M, zte[l], p);
        zdqdt[l] = fmaxf(zqsa - zqe[l], 0.0f);
        zdtdt[l] = ztp - zte[l];
        zpt = zdtdt[l] * d * zacpe[l] + zpt;
        zpq = zdqdt[l] * d * ALV + zpq;
        zbeta = fminf(1.0f, zqe[l] / zqsa) * d + zbeta;
        zpbeta = zpbeta + d;
        zi = zdqtot[l] * d * ALV + zi;
      }
    }
  }
  zi = fmaxf(0.0f, zi);
  float zrhm = 0;
  if (zpbeta > 0) zrhm = zbeta / zpbeta;
  float beta = 1.0f;
  if (zrhm >= K(rhbeta)) beta = powi((1.0f - zrhm) / (1.0f - K(rhbeta)), (int)K(nbeta));
  float zat = 0, zaq = 0;
  if (zpt > 0) zat = zi * (1.0f - beta) / zpt;
  if (zpq > 0) zaq = zi * beta / zpq;
  bool shallow = false;
  if (itop > 0 && (dp * sg[itop - 1] > K(pdeep) || zi <= 0)) shallow = true;
  float icctot = 0;
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const int level = l + 1;
    float dprcl = 0, icclev = 0;
    if (ilift >= level && itop <= level && zi > 0 && !shallow) {
      LV(dqdt, l, h) = zaq * zdqdt[l] / DELTSEC2 - zdqt[l];
      LV(dtdt, l, h) = zat * zdtdt[l] / DELTSEC2 + LV(dtdt, l, h);
      dprcl = zat * zdtdt[l] * ds[l] / zalcpe[l] / DELTSEC2;
    }
    if (dprcl > 0) {
      icclev = 1;
      icctot = icctot + 1;
    }
    if (sg[l] < 0.4f && icclev == 1.0f && itop == level) icclev = 2;
    LV(dprcl, l, h) = dprcl;
    LV(dprscl, l, h) = 0;
    LV(icclev, l, h) = icclev;
  }
  F(icctot)[h] = icctot;
  if (!shallow) return;
  const float zkonst1 = K(ga) * DELTSEC2 / K(gascon);
  const float zkonst2 = K(rkshallow) * zkonst1 * K(ga) / K(gascon);
  int kktop = 1;
  for (int level = 2; level <= NLEM; level++)
    if (dp * sg[level - 1] < K(pdeepth)) kktop = level;
  const int ktop = itop > kktop ? itop : kktop;
  float zkdiff[NL], skp[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    zkdiff[l] = 0;
    skp[l] = skap[l];
  }
#pragma unroll
  for (int l = 0; l < NLEM; l++) {
    const int lp = l + 1, level = l + 1;
    const float shh = T(sigmah)[l];
    const float zfac = zkonst2 * shh * shh * T(vdiff_rdsig)[l];
    const float d = ds[l], dspp = ds[lp];
    const float zrthl = (dspp + d) / (zte[l] * dspp + zte[lp] * d);
    if (level >= ktop && level < ilift) zkdiff[l] = zfac * zrthl * zrthl;
    if (level + 1 == ktop) {
      float zzqs = sat_capped(M, zte[l], dp * sg[l]);
      zzqs = zzqs * (1.0f / (1.0f - MOIST * zzqs));
      float zzqsp = sat_capped(M, zte[lp], dp * sg[lp]);
      zzqsp = zzqsp * (1.0f / (1.0f - MOIST * zzqsp));
      const float zrh = zqe[l] / zzqs, zrhp = zqe[lp] / zzqsp;
      const float zzf = fabsf((zrh - 0.8f) / 0.2f * (zrhp - zrh));
      zkdiff[l] = zfac * zrthl * zrthl * zzf;
    }
  }
  float zqn[NL], ztn[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) zqn[l] = zqe[l];
  diffuse<NL>(M, zqn, zkdiff);
#pragma unroll
  for (int l = 0; l < NLEM; l++) zkdiff[l] = zkdiff[l] * T(vdiff_skaph)[l];
#pragma unroll
  for (int l = 0; l < NL; l++) ztn[l] = zte[l];
  diffuse_potential<NL>(M, ztn, zkdiff, skp);
#pragma unroll
  for (int l = 0; l < NL; l++) {
    LV(dtdt, l, h) = LV(dtdt, l, h) + (ztn[l] - zte[l]) / DELTSEC2;
    LV(dqdt, l, h) = LV(dqdt, l, h) + (zqn[l] - zqe[l]) / DELTSEC2;
  }
}

template <int NL>
__device__ __forceinline__ void mkdca(const DevMember& M, int h) {
  // physics/precipitation/wgsl/mkdca.wgsl:13-97
  const int NH = M.nhor, NLEM = NL - 1;
  const float DELTSEC2 = K(deltsec2);
  const float *skap = T(vdiff_skap), *ds = T(dsigma);
  float zt[NL], zth[NL], zq[NL], zqn[NL];
  bool mixed[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    zt[l] = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
    zth[l] = zt[l] / skap[l];
    zq[l] = LV(dq, l, h) + LV(dqdt, l, h) * DELTSEC2;
    zqn[l] = zq[l];
    mixed[l] = false;
  }
  float limited = 0;
  int passes = 0;
  for (;;) {
    bool unstable = false;
    int ibase = 0, itop = NL;
    float ztht = 0, zqt = 0, zsum1 = 0, zsum2 = 0, zsumq1 = 0, zsumq2 = 0;
#pragma unroll
    for (int l = 0; l < NLEM; l++) {
      const int lp = l + 1;
      if (zth[l] < zth[lp]) {
        ibase = l + 2;
        itop = l;
        zsum1 = skap[lp] * ds[lp] + skap[l] * ds[l];
        zsum2 = skap[lp] * ds[lp] * zth[lp] + skap[l] * ds[l] * zth[l];
        zsumq1 = ds[lp] + ds[l];
        zsumq2 = ds[lp] * zqn[lp] + ds[l] * zqn[l];
        ztht = zsum2 / zsum1;
        zqt = zsumq2 / zsumq1;
        unstable = true;
      }
    }
    if (!unstable) break;
#pragma unroll
    for (int level = NLEM - 1; level >= 1; level--) {
      const int l = level - 1;
      if (itop == level && zth[l] < ztht) {
        itop = level - 1;
        zsum1 = zsum1 + skap[l] * ds[l];
        zsum2 = zsum2 + skap[l] * ds[l] * zth[l];
        zsumq1 = zsumq1 + ds[l];
        zsumq2 = zsumq2 + ds[l] * zqn[l];
        ztht = zsum2 / zsum1;
        zqt = zsumq2 / zsumq1;
      }
    }
#pragma unroll
    for (int l = 0; l < NL; l++) {
      const int level = l + 1;
      if (level > itop && level <= ibase) {
        zth[l] = ztht;
        zqn[l] = zqt;
        mixed[l] = true;
      }
    }
    passes = passes + 1;
    if (passes > 2 * NL) {
      limited = 1;
      break;
    }
  }
#pragma unroll
  for (int l = 0; l < NL; l++) {
    LV(dca_dtdt, l, h) = mixed[l] ? (zth[l] * skap[l] - zt[l]) / DELTSEC2 : 0.0f;
    LV(dca_dqdt, l, h) = (zqn[l] - zq[l]) / DELTSEC2;
  }
  F(dca_limited)[h] = limited;
}

__device__ __forceinline__ float evaporate(const DevMember& M, float* flux, float deficit, float zlcpe, float zqsat,
                                           float zcor, float zt, float ds) {
  // physics/precipitation/wgsl/mkrain.wgsl:8-15
  const float zdqdt = fminf(K(gamma) * deficit * ds / K(deltsec2) /
                                (1.0f + zlcpe * K(ra2) * (K(tmelt) - K(ra4)) * zqsat * zcor / ((zt - K(ra4)) * (zt - K(ra4)))),
                            *flux);
  *flux = *flux - zdqdt;
  return zdqdt / ds;
}


template <int NL>
__device__ __forceinline__ void mkclouds(const DevMember& M, int h) {
  // physics/precipitation/wgsl/mkclouds.wgsl:18-86
  const int NH = M.nhor;
  const float ZCLMAX = 1.0f, ZCCA = 0.245f, ZCCB = 0.125f, ZCCMAX = 0.8f, ZCCMIN = 0.05f;
  const float DELTSEC2 = K(deltsec2), GA = K(ga), GASCON = K(gascon);
  const float dp = F(dp)[h];
  float cover[NL], zzf[NL];
  float zcctot = ZCCMIN, zccconv = ZCCMIN;
  const float dprc = F(dprc)[h];
  if (dprc > 0) {
    zcctot = ZCCA + ZCCB * logf(dprc * K(zrfac));
    zcctot = fminf(ZCCMAX, fmaxf(ZCCMIN, zcctot));
    const float levels = F(icctot)[h];
    zccconv = levels == 0.0f ? 1.0f : 1.0f - powf(1.0f - zcctot, 1.0f / levels);
  }
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float zt = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
    const float zq = LV(dq, l, h) + LV(dqdt, l, h) * DELTSEC2;
    const float zqsat = sat_divided(M, zt, dp * T(sigma)[l]);
    float zrh = zq / zqsat;
    const float icclev = LV(icclev, l, h);
    float dcc = 0;
    if (icclev > 0) dcc = zccconv;
    if (icclev == 2.0f && zcctot >= 0.3f) dcc = fminf(dcc + 2.0f * (zcctot - 0.3f), ZCLMAX);
    zrh = zrh * (1.0f - dcc);
    const float zrcrit = T(rcrit)[l];
    float zwfac = 1.0f;
    if (T(sigma)[l] > 0.8f && K(clwfac) > 0) {
      const float dw = LV(dw, l, h);
      zwfac = K(clwfac) * fmaxf(0.0f, K(clwcrit1) - dw);
      if (K(clwcrit2) > dw) zwfac = 1.0f;
    }
    const float excess = fmaxf(0.0f, (zrh - zrcrit) / (1.0f - zrcrit));
    cover[l] = fminf(dcc + zwfac * (excess * excess), ZCLMAX);
  }
  float dqvi = 0, zzh = 0;
#pragma unroll
  for (int l = NL - 1; l >= 1; l--) {
    dqvi = LV(dq, l, h) * T(dsigma)[l] + dqvi;
    const float zdh = -LV(dt, l, h) * GASCON / GA * T(rad_lnsig)[l];
    zzf[l] = zzh + zdh * 0.5f;
    zzh = zzh + zdh;
  }
  dqvi = LV(dq, 0, h) * T(dsigma)[0] + dqvi;
  zzf[0] = zzh - LV(dt, 0, h) * GASCON / GA * T(rad_lnsig)[0];
  dqvi = dqvi * dp / GA;
  F(dqvi)[h] = dqvi;
  zzh = 700.0f * logf(1.0f + dqvi);
  float clear = 1.0f;
#pragma unroll
  for (int l = 0; l < NL; l++) {
    float dql = 0;
    if (zzh > 0 && cover[l] > 0) {
      dql = 0.00021f * expf(-zzf[l] / zzh) * GASCON * LV(dt, l, h) / (T(sigma)[l] * dp);
      dql = fmaxf(dql, 1.0e-9f);
    }
    LV(dql, l, h) = dql;
    LV(dcc, l, h) = cover[l];
    clear = clear * (1.0f - cover[l]);
  }
  LV(dql, NL, h) = 0;
  LV(dcc, NL, h) = 1.0f - clear;
}

// ----------------------------------------------------------------- land

struct SnowLayer {
  float ctop, soilz1, diff1, cap1;
};

__device__ __forceinline__ SnowLayer snow_layer(const DevMember& M, float snowz, float cap0, float diff0, float soilz0) {
  // physics/land/wgsl/soil.wgsl:63-77
  SnowLayer r;
  const float zsntop = fminf(snowz * 1000.0f / K(rhosnow), K(dztop));
  r.ctop = cap0 * K(snowcap) * K(dztop) / (cap0 * zsntop + K(snowcap) * (K(dztop) - zsntop));
  const float zsnowz = fminf(snowz * 1000.0f / K(rhosnow), 1.0f) - zsntop;
  r.soilz1 = soilz0 + zsnowz;
  r.diff1 = diff0 * K(snowdiff) * r.soilz1 / (K(snowdiff) * soilz0 + diff0 * zsnowz);
  r.cap1 = cap0 * K(snowcap) * r.soilz1 / (cap0 * zsnowz + K(snowcap) * soilz0);
  return r;
}

__device__ __forceinline__ float surface_temperature(const DevMember& M, SnowLayer layer, float dtsm, float flux, float soil1) {
  // physics/land/wgsl/soil.wgsl:84-91
  const float conduction = 2.0f * layer.diff1;
  return (layer.ctop * K(dztop) * dtsm / K(deltsec) + flux + conduction * soil1 / layer.soilz1) /
         (layer.ctop * K(dztop) / K(deltsec) + conduction / layer.soilz1);
}

__device__ __forceinline__ void mktsoil(const DevMember& M, int h, float ftop, const float* soilz, const float* cap,
                        const float* diff) {
  // physics/land/wgsl/soil.wgsl:23-55
  const int NH = M.nhor, NS = 5, n = NS - 1;
  const float DELTSEC = K(deltsec);
  float zdiff[5], zcap[5], zebs[5], ztn[5];
  zdiff[n - 1] = 2.0f * diff[n] * diff[n - 1] / (diff[n] * soilz[n - 1] + diff[n - 1] * soilz[n]);
  zcap[n] = cap[n] * soilz[n] / DELTSEC;
  zebs[n] = 1.0f / (zcap[n] + zdiff[n - 1]);
  ztn[n] = zcap[n] * LV(dsoilt, n, h) * zebs[n];
#pragma unroll
  for (int l = n - 1; l >= 1; l--) {
    zdiff[l - 1] = 2.0f * diff[l] * diff[l - 1] / (diff[l] * soilz[l - 1] + diff[l - 1] * soilz[l]);
    zcap[l] = cap[l] * soilz[l] / DELTSEC;
    zebs[l] = 1.0f / (zcap[l] + zdiff[l - 1] + zdiff[l] * (1.0f - zdiff[l] * zebs[l + 1]));
    ztn[l] = zebs[l] * (zcap[l] * LV(dsoilt, l, h) + zdiff[l] * ztn[l + 1]);
  }
  zcap[0] = cap[0] * soilz[0] / DELTSEC;
  zebs[0] = 1.0f / (zcap[0] + zdiff[0] * (1.0f - zdiff[0] * zebs[1]));
  ztn[0] = zebs[0] * (zcap[0] * LV(dsoilt, 0, h) + zdiff[0] * ztn[1] + ftop);
#pragma unroll
  for (int l = 1; l < NS; l++) ztn[l] = ztn[l] + zebs[l] * zdiff[l - 1] * ztn[l - 1];
  const bool glacier = F(dglac)[h] > 0.5f;
#pragma unroll
  for (int l = 0; l < NS; l++) LV(dsoilt, l, h) = glacier ? fminf(ztn[l], K(tmelt)) : ztn[l];
}

template <int NL>
__device__ __forceinline__ void soil(const DevMember& M, int h) {
  // physics/land/wgsl/soil.wgsl:95-185
  const int NH = M.nhor;
  const float DELTSEC = K(deltsec), TMELT = K(tmelt), ALS = K(als), ALV = K(alv);
  const float dls = F(dls)[h];
  const bool land = dls > 0;
  const float snowStart = F(dsnowz)[h];
  float dwater = 0, dsmelt = 0;
  if (snowStart == 0.0f) F(dsnowt)[h] = TMELT;
  if (land) {
    const float glac = F(dglac)[h];
    float soilz[5], cap[5], diff[5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      cap[l] = K(sicecap) * glac + K(soilcap) * (1.0f - glac);
      diff[l] = K(sicediff) * glac + K(soildiff) * (1.0f - glac);
      soilz[l] = T(dsoilz)[l];
    }
    const float zhfla = F(dshfl)[h] + F(dlhfl)[h] + LV(dflux, NL, h);
    const float devap = F(devap)[h];
    float zdsnowz = 0;
    if (F(dprs)[h] > 0) zdsnowz = F(dprs)[h];
    if (snowStart > 0) zdsnowz = zdsnowz + devap;
    float zsnowz = fmaxf(0.0f, snowStart + zdsnowz * DELTSEC);
    zdsnowz = (zsnowz - snowStart) / DELTSEC;
    dwater = devap + F(dprl)[h] + F(dprc)[h] - zdsnowz;
    float dsnowz = zsnowz;
    const float dtsm = F(land_dtsm)[h];
    const float soil1 = LV(dsoilt, 0, h);
    SnowLayer layer = snow_layer(M, dsnowz, cap[0], diff[0], soilz[0]);
    float dts = surface_temperature(M, layer, dtsm, zhfla, soil1);
    float zhfls = 2.0f * layer.diff1 / layer.soilz1 * (dts - soil1);
    float zhflm = 0;
    if (dsnowz > 0 && dts > TMELT) {
      dts = TMELT;
      zhflm = fmaxf(0.0f, zhfla - layer.ctop * K(dztop) * (dts - dtsm) / DELTSEC +
                              2.0f * layer.diff1 * (soil1 - dts) / layer.soilz1);
      zsnowz = fmaxf(0.0f, dsnowz - zhflm * DELTSEC / ((ALS - ALV) * 1000.0f));
      zdsnowz = (zsnowz - dsnowz) / DELTSEC;
      zhflm = -zdsnowz * 1000.0f * (ALS - ALV);
      dsnowz = zsnowz;
      dsmelt = -zdsnowz;
      zhfls = 2.0f * layer.diff1 / layer.soilz1 * (dts - soil1);
      dwater = dwater - zdsnowz;
      layer = snow_layer(M, dsnowz, cap[0], diff[0], soilz[0]);
    }
    if (dsnowz == 0.0f && zhflm > 0) {
      dts = surface_temperature(M, layer, dtsm, zhfla - zhflm, soil1);
      zhfls = 2.0f * layer.diff1 / layer.soilz1 * (dts - soil1);
    }
    soilz[0] = layer.soilz1;
    cap[0] = layer.cap1;
    diff[0] = layer.diff1;
    mktsoil(M, h, zhfls, soilz, cap, diff);
    if (dsnowz > 0) F(dsnowt)[h] = dts;
    if (K(dsmax) > 0 && dsnowz > K(dsmax)) {
      zdsnowz = (K(dsmax) - dsnowz) / DELTSEC;
      dsnowz = K(dsmax);
      dwater = dwater - zdsnowz;
      dts = dts + zdsnowz * 1000.0f * (ALS - ALV) * DELTSEC / layer.ctop / K(dztop);
      dsmelt = dsmelt - zdsnowz;
    }
    F(dsnowz)[h] = dsnowz;
    F(land_dts)[h] = dts;
    F(dsndch)[h] = dsnowz - snowStart;
  } else {
    F(dsndch)[h] = snowStart;
  }
  F(dwater)[h] = dwater;
  F(dsmelt)[h] = dsmelt;
  float drunoff = 0;
  if (land) {
    const float dwatc = F(dwatc)[h] + DELTSEC * dwater;
    const float dwmax = F(dwmax)[h];
    drunoff = fmaxf(0.0f, dwatc - dwmax) / DELTSEC;
    F(dwatc)[h] = fmaxf(fminf(dwmax, dwatc), 0.0f);
  }
  F(drunoff)[h] = drunoff;
}

__device__ __forceinline__ float river(const DevMember& M, int h) {
  return F(driver)[h] * F(darea)[h] * F(dls)[h];
}

__device__ __forceinline__ void eastward(const DevMember& M, int row, int lon, float* out, float* into) {
  // physics/land/wgsl/mkradv.wgsl:13-22
  const int h = row + lon;
  const int next = row + (lon + 1) % M.nlon;
  const float u = F(duroff)[h];
  const float water = u > 0 ? river(M, h) : river(M, next);
  *out = u * water / F(darea)[h];
  *into = u * water / F(darea)[next];
}

__device__ __forceinline__ void mkradv(const DevMember& M, int h) {
  // physics/land/wgsl/mkradv.wgsl:24-59; each lane writes only its own
  // drunoff and reads only driver, so the update needs no staging.
  const int NLON = M.nlon;
  const int lat = h / NLON, lon = h % NLON, row = lat * NLON;
  float runoff = F(drunoff)[h], o, in;
  if (lon == 0) {
    eastward(M, row, 0, &o, &in);
    runoff = runoff - o;
    eastward(M, row, NLON - 1, &o, &in);
    runoff = runoff + in;
  } else {
    eastward(M, row, lon - 1, &o, &in);
    runoff = runoff + in;
    eastward(M, row, lon, &o, &in);
    runoff = runoff - o;
  }
  const float area = F(darea)[h];
  const float v = F(dvroff)[h];
  const float below = lat + 1 < M.nlat ? river(M, h + NLON) : 0.0f;
  if (v < 0)
    runoff = runoff + v * river(M, h) / area;
  else
    runoff = runoff + v * below / area;
  float fromAbove = 0;
  if (lat > 0) {
    const int above = h - NLON;
    const float va = F(dvroff)[above];
    fromAbove = va < 0 ? -va * river(M, above) : -va * river(M, h);
  }
  F(drunoff)[h] = runoff + fromAbove / area;
}

// momint of a 0..13 climatology; offset 0 for nstep, 3 for nstep + 1
// (sea-ice/wgsl/ice.wgsl:21-28, land/wgsl/surface.wgsl:43-49).
__device__ __forceinline__ float clim_at(const DevMember& M, const float* field, int h, int offset) {
  const float* ss = F(surface_step);
  const float w = ss[offset + 2];
  return (1.0f - w) * field[(int)ss[offset] * M.nhor + h] + w * field[(int)ss[offset + 1] * M.nhor + h];
}

__device__ __forceinline__ float snow_albedo(const DevMember& M, float forest, float maxF, float maxB, float minF,
                                             float minB, float ts) {
  const float zalbmax = forest * maxF + (1.0f - forest) * maxB;
  const float zalbmin = forest * minF + (1.0f - forest) * minB;
  const float zdalb = (zalbmax - zalbmin) * (ts - 263.16f) / (K(tmelt) - 263.16f);
  return fmaxf(zalbmin, fminf(zalbmax, zalbmax - zdalb));
}

__device__ __forceinline__ float glacier_albedo(const DevMember& M, float mx, float mn, float ts) {
  return fmaxf(mn, fminf(mx, mx - (mx - mn) * (ts - 263.16f) / (K(tmelt) - 263.16f)));
}

// simba's vegstep for one land column below 90 % glacier (physics/
// vegetation/wgsl/vegstep.wgsl:31-99): light- and water-limited primary
// production, vegetation and soil carbon, forest and soil cover; with
// nveg = 2 also roughness, soil water capacity, surface humidity factor,
// albedo and forest cover.
template <int NL>
__device__ __forceinline__ void vegstep(const DevMember& M, int h) {
  if (!(F(dls)[h] > 0.0f && F(dglac)[h] < 0.90f)) return;
  const int NH = M.nhor;
  // simba's fixed parameters (setup/index.ts:263-282)
  constexpr float VWS_CRIT = 0.25f, Q10 = 2.0f, ZLAIMAX = 9.0f, ZLAIMIN = 0.1f, VALB_MIN = 0.12f;
  constexpr float VALB_MAX = 0.3f, VSALB_MIN = 0.2f, VWMAX_MAX = 0.5f, VWMAX_MIN = 0.05f, VZ0_MIN = 0.05f;
  constexpr float CVEG_A = 2.9f, CVEG_C = 1.570796326794897f, CVEG_F = 3.0f, CVEG_K = 0.5f, CVEG_L = 0.25f;
  constexpr float CT_CRIT = 5.0f, CO2_REF = 360.0f, CO2_COMP = 0.0f, CO2_SENS = 0.3f;
  const float co2 = F(co2)[0];
  float zbeta = 0.0f;
  if (!(co2 < 1.0e-8f)) zbeta = fmaxf(0.0f, 1.0f + CO2_SENS * logf((co2 - CO2_COMP) / (CO2_REF - CO2_COMP)));
  const float ts = LV(dt, NL, h);
  const float dp = F(dp)[h];
  const float dwatc = F(dwatc)[h];
  const float dsnow = F(dsnow)[h];
  float zwmax = F(dwmax)[h];
  const float zglacfree = 1.0f - F(dglac)[h];
  const float zsvp = K(ra1) * expf(K(ra2) * (ts - K(tmelt)) / (ts - K(ra4)));
  const float zvpd = fmaxf(0.01f, zsvp - dp * LV(dq, NL - 1, h) / 0.622f);
  float dcveg = F(dcveg)[h];
  float dcsoil = F(dcsoil)[h];
  const float zlaim = ZLAIMIN + 0.63661977236758f * ZLAIMAX * atanf(CVEG_L * 2.0f * F(dagg)[h] * dcveg);
  const float zvegm = 1.0f - expf(-CVEG_K * zlaim);
  float zveg = 0.0f;
  if (zwmax > 0.0f) zveg = fminf(fminf(1.0f, zvegm), fmaxf(0.0f, dwatc / (zwmax * VWS_CRIT)));
  const float zft = fminf(1.0f, fmaxf(0.0f, (ts - K(tmelt)) / CT_CRIT));
  const float zgppl = K(rlue) * zbeta * zft * zveg * LV(dfd, NL, h);
  float zgppw = 0.0f;
  if (!(co2 < 1.0e-8f)) zgppw = K(co2conv) * dp * zveg * (-F(devap)[h]) / zvpd * (0.3f * co2);
  zgppw = fminf(1.0f, fmaxf(0.0f, zgppw));
  const float zgpp = fminf(zgppl, zgppw);
  const float znpp = F(dgrow)[h] * 0.5f * zgpp * zglacfree;
  F(dnpp)[h] = znpp;
  const float zlitter = dcveg / K(tau_veg);
  const float zres = powf(Q10, (ts - K(tmelt) - 10.0f) / 10.0f) * dcsoil / K(tau_soil);
  dcveg = dcveg + (znpp - zlitter) * K(deltsec) * K(ncveg);
  dcsoil = dcsoil + (zlitter - zres) * K(deltsec) * K(ncveg);
  const float zforest = fminf(1.0f, fmaxf(0.0f, (atanf(dcveg - CVEG_A) - K(cveg_e)) / (CVEG_C - K(cveg_e))));
  const float dvsoil = fminf(1.0f, fmaxf(0.0f, (atanf(dcveg - CVEG_F) - K(cveg_g)) / (CVEG_C - K(cveg_g))));
  F(dcveg)[h] = dcveg;
  F(dcsoil)[h] = dcsoil;
  F(dvsoil)[h] = dvsoil;
  F(dlai)[h] = -logf(1.0f - zveg) / CVEG_K;
  const float zvz0 = F(dmr)[h] * zforest + VZ0_MIN * (1.0f - zforest);
  float zvalb = VALB_MIN * zveg + VALB_MAX * (1.0f - zveg);
  zwmax = VWMAX_MAX * dvsoil + VWMAX_MIN * (1.0f - dvsoil);
  float zvrhs = F(dsc)[h] * fminf(1.0f, fmaxf(0.0f, dwatc / (zwmax * VWS_CRIT)));
  if (dsnow > 0.0f) {
    float zalbsn = (K(albsmax) - K(albsmin)) * (ts - 263.16f) / (K(tmelt) - 263.16f);
    zalbsn = fmaxf(K(albsmin), fminf(K(albsmax), K(albsmax) - zalbsn));
    zvalb = zvalb + (zalbsn - zvalb) * dsnow / (dsnow + 0.01f);
    zvalb = VSALB_MIN * zforest + zvalb * (1.0f - zforest);
    zvrhs = 1.0f;
  }
  if (K(nveg) == 2.0f) {
    const float z0clim = F(dz0climo)[h];
    F(dz0)[h] = sqrtf(zvz0 * zvz0 + z0clim * z0clim);
    F(dwmax)[h] = zwmax;
    F(drhs)[h] = zvrhs;
    F(dalb)[h] = zvalb;
    F(dforest)[h] = zforest;
  }
}

// glaciermod's glacierstep with glacierstop's promotion at each calendar
// year's end (physics/glacier/wgsl/glacierstep.wgsl:1-25): land that kept
// glacelim of snow all year becomes glacier.
__device__ __forceinline__ void glacierstep(const DevMember& M, int h) {
  if (K(nglacier) == 0.0f) return;
  const bool land = F(dls)[h] > 0.5f;
  if (land && F(dsnowz)[h] < K(glacelim)) F(dpersist)[h] = 0.0f;
  if (land && F(dsnowz)[h] < 0.1f) F(dglac)[h] = 0.0f;
  if (F(surface_step)[6] > 0.5f) {
    if (F(dpersist)[h] > 0.5f) F(dglac)[h] = 1.0f;
    F(dpersist)[h] = F(dls)[h] >= 0.5f ? 1.0f : 0.0f;
  }
}

template <int NL>
__device__ __forceinline__ void land_surface(const DevMember& M, int h) {
  // physics/land/wgsl/surface.wgsl:67-142
  const int NH = M.nhor;
  const float dls = F(dls)[h];
  if (dls > 0) F(driver)[h] = F(driver)[h] + F(drunoff)[h] * K(deltsec);
  const float dts = F(land_dts)[h];
  float albclim = clim_at(M, F(dalbcl), h, 3);
  float albclim1 = clim_at(M, F(dalbcl1), h, 3);
  float albclim2 = clim_at(M, F(dalbcl2), h, 3);
  if (K(nwetsoil) != 0 && dls > 0.5f) {
    if (dts > 273.15f) {
      const float bf = F(dwater)[h] / K(wsmax);
      const float al = 1.0f / 5.2f * powf(powf(fmaxf(bf, 1.0e-3f), (1.0f - 4.0f) / 4.0f) - 1.0f, 1.0f / 4.0f);
      const float dry = expf(-powi(bf * 25.0f, 6));
      const float wet = expf(-powi((1.0f - bf) * 30.0f, 9));
      albclim = K(alblandmax) * dry + al * (1.0f - dry - wet) + K(doceanalb1) * wet;
      albclim1 = albclim;
      albclim2 = K(alblandmax) * dry + al * (1.0f - dry - wet) + K(doceanalb2) * wet;
    } else {
      albclim = K(albland);
      albclim1 = K(dgroundalb1);
      albclim2 = K(dgroundalb2);
    }
  }
  F(dalbclim)[h] = albclim;
  F(dalbclim1)[h] = albclim1;
  F(dalbclim2)[h] = albclim2;
  if (dls > 0) {
    F(land_dtsm)[h] = dts;
    float dqs = K(rdbrv) * K(ra1) * expf(K(ra2) * (dts - K(tmelt)) / (dts - K(ra4))) / F(dp)[h];
    dqs = dqs / (1.0f - (1.0f / K(rdbrv) - 1.0f) * dqs);
    F(land_dqs)[h] = dqs;
    const float dsnow = F(dsnowz)[h];
    F(dsnow)[h] = dsnow;
    if (dsnow > 0) {
      const float forest = F(dforest)[h];
      const float s0 = snow_albedo(M, forest, K(albsmaxf), K(albsmax), K(albsminf), K(albsmin), dts);
      const float s1 = snow_albedo(M, forest, K(albsmaxf1), K(albsmax1), K(albsminf1), K(albsmin1), dts);
      const float s2 = snow_albedo(M, forest, K(albsmaxf2), K(albsmax2), K(albsminf2), K(albsmin2), dts);
      F(dalb)[h] = albclim + (s0 - albclim) * dsnow / (dsnow + 0.01f);
      F(dsalb)[h] = albclim1 + (s1 - albclim1) * dsnow / (dsnow + 0.01f);
      F(dsalb)[NH + h] = albclim2 + (s2 - albclim2) * dsnow / (dsnow + 0.01f);
      F(drhs)[h] = 1.0f;
    } else {
      F(dalb)[h] = albclim;
      F(dsalb)[h] = albclim1;
      F(dsalb)[NH + h] = albclim2;
      const float dwmax = F(dwmax)[h];
      if (dwmax > 0) F(drhs)[h] = fminf(1.0f, F(dwatc)[h] / (K(drhsfull) * dwmax));
    }
    F(dtsoil)[h] = LV(dsoilt, 0, h);
    F(dtd2)[h] = LV(dsoilt, 1, h);
    F(dtd3)[h] = LV(dsoilt, 2, h);
    F(dtd4)[h] = LV(dsoilt, 3, h);
    F(dtd5)[h] = LV(dsoilt, 4, h);
    LV(dt, NL, h) = dts;
    LV(dq, NL, h) = dqs;
  }
  if (K(nveg) > 0.0f) vegstep<NL>(M, h);
  if (F(dglac)[h] > 0.5f && dls > 0) {
    F(dalb)[h] = glacier_albedo(M, K(albgmax), K(albgmin), dts);
    F(dsalb)[h] = glacier_albedo(M, K(albgmax1), K(albgmin1), dts);
    F(dsalb)[NH + h] = glacier_albedo(M, K(albgmax2), K(albgmin2), dts);
    F(drhs)[h] = 1.0f;
  }
  glacierstep(M, h);
}

// --------------------------------------------------- sea ice and ocean

template <int NL>
__device__ __forceinline__ void sea_ice(const DevMember& M, int h) {
  // physics/sea-ice/wgsl/ice.wgsl:36-225
  const int NH = M.nhor;
  const float XDT = K(xdt), XMIND = K(xmind), XMAXD = K(xmaxd), CICEMIN = K(cicemin);
  const float TFREEZE = K(tfreeze), TMELT = K(tmelt), CRHOI = K(crhoi), CRHOS = K(crhos);
  const float CRHOSN = K(crhosn), CPI = K(cpi), CKAPI = K(ckapi), CKAPSN = K(ckapsn);
  const float CLFI = K(clfi), CLFSN = K(clfsn);
  const bool sea = F(dls)[h] < 0.5f;
  float xheat = 0, xprs = 0;
  if (sea) {
    const float naccua = F(sea_naccua)[h] + 1.0f;
    xheat = (F(dshfl)[h] + LV(dswfl, NL, h) + LV(dlwfl, NL, h) + F(dlhfl)[h] + F(sea_cheata)[h]) / naccua;
    xprs = (F(dprs)[h] + F(sea_cprsa)[h]) / naccua;
  }
  F(sea_cheata)[h] = 0;
  F(sea_cprsa)[h] = 0;
  F(sea_naccua)[h] = 0;
  const float zrhoilfdt = CRHOI * CLFI / XDT;
  F(xclssto)[h] = clim_at(M, F(xclsst), h, 0);
  F(xclsst2)[h] = clim_at(M, F(xclsst), h, 3);
  const float cliced = fmaxf(clim_at(M, F(xcliced), h, 3), 0.0f);
  float clicec = fmaxf(clim_at(M, F(xclicec), h, 3), 0.0f);
  if (cliced <= 0) clicec = 0;
  F(xcliced2)[h] = cliced;
  F(xclicec2)[h] = clicec;
  const float xls = F(xls)[h], xsst = F(xsst)[h];
  float xiced = F(xiced)[h], xsnow = F(xsnow)[h], xicecc = F(xicecc)[h];
  float xcflux = F(xcflux)[h], xcfluxf = F(xcfluxf)[h];
  F(ice_snowold)[h] = xsnow;
  const float zicedold = xls < 0.5f ? xiced : 0.0f;
  float ximelt = 0, xsmelt = 0, xsmflx = 0, xscflx = 0, xsndch = 0, xcfluxr = 0;
  if (F(xicec)[h] >= CICEMIN) {
    xsnow = xsnow + xprs * XDT;
    xprs = 0;
  }
  float xts = F(xts)[h], xqmelt = 0;
  const float ztso = fminf(TMELT, xts);
  const float zhsnow = 1.0e3f / CRHOSN * xsnow;
  if (xiced >= XMIND) {
    const float zcpdt = XMIND * CRHOI * CPI / XDT;
    const float zckap = (CKAPSN * zhsnow + CKAPI * xiced) / (zhsnow + xiced);
    float zkapz = zckap / (xiced + zhsnow);
    if (xiced >= XMAXD && XMAXD >= 0) zkapz = 0;
    const float zflx = xheat + zkapz * xsst;
    xts = (zcpdt * ztso + zflx) / (zcpdt + zkapz);
    if (xts > TMELT) {
      xqmelt = xheat + zkapz * (xsst - TMELT) - zcpdt * (TMELT - ztso);
      xts = TMELT;
    }
  } else {
    xts = xsst;
    if (xiced > 0) xqmelt = xheat;
  }
  F(xts)[h] = xts;
  F(xqmelt)[h] = xqmelt;
  if (xiced < XMIND) {
    xcflux = xheat;
    xcfluxf = 0;
  } else {
    const float zhsnowc = 1.0e3f / CRHOSN * xsnow;
    float zckap = (CKAPSN * zhsnowc + CKAPI * xiced) / (zhsnowc + xiced);
    if (XMAXD >= 0 && xiced >= XMAXD) zckap = 0;
    xcflux = zckap * (xts - xsst) / (zhsnowc + xiced);
    xcfluxf = zckap * (TFREEZE - xsst) / (zhsnowc + xiced);
  }
  if (F(xicec)[h] < CICEMIN && xprs > 0) {
    xcflux = xcflux - xprs * 1000.0f * CLFSN;
    xsmelt = xprs + xsmelt;
    xsmflx = xsmflx + xprs * 1000.0f * CLFSN;
    xprs = 0;
  }
  float zqmelt = xqmelt;
  if (xiced > 0) {
    if (xqmelt > xsnow * 1000.0f * CLFSN / XDT) {
      xsmelt = xsmelt + xsnow / XDT;
      xsmflx = xsmflx + xsnow * 1000.0f * CLFSN / XDT;
      zqmelt = xqmelt - xsnow * 1000.0f * CLFSN / XDT;
      xsnow = 0;
    } else if (xqmelt > 0 && xsnow > 0) {
      xsmelt = xsmelt + xqmelt / CLFSN / 1000.0f;
      xsnow = xsnow - xqmelt * XDT / CLFSN / 1000.0f;
      xsmflx = xsmflx + xqmelt;
      zqmelt = 0;
    }
    xcflux = xcflux + zqmelt;
  } else {
    xsmelt = xsmelt + xsnow / XDT;
    xcflux = xcflux - xsnow * 1000.0f * CLFSN / XDT;
    xsmflx = xsmflx + xsnow * 1000.0f * CLFSN / XDT;
    xsnow = 0;
  }
  if (xls < 0.5f && 1.0e3f * xsnow > (CRHOS - CRHOI) * xiced) {
    const float zhice = (1.0e3f * xsnow + CRHOI * xiced) / CRHOS;
    const float zdice = zhice - xiced;
    const float zdsnow = -zdice * CRHOI / 1.0e3f;
    xsnow = xsnow + zdsnow;
    xscflx = zdsnow * 1.0e3f * CLFSN / XDT;
    xsmelt = xsmelt - zdsnow / XDT;
    xsmflx = xsmflx - zdsnow * 1.0e3f * CLFSN / XDT;
  }
  if (xls < 0.5f) xcflux = xcflux + F(xoflux)[h];
  if (xsst <= TFREEZE || xiced > 0) {
    ximelt = xcflux + xscflx - xcfluxf;
    xiced = xiced - ximelt / zrhoilfdt;
    xcflux = xcfluxf;
  }
  if (xiced <= 0) {
    xcflux = xcflux - xiced * zrhoilfdt;
    ximelt = ximelt + xiced * zrhoilfdt;
    xiced = 0;
  }
  if (fabsf(xiced) < 1.0e-9f) xiced = 0;
  if (xiced > zicedold) xicecc = fminf(xicecc + (1.0f - xicecc) * (xiced - zicedold) / 0.5f, 1.0f);
  if (xiced < zicedold) xicecc = fmaxf(xicecc + xicecc * (xiced - zicedold) / (2.0f * zicedold), 0.0f);
  if (xiced <= 0) xicecc = 0;
  if (XMAXD >= 0 && xiced > XMAXD) {
    const float zcflux = (xiced - XMAXD) * zrhoilfdt;
    xiced = XMAXD;
    xcfluxr = xcfluxr + zcflux;
    ximelt = ximelt + zcflux;
    xsndch = xsndch + zcflux * 1000.0f / CRHOI / zrhoilfdt;
  }
  float zflx = F(ice_zflx)[h];
  if (xls < 1.0f) zflx = fmaxf(0.0f, XMAXD - xiced) * zrhoilfdt;
  F(ice_zflx)[h] = zflx;
  const float xgw = F(xgw)[h];
  F(ice_part1)[h] = xls < 1.0f ? xcfluxr * xgw : 0.0f;
  F(ice_part2)[h] = xls < 1.0f ? zflx * xgw : 0.0f;
  F(xiced)[h] = xiced;
  F(xsnow)[h] = xsnow;
  F(xicecc)[h] = xicecc;
  F(xicec)[h] = xicecc;
  F(xcflux)[h] = xcflux;
  F(xcfluxf)[h] = xcfluxf;
  F(xcfluxr)[h] = xcfluxr;
  F(ximelt)[h] = ximelt;
  F(xsmelt)[h] = xsmelt;
  F(xsmflx)[h] = xsmflx;
  F(xsndch)[h] = xsndch;
}

__device__ __forceinline__ void oceanstep(const DevMember& M, int h, float heat) {
  // physics/ocean/wgsl/oceanstep.wgsl:11-46
  F(yclsst2)[h] = clim_at(M, F(yclsst), h, 3);
  const float yls = F(yls)[h], yiced = F(xiced)[h], ymld = F(ymld)[h], yfldo = F(yfldo)[h];
  const float zcpsdt = K(dtmix) / (K(crhos) * K(cps));
  const float ysst = F(ysst)[h];
  float zsst = ysst, yiflux = 0;
  if (yls < 1.0f && yiced <= 0) zsst = zsst + (heat + yfldo) * zcpsdt / ymld;
  if (yls < 1.0f && yiced > 0) yiflux = yiflux - yfldo - heat;
  if (ysst < K(tfreeze) && yls < 1.0f) {
    yiflux = yiflux + (K(tfreeze) - zsst) * ymld / zcpsdt;
    zsst = K(tfreeze);
  }
  if (yiced > 0 && yls < 1.0f) {
    const float zflx = (K(tfreeze) - zsst) * ymld / zcpsdt;
    zsst = zsst + zflx * zcpsdt / ymld;
    yiflux = yiflux + zflx;
  }
  F(ysst)[h] = zsst;
  F(xsst)[h] = zsst;
  F(xmld)[h] = ymld;
  F(xoheat)[h] = -yiflux;
}

__device__ __forceinline__ void ice_finish(const DevMember& M, int h) {
  // physics/sea-ice/wgsl/finish.wgsl:8-46
  const float xls = F(xls)[h];
  float xcflux = F(xcflux)[h], xcfluxr = F(xcfluxr)[h];
  const float s1 = F(ice_sums)[0], s2 = F(ice_sums)[1];
  if (s1 > 0 && s2 > 0 && xls < 1.0f) {
    float zflx = F(ice_zflx)[h];
    const float zfac = s1 / s2;
    if (zfac <= 1.0f) zflx = zflx * zfac;
    xcfluxr = xcfluxr - zflx;
    xcflux = xcflux - zflx;
  }
  F(xcfluxr)[h] = xcfluxr;
  const float xiced = F(xiced)[h];
  const float xicec = F(xicec)[h] >= K(thicec) ? 1.0f : 0.0f;
  float xsnow = F(xsnow)[h], xsmelt = F(xsmelt)[h];
  if (xiced <= 0 && xsnow > 0) {
    const float zcflux = xsnow * 1000.0f * K(clfsn) / K(xdt);
    xsmelt = xsmelt + xsnow / K(xdt);
    F(xsmflx)[h] = F(xsmflx)[h] + zcflux;
    xcflux = xcflux - zcflux;
    xsnow = 0;
  }
  F(xicec)[h] = xicec;
  F(xsnow)[h] = xsnow;
  F(xsmelt)[h] = xsmelt;
  F(xcflux)[h] = xcflux;
  const bool sea = xls < 0.5f;
  oceanstep(M, h, sea ? xcflux : 0.0f);
  F(xoflux)[h] = F(xoheat)[h];
  if (sea) F(xsndch)[h] = F(xsndch)[h] + (xsnow - F(ice_snowold)[h]);
}

template <int NL>
__device__ __forceinline__ void seastep(const DevMember& M, int h) {
  // physics/sea/wgsl/seastep.wgsl:32-75
  const int NH = M.nhor;
  const bool coupled = K(ice_coupled) != 0;
  if (coupled)
    ice_finish(M, h);
  else
    F(sea_naccua)[h] = F(sea_naccua)[h] + 1.0f;
  if (!(F(dls)[h] < 0.5f)) return;
  const float dts = F(xts)[h];
  const float dicec = F(xicec)[h];
  if (coupled) {
    F(sea_dsst)[h] = F(xsst)[h];
    F(sea_dts)[h] = dts;
    F(dmld)[h] = F(xmld)[h];
    F(dicec)[h] = dicec;
    F(diced)[h] = F(xiced)[h];
    F(dsnow)[h] = F(xsnow)[h];
    F(dsmelt)[h] = F(xsmelt)[h];
    F(dsndch)[h] = F(xsndch)[h];
  } else {
    F(sea_cheata)[h] = F(dshfl)[h] + LV(dswfl, NL, h) + LV(dlwfl, NL, h) + F(dlhfl)[h] + F(sea_cheata)[h];
    F(sea_cprsa)[h] = F(dprs)[h] + F(sea_cprsa)[h];
  }
  const float dp = F(dp)[h];
  LV(dt, NL, h) = dts;
  float dqs = K(rdbrv) * K(ra1) * expf(K(ra2) * (dts - K(tmelt)) / (dts - K(ra4))) / dp;
  dqs = dqs / (1.0f - (1.0f / K(rdbrv) - 1.0f) * dqs);
  F(sea_dqs)[h] = dqs;
  LV(dq, NL, h) = dqs;
  const float taux = F(dtaux)[h], tauy = F(dtauy)[h];
  const float zz0 = fmaxf(K(charnock) * sqrtf(taux * taux + tauy * tauy) * K(gascon) * dts / (K(ga) * dp), K(dz0sea));
  F(drhs)[h] = K(drhssea) * (1.0f - dicec) + K(drhsice) * dicec;
  F(dalb)[h] = K(albsea) * (1.0f - dicec) + dicec * fminf(K(albice), 0.5f + 0.025f * (273.0f - dts));
  F(dsalb)[h] = K(doceanalb1) * (1.0f - dicec) + dicec * fminf(K(dicealbmx1), K(dicealbmn1) + 0.025f * (273.0f - dts));
  F(dsalb)[NH + h] =
      K(doceanalb2) * (1.0f - dicec) + dicec * fminf(K(dicealbmx2), K(dicealbmn2) + 0.025f * (273.0f - dts));
  F(dz0)[h] = zz0 * (1.0f - dicec) + K(dz0ice) * dicec;
}

template <int NL>
__device__ __forceinline__ void carbonstep(const DevMember& M, int h) {
  // physics/carbon/wgsl/carbonstep.wgsl:22-50
  const int NH = M.nhor;
  const float precip = (F(dprc)[h] + F(dprl)[h]) * 3.154e9f;
  F(localprecip)[h] = precip;
  F(localavgprecip)[h] = F(localavgprecip)[h] + precip * K(timeweight);
  if (F(carbon_step)[0] == 0.0f) return;
  const float tsurf = LV(dt, NL, h);
  if (K(ncarbon) != 0) {
    float weathering = 0;
    if (F(dls)[h] > 0.5f && tsurf >= 273.15f) {
      const float pco2 = F(co2)[0] * F(dp)[h] * 1.0e-5f;
      weathering = powf(pco2 / K(co2earth), K(cbeta)) * expf(K(kact) * (tsurf - 288.0f)) * powf(precip / K(pearth), 0.65f);
      if (K(nsupply) != 0) {
        const float wmaxp = fmaxf(K(wmax) / K(vearth), 2.0e-15f);
        weathering = wmaxp * (1.0f - expf(-weathering / wmaxp));
      }
    }
    F(localweathering)[h] = weathering * (K(plarad) * K(plarad)) / (K(rad_earthsq) * 0.29f);
  }
  F(localavgtemps)[h] = F(localavgtemps)[h] + tsurf * K(timeweight);
  F(localavgweather)[h] = F(localavgweather)[h] + F(localweathering)[h] * K(timeweight);
}

template <int NL>
__device__ __forceinline__ void physics_out(const DevMember& M, int h) {
  // dynamics/wgsl/physics-out.wgsl:9-21, this column's levels
  const int NH = M.nhor;
  const float PSURF = K(psurf), CV = K(cv), CT = K(ct), WW = K(ww);
  const float scale = sqrtf(T(rcsq)[h]);
  const float dp = F(dp)[h];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    LV(gudt, l, h) = LV(dudt, l, h) / scale / CV / WW + LV(gudt, l, h);
    LV(gvdt, l, h) = LV(dvdt, l, h) / scale / CV / WW + LV(gvdt, l, h);
    LV(gtdt, l, h) = LV(dtdt, l, h) / CT / WW + LV(gtdt, l, h);
    LV(gqdt, l, h) = LV(dqdt, l, h) * dp / WW / PSURF + LV(gqdt, l, h);
  }
}

// The first physics pass runs as four column kernels, so that each gets
// the registers of its own routines (one fused kernel holds the union and
// runs one wave per SIMD).
template <int NL>
__device__ __forceinline__ void surface_fluxes(const DevMember& M, int h) {
  fixer_apply<NL>(M, h);
  surflx<NL>(M, h);
  vdiff<NL>(M, h);
}

template <int NL>
__device__ __forceinline__ void shortwave(const DevMember& M, int h) {
  solang(M, h);
  if (K(ozone) != 0) mko3<NL>(M, h);
  swr<NL>(M, h);
}

template <int NL>
__device__ __forceinline__ void longwave(const DevMember& M, int h) {
  lwr<NL>(M, h);
  radstep<NL>(M, h);
}

template <int NL>
__device__ __forceinline__ void convection(const DevMember& M, int h) {
  kuo<NL>(M, h);
  mkdca<NL>(M, h);
}

// The rain column (physics/precipitation/wgsl/column.wgsl:12-27 up to
// mkrain): mkdca's apply, mklsp (mklsp.wgsl:3-39) and mkrain
// (mkrain.wgsl:20-95) in one pass down the column. Each routine treats a
// level on its own but for mkrain's running fluxes, which run top down,
// so the three run level by level, each level's tendencies kept in
// registers between them: every value gets the same operations in the
// same order as three passes over the column, and each level field is read
// and written once.
template <int NL>
__device__ __forceinline__ void rain(const DevMember& M, int h) {
  const int NH = M.nhor;
  const float DELTSEC2 = K(deltsec2), MOIST = 1.0f / K(rdbrv) - 1.0f;
  const float ALV = K(alv), ALS = K(als), ACPD = K(acpd), ADV = K(adv), TMELT = K(tmelt);
  const bool dca = F(dca_skip)[0] == 0.0f;
  const float dp = F(dp)[h];
  float zprc = 0, zprl = 0, zprsc = 0, zprsl = 0;
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const float ds = T(dsigma)[l];
    const float dq = LV(dq, l, h), t = LV(dt, l, h);
    float dtdt = LV(dtdt, l, h), dqdt = LV(dqdt, l, h);
    // mkdca's apply
    if (dca) {
      dtdt = dtdt + LV(dca_dtdt, l, h);
      dqdt = dqdt + LV(dca_dqdt, l, h);
    }
    // mklsp: large-scale condensation at level l
    float dprll = 0;
    const float dprsll = 0;
    {
      const float zq = dq + dqdt * DELTSEC2;
      const float zt = t + dtdt * DELTSEC2;
      const float zlcpe = ALV / (ACPD * (1.0f + ADV * dq));
      const float p = dp * T(sigma)[l];
      float zqsat = sat_capped(M, zt, p);
      float zcor = 1.0f / (1.0f - MOIST * zqsat);
      zqsat = zqsat * zcor;
      if (zqsat < 0) zqsat = 0;
      if (zq > zqsat) {
        const float c = K(ra2) * (TMELT - K(ra4));
        float zdqdt = (zqsat - zq) / (1.0f + zlcpe * c * zqsat * zcor / ((zt - K(ra4)) * (zt - K(ra4))));
        float ztn = zt - zdqdt * zlcpe;
        float zqn = zq + zdqdt;
        zqsat = sat_capped(M, ztn, p);
        zcor = 1.0f / (1.0f - MOIST * zqsat);
        zqsat = zqsat * zcor;
        zdqdt = (zqsat - zqn) / (1.0f + zlcpe * c * zqsat * zcor / ((ztn - K(ra4)) * (ztn - K(ra4))));
        ztn = ztn - zdqdt * zlcpe;
        zqn = zqn + zdqdt;
        const float lsp_dqdt = (zqn - zq) / DELTSEC2;
        const float lsp_dtdt = (ztn - zt) / DELTSEC2;
        dqdt = dqdt + lsp_dqdt;
        dtdt = dtdt + lsp_dtdt;
        dprll = -fminf(lsp_dqdt, 0.0f) * ds;
      }
    }
    LV(dprsll, l, h) = dprsll;
    LV(dprll, l, h) = dprll;
    // mkrain: the fluxes from above melt or freeze, then evaporate
    float zdtdt = 0, zdqdtl = 0, zdqdts = 0;
    const float dprscl = LV(dprscl, l, h);
    zprc = zprc + LV(dprcl, l, h) - dprscl;
    zprl = zprl + dprll - dprsll;
    zprsc = zprsc + dprscl;
    zprsl = zprsl + dprsll;
    const int surface = l == NL - 1 ? 2 : 1;
    for (int below = 0; below < surface; below++) {
      const float temperature = below ? LV(dt, l + 1, h) : t;
      const float zlcp = (ALS - ALV) / (ACPD * (1.0f + ADV * dq));
      if (zprsc + zprsl > 0 && temperature > TMELT) {
        zdtdt = -(zprsc + zprsl) / ds * zlcp + zdtdt;
        zprc = zprsc + zprc;
        zprl = zprsl + zprl;
        zprsc = 0;
        zprsl = 0;
      } else if (zprc + zprl > 0 && temperature <= TMELT) {
        zdtdt = (zprc + zprl) / ds * zlcp + zdtdt;
        zprsc = zprc + zprsc;
        zprsl = zprl + zprsl;
        zprc = 0;
        zprl = 0;
      }
    }
    const float zq = dq + dqdt * DELTSEC2;
    const float zt = t + dtdt * DELTSEC2;
    float zqsat = sat_capped(M, zt, dp * T(sigma)[l]);
    const float zcor = 1.0f / (1.0f - MOIST * zqsat);
    zqsat = zqsat * zcor;
    if (zqsat < 0) zqsat = 0;
    if (zq < zqsat) {
      const float deficit = zqsat - zq;
      const float water = ALV / (ACPD * (1.0f + ADV * dq));
      const float ice = ALS / (ACPD * (1.0f + ADV * dq));
      if (zprl > 0) {
        const float e = evaporate(M, &zprl, deficit, water, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * water;
        zdqdtl = zdqdtl + e;
      }
      if (zprc > 0) {
        const float e = evaporate(M, &zprc, deficit, water, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * water;
        zdqdtl = zdqdtl + e;
      }
      if (zprsl > 0) {
        const float e = evaporate(M, &zprsl, deficit, ice, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * ice;
        zdqdts = zdqdts + e;
      }
      if (zprsc > 0) {
        const float e = evaporate(M, &zprsc, deficit, ice, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * ice;
        zdqdts = zdqdts + e;
      }
    }
    LV(dqdt, l, h) = dqdt + zdqdtl + zdqdts;
    LV(dtdt, l, h) = dtdt + zdtdt;
  }
  F(dprc)[h] = (zprc + zprsc) * dp / K(ga) / 1000.0f;
  F(dprl)[h] = (zprl + zprsl) * dp / K(ga) / 1000.0f;
  F(dprs)[h] = (zprsc + zprsl) * dp / K(ga) / 1000.0f;
}

template <int NL>
__device__ __forceinline__ void clouds(const DevMember& M, int h) {
  mkclouds<NL>(M, h);
}

template <int NL>
__device__ __forceinline__ void soil_step(const DevMember& M, int h) {
  soil<NL>(M, h);
}

template <int NL>
__device__ __forceinline__ void physics_c(const DevMember& M, int h) {
  land_surface<NL>(M, h);
  if (K(ice_coupled) != 0) sea_ice<NL>(M, h);
}

template <int NL>
__device__ __forceinline__ void physics_d(const DevMember& M, int h) {
  seastep<NL>(M, h);
  carbonstep<NL>(M, h);
  physics_out<NL>(M, h);
}

// Deterministic sum of two grid fields of one member by one workgroup:
// each lane sums a strided share in order, then a fixed tree
// (reduction/wgsl/sum.wgsl:12-34), in FP64 (the planet's mass and energy
// budgets), rounded to FP32 once at the end.
// The planet-wide reductions run one workgroup of kReduceThreads per member.
constexpr int kReduceThreads = 1024;

// grid points per thread of a planet-wide reduction (T63: 18432 points)
constexpr int kReducePts = 18;

__device__ __forceinline__ void block_sum2(const DevMember& M, const float* a, const float* b, float* out) {
  __shared__ double sa[kReduceThreads], sb[kReduceThreads];
  // every load issued before the sums, which run in point order (points
  // past the grid load its last point, unconditionally, and are not summed)
  float va[kReducePts], vb[kReducePts];
#pragma unroll
  for (int j = 0; j < kReducePts; j++) {
    const int h = min((int)threadIdx.x + kReduceThreads * j, M.nhor - 1);
    va[j] = a[h];
    vb[j] = b[h];
  }
  double x = 0, y = 0;
#pragma unroll
  for (int j = 0; j < kReducePts; j++)
    if (threadIdx.x + kReduceThreads * j < M.nhor) {
      x += (double)va[j];
      y += (double)vb[j];
    }
  sa[threadIdx.x] = x;
  sb[threadIdx.x] = y;
  __syncthreads();
  for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = (float)sa[0];
    out[1] = (float)sb[0];
  }
}

#undef F
#undef T
#undef K
#undef LV

}  // namespace

// One lane per column; `waves` is the occupancy the register allocation
// must allow (waves per SIMD).
#define EXO_COLUMN_KERNEL_W(kname, fn, waves)                                                  \
  extern "C" __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(waves)))      \
  void kname(const DevMember* __restrict__ mems, LaunchPlan plan) {                            \
    int tile;                                                                                  \
    const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];                             \
    const int h = tile * 64 + threadIdx.x;                                                     \
    if (h >= M.nhor) return;                                                                   \
    if (M.nlev == 10)                                                                          \
      fn<10>(M, h);                                                                            \
    else                                                                                       \
      fn<5>(M, h);                                                                             \
  }
#define EXO_COLUMN_KERNEL(kname, fn) EXO_COLUMN_KERNEL_W(kname, fn, 1)

EXO_COLUMN_KERNEL(k_fixer_column, fixer_column)
EXO_COLUMN_KERNEL_W(k_surface_fluxes, surface_fluxes, 2)
EXO_COLUMN_KERNEL(k_shortwave, shortwave)
EXO_COLUMN_KERNEL(k_longwave, longwave)
EXO_COLUMN_KERNEL(k_convection, convection)
EXO_COLUMN_KERNEL(k_rain, rain)
EXO_COLUMN_KERNEL(k_clouds, clouds)
EXO_COLUMN_KERNEL(k_soil, soil_step)
EXO_COLUMN_KERNEL(k_physics_c, physics_c)
EXO_COLUMN_KERNEL(k_physics_d, physics_d)
EXO_COLUMN_KERNEL(k_storm, storm_diagnostics)

// The column physics' transcendental routines over arrays, for callers
// that check their accuracy (exo_math_eval): fn 0 log2_pos, 1 pow_pos(a,
// b), 2 log10_pos, 3 log_pos, 4 exp_neg, 5 the library's expf, 6 logf,
// 7 powf(a, b).
extern "C" __global__ __launch_bounds__(256) void k_math_eval(int fn, const float* __restrict__ a,
                                                             const float* __restrict__ b, float* __restrict__ out, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float x = a[i], y = b[i];
  float r;
  switch (fn) {
    case 0: r = log2_pos(x); break;
    case 1: r = pow_pos(x, y); break;
    case 2: r = log10_pos(x); break;
    case 3: r = log_pos(x); break;
    case 4: r = exp_neg(x); break;
    case 5: r = expf(x); break;
    case 6: r = logf(x); break;
    default: r = powf(x, y); break;
  }
  out[i] = r;
}

extern "C" __global__ __launch_bounds__(64) void k_mkradv(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int h = tile * 64 + threadIdx.x;
  if (h >= M.nhor) return;
  mkradv(M, h);
}

// fixer's latitude pass (fixer/wgsl/row.wgsl:4-33): one wave per latitude
// row, four rows a workgroup. Lanes sum strided shares, then a fixed
// butterfly (the same order on every run), in FP64; a row that rescales
// loads every level's values before it stores any.
constexpr int kFixerRowPts = 3;  // longitudes per lane: 192 at T63
constexpr int kMaxLevels = 10;

extern "C" __global__ __launch_bounds__(256) void k_fixer_rows(const DevMember* __restrict__ mems, LaunchPlan plan) {
  int tile;
  const DevMember& M = mems[plan_find(plan, blockIdx.x, &tile)];
  const int row = tile * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= M.nlat) return;
  const int NH = M.nhor, NLON = M.nlon, first = row * NLON;
  float* fixq = dev_arena(M) + M.slot[SLOT_fixq];
  float* fixneg = dev_arena(M) + M.slot[SLOT_fixneg];
  float* fixpos = dev_arena(M) + M.slot[SLOT_fixpos];
  // loads past the row's end read its last point and are not used
  float neg[kFixerRowPts], pos[kFixerRowPts];
#pragma unroll
  for (int k = 0; k < kFixerRowPts; k++) {
    const int j = min(lane + 64 * k, NLON - 1);
    neg[k] = fixneg[first + j];
    pos[k] = fixpos[first + j];
  }
  double zneg = 0, zpos = 0;
#pragma unroll
  for (int k = 0; k < kFixerRowPts; k++)
    if (lane + 64 * k < NLON) {
      zneg += (double)neg[k];
      zpos += (double)pos[k];
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    zneg += __shfl_xor(zneg, o);
    zpos += __shfl_xor(zpos, o);
  }
  if (zneg <= 0 || zpos < zneg) return;
  const float factor = (float)((zpos - zneg) / zpos);
  const int L = M.nlev;
  float q[kFixerRowPts][kMaxLevels];
#pragma unroll
  for (int k = 0; k < kFixerRowPts; k++)
#pragma unroll
    for (int l = 0; l < kMaxLevels; l++)
      q[k][l] = fixq[min(l, L - 1) * NH + first + min(lane + 64 * k, NLON - 1)];
#pragma unroll
  for (int k = 0; k < kFixerRowPts; k++) {
    const int j = lane + 64 * k;
    if (j >= NLON) continue;
#pragma unroll
    for (int l = 0; l < kMaxLevels; l++)
      if (l < L) fixq[l * NH + first + j] = fmaxf(q[k][l] * factor, 0.0f);
    fixneg[first + j] = 0;
    fixpos[first + j] = pos[k] * factor;
  }
}

// fixer's global sum of the rows' results: one workgroup per member.
extern "C" __global__ __launch_bounds__(kReduceThreads) void k_fixer_sum(const DevMember* __restrict__ mems,
                                                                          LaunchPlan plan) {
  const DevMember& M = mems[plan.member[blockIdx.x]];
  block_sum2(M, dev_arena(M) + M.slot[SLOT_fixneg], dev_arena(M) + M.slot[SLOT_fixpos], dev_arena(M) + M.slot[SLOT_fixsum]);
}

// mkdca-limit.wgsl:5-22: dca_skip = 1 when any column gave up.
extern "C" __global__ __launch_bounds__(kReduceThreads) void k_dca_limit(const DevMember* __restrict__ mems,
                                                                          LaunchPlan plan) {
  const DevMember& M = mems[plan.member[blockIdx.x]];
  __shared__ float part[kReduceThreads];
  const float* lim = dev_arena(M) + M.slot[SLOT_dca_limited];
  float flag = 0;
#pragma unroll
  for (int j = 0; j < kReducePts; j++) flag = fmaxf(flag, lim[min((int)threadIdx.x + kReduceThreads * j, M.nhor - 1)]);
  part[threadIdx.x] = flag;
  __syncthreads();
  for (int s = kReduceThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] = fmaxf(part[threadIdx.x], part[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) dev_arena(M)[M.slot[SLOT_dca_skip]] = part[0];
}

// getiflx's global sums (physics/surface/index.ts:72-96), members whose
// sea ice runs this step only.
extern "C" __global__ __launch_bounds__(kReduceThreads) void k_ice_sums(const DevMember* __restrict__ mems,
                                                                         LaunchPlan plan) {
  const DevMember& M = mems[plan.member[blockIdx.x]];
  if (M.k[CON_ice_coupled] == 0.0f) return;
  block_sum2(M, dev_arena(M) + M.slot[SLOT_ice_part1], dev_arena(M) + M.slot[SLOT_ice_part2],
             dev_arena(M) + M.slot[SLOT_ice_sums]);
}

namespace exo {

void enqueue_physics(Engine& e, const std::vector<int>& active) {
  Stepper s(e, active);
  const LaunchPlan one = s.plan([](const DevMember&) { return 1; });
  s.column(k_fixer_column);
  const LaunchPlan rows = s.plan([](const DevMember& d) { return (d.nlat + 3) / 4; });
  launch(e, k_fixer_rows, dim3(rows.first[rows.n]), dim3(256), e.d_desc, rows);
  launch(e, k_fixer_sum, dim3(one.n), dim3(kReduceThreads), e.d_desc, one);
  s.column(k_surface_fluxes);
  s.column(k_shortwave);
  s.column(k_longwave);
  s.column(k_convection);
  launch(e, k_dca_limit, dim3(one.n), dim3(kReduceThreads), e.d_desc, one);
  s.column(k_rain);
  s.column(k_clouds);
  s.column(k_soil);
  s.column(k_mkradv);
  s.column(k_physics_c);
  launch(e, k_ice_sums, dim3(one.n), dim3(kReduceThreads), e.d_desc, one);
  s.column(k_physics_d);
  // hurricanestep (stepper/index.ts:137-158): members with storm
  // diagnostics on (physics/storm/index.ts:27)
  const LaunchPlan storms = s.plan([](const DevMember& d) {
    return d.k[CON_nstormdiag] > 0.0f ? (d.nhor + 63) / 64 : 0;
  });
  if (storms.first[storms.n] > 0) launch(e, k_storm, dim3(storms.first[storms.n]), dim3(64), e.d_desc, storms);
}

}  // namespace exo
