/*
 * FP64 CPU oracle of gridpointd's parameterizations: TEST INFRASTRUCTURE
 * ONLY (see exo_oracle.c). Each routine restates one reference WGSL kernel
 * (file:line given) in real precision, in the reference's dispatch order
 * (stepper/index.ts:148-161): fixer, fluxstep, radstep, rainstep,
 * surfstep, carbonstep, hurricanestep. The reference's f32 "precise_*" helpers
 * (shared/gpu/math/wgsl/precise.wgsl) are libm's log/exp/pow here.
 *
 * Pinned by physics on crafted columns (tests/test_oracle_physics_pins.py):
 * every parameterization (surflx, vdiff, lwr, swr, solang, mko3, radstep,
 * kuo, mkdca, mklsp, mkrain, mkclouds, soil, mkradv, land, sea_ice, the
 * ocean mixed layer, carbon and the fixer), each against laws its results
 * must obey (Kirchhoff, energy and water conservation, backward-Euler
 * diffusion, textbook latent heats, Tetens saturation, random cloud
 * overlap, Green's ozone profile, Charnock's roughness, the log law,
 * Slingo's cloud cover, the Walker-Hays-Kasting weathering law), and
 * hurricanemod's storm diagnostics (tests/test_oracle_storm_pins.py: the
 * pressure-level choice, the kinematic indices, Bolton's saturation, the
 * dry-adiabatic CAPE in closed form, tropical CAPE and potential
 * intensity, the trigonometric root of the ventilation cubic).
 *
 * The details below are pinned on their own (tests/test_oracle_fit_pins.py):
 * lwr's absorptivity fits (zero at zero path, continuous across the branch
 * points the reference's offsets are derived from, monotone, bounded),
 * swr's ozone and vapour absorptions (zero, rising, saturating), the ice
 * surface temperature (the implicit step closes the top layer's energy
 * balance), the glacier and wet-soil albedo curves (endpoints, ramp, the
 * inverted retention curve), the cloud liquid water (Kiehl et al.'s CCM3
 * profile at hypsometric heights), the weathering supply limit
 * (asymptotes, inverse) and the storm's outflow ratio (T_s over the
 * environment's temperature at the level of neutral buoyancy).
 *
 * The two coefficients no property fixes, the slope of the CO2
 * absorptivity's logarithmic branch (0.0546 per decade of path above 1)
 * and the exponent of the ozone absorption's ultraviolet term (0.805),
 * are pinned to the literals read out of lwr.wgsl:122 and swr.wgsl:36
 * (tests/golden/make_fit_constants.py): the oracle's rise per decade and
 * its large-path log-log slope must recover them.
 */
#define EXO_ORACLE_INTERNAL
#include <tgmath.h>
#include <stdlib.h>
#include <string.h>

#include "exo_oracle.h"

#define F(s) A(*v, s)
#define T(s) TB(*v, s)
#define K(s) KK(*v, s)
#define LV(s, l, h) (F(s)[(long)(l) * NH + (h)])

static real minf_(real a, real b) { return a < b ? a : b; }
static real maxf_(real a, real b) { return a > b ? a : b; }
#define MIN(a, b) minf_((a), (b))
#define MAX(a, b) maxf_((a), (b))

static real powi(real x, int n) {
  real r = 1;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

/* ----------------------------------------------------------------- fixer */

/* physics/fixer/wgsl/column.wgsl:4-36 */
static void fixer_column(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  for (int h = 0; h < NH; h++) {
    real neg = 0, pos = 0;
    for (int l = 0; l < L; l++) {
      const real q = LV(dq, l, h) * T(dsigma)[l];
      if (LV(dq, l, h) < 0) neg += q;
      if (LV(dq, l, h) > 0) pos += q;
    }
    real zneg = -neg * F(dp)[h] * T(zgw)[h];
    real zpos = pos * F(dp)[h] * T(zgw)[h];
    real factor = 1;
    const int fixed = zneg > 0 && zpos >= zneg;
    if (fixed) {
      factor = (zpos - zneg) / zpos;
      zneg = 0;
      zpos = zpos * factor;
    }
    for (int l = 0; l < L; l++) {
      const real q = LV(dq, l, h);
      LV(fixq, l, h) = fixed ? MAX(q * factor, 0.0) : q;
    }
    F(fixneg)[h] = zneg;
    F(fixpos)[h] = zpos;
  }
}

/* physics/fixer/wgsl/row.wgsl:4-33 */
static void fixer_row(const M* v) {
  const int NH = v->NHOR, L = v->NLEV, NLON = v->NLON;
  for (int row = 0; row < v->NLAT; row++) {
    const int first = row * NLON;
    real zneg = 0;
    for (int j = 0; j < NLON; j++) zneg += F(fixneg)[first + j];
    if (zneg <= 0) continue;
    real zpos = 0;
    for (int j = 0; j < NLON; j++) zpos += F(fixpos)[first + j];
    if (zpos < zneg) continue;
    const real factor = (zpos - zneg) / zpos;
    for (int j = 0; j < NLON; j++) {
      for (int l = 0; l < L; l++) LV(fixq, l, first + j) = MAX(LV(fixq, l, first + j) * factor, 0.0);
      F(fixneg)[first + j] = 0;
      F(fixpos)[first + j] = F(fixpos)[first + j] * factor;
    }
  }
}

/* reduction/wgsl/sum.wgsl:12-34 (exact-order sum in FP64) */
static void global_sum(const M* v, const real* a, const real* b, real* out) {
  real s1 = 0, s2 = 0;
  for (int h = 0; h < v->NHOR; h++) {
    s1 += a[h];
    s2 += b[h];
  }
  out[0] = s1;
  out[1] = s2;
}

/* physics/fixer/wgsl/apply.wgsl:4-23 */
static void fixer_apply(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real zzneg = F(fixsum)[0], zzpos = F(fixsum)[1];
  for (int h = 0; h < NH; h++)
    for (int l = 0; l < L; l++) {
      real q = LV(fixq, l, h);
      if (zzneg > 0) q = MAX(q * ((zzpos - zzneg) / zzpos), 0.0);
      LV(gqdt, l, h) = LV(gqdt, l, h) + F(dp)[h] * (q - LV(dq, l, h)) / K(delt2) / K(psurf);
      LV(dq, l, h) = q;
    }
}

/* ---------------------------------------------------------- surface flux */

/* physics/surface-flux/wgsl/surflx.wgsl:30-128 */
static void surflx(const M* v) {
  const int NH = v->NHOR, L = v->NLEV, low = L - 1;
  const real VK = 0.4, GA = K(ga), GASCON = K(gascon), DELTSEC2 = K(deltsec2);
  const real moisture = 1.0 / K(rdbrv) - 1.0;
  for (int h = 0; h < NH; h++) {
    const real u = LV(du, low, h), vv = LV(dv, low, h), t = LV(dt, low, h), q = LV(dq, low, h);
    const real ts = LV(dt, L, h), qs = LV(dq, L, h);
    const real dp = F(dp)[h], dls = F(dls)[h], dsigma = T(dsigma)[low];
    const real zfac = T(flux_sigfac)[0], zlnsig = T(flux_lnsig)[0];
    real dtsa = t * zfac * (1.0 + moisture * q);
    const real zabsu2 = MAX(K(zumin), u * u + vv * vv);
    const real znl = -GASCON * 0.5 * (t + dtsa) * zlnsig / GA;
    const real zbz0 = znl / F(dz0)[h];
    const real zri = GA * znl / (zabsu2 * dtsa) * (dtsa - ts * (1.0 + moisture * qs));
    const real zkblnz = VK / log(zbz0 + 1.0);
    const real zkblnz2 = zkblnz * zkblnz;
    real zrifm, zrifh;
    if (zri <= 0) {
      const real zdenom = 1.0 + 3.0 * K(vdiff_c) * K(vdiff_b) * sqrt(-1.0 * zri * (zbz0 + 1.0)) * zkblnz2;
      zrifm = 1.0 - 2.0 * K(vdiff_b) * zri / zdenom;
      if (dls < 1.0) {
        const real zdth = -(dtsa - ts * (1.0 + moisture * qs));
        zrifh = pow(1.0 + pow(0.0016 * pow(zdth, 1.0 / 3.0) / sqrt(zabsu2) / zkblnz2, 1.25), 0.8);
      } else {
        zrifh = 1.0 - 3.0 * K(vdiff_b) * zri / zdenom;
      }
    } else {
      const real zdenom = sqrt(1.0 + K(vdiff_d) * zri);
      zrifh = 1.0 / (1.0 + 3.0 * K(vdiff_b) * zri * zdenom);
      zrifm = 1.0 / (1.0 + 2.0 * K(vdiff_b) * zri / zdenom);
    }
    const real ztrans = sqrt(zabsu2) * zkblnz2;
    const real transh = ztrans * zrifh, transm = ztrans * zrifm;
    F(dtransh)[h] = transh;
    F(dtransm)[h] = transm;
    const real zkonst1 = GA * DELTSEC2 / (GASCON * dsigma);
    const real zkonst2 = dsigma / DELTSEC2 / GA;
    const real zkdiffm = zkonst1 * transm / ts;
    const real zun = u / (1.0 + zkdiffm), zvn = vv / (1.0 + zkdiffm);
    LV(dudt, low, h) = LV(dudt, low, h) + (zun - u) / DELTSEC2;
    LV(dvdt, low, h) = LV(dvdt, low, h) + (zvn - vv) / DELTSEC2;
    real dtdt = LV(dtdt, low, h);
    if (K(ndheat) != 0)
      dtdt = dtdt + -(zun * zun - u * u + zvn * zvn - vv * vv) / DELTSEC2 * 0.5 / K(acpd) / (1.0 + K(adv) * q);
    const real taux = dp * zkonst2 * zkdiffm * zun, tauy = dp * zkonst2 * zkdiffm * zvn;
    F(dtaux)[h] = taux;
    F(dtauy)[h] = tauy;
    const real ustar = sqrt(sqrt(taux * taux + tauy * tauy) * GASCON * ts / dp);
    F(dust3)[h] = ustar * ustar * ustar;
    const real zkdiffh = zkonst1 * transh / ts;
    const real ztn = (t + zkdiffh * ts) / (1.0 + zkdiffh * zfac);
    dtdt = dtdt + (ztn - t) / DELTSEC2;
    LV(dtdt, low, h) = dtdt;
    const real moist = MAX(1.0, 1.0 + K(adv) * q);
    F(dshfl)[h] = zkonst2 * zkdiffh * dp * K(acpd) * moist * (ztn * zfac - ts);
    F(dshdt)[h] = -1.0 * zkonst2 * zkdiffh * dp * K(acpd) * moist;
    dtsa = ztn * zfac;
    F(dtsa)[h] = dtsa;
    const real zkdiffq = F(drhs)[h] * zkonst1 * transh / ts;
    real zqn = MAX(q, (q + zkdiffq * qs) / (1.0 + zkdiffq));
    if (dls > 0) zqn = MIN(zqn, F(dwatc)[h] / K(deltsec) * 1000.0 / dp / zkonst2 + q);
    LV(dqdt, low, h) = LV(dqdt, low, h) + (zqn - q) / DELTSEC2;
    const real devap = -dp * zkonst2 / 1000.0 * (zqn - q);
    F(devap)[h] = devap;
    const real latent = (ts > K(tmelt) || dls < 0.5) ? K(alv) : K(als);
    const real dlhfl = devap * latent * 1000.0;
    F(dlhfl)[h] = dlhfl;
    const real curvature = (ts - K(ra4)) * (ts - K(ra4));
    const real dlhdt = -1.0 * latent * zkdiffq * zkonst2 * dp * K(ra2) * (K(tmelt) - K(ra4)) * qs / curvature;
    F(dlhdt)[h] = dlhfl == 0.0 ? 0.0 : dlhdt;
  }
}

/* physics/diffusion/wgsl/implicit.wgsl:4-26 */
static void diffuse(const M* v, real* x, const real* k) {
  const int L = v->NLEV;
  real ebs[16];
  const real* ds = T(dsigma);
  ebs[0] = k[0] / (ds[0] + k[0]);
  x[0] = ds[0] * x[0] / (ds[0] + k[0]);
  for (int l = 1; l < L - 1; l++) {
    const real den = ds[l] + k[l] + k[l - 1] * (1.0 - ebs[l - 1]);
    ebs[l] = k[l] / den;
    x[l] = (x[l] * ds[l] + k[l - 1] * x[l - 1]) / den;
  }
  x[L - 1] = (x[L - 1] * ds[L - 1] + k[L - 2] * x[L - 2]) / (ds[L - 1] + k[L - 2] * (1.0 - ebs[L - 2]));
  for (int l = L - 2; l >= 0; l--) x[l] = x[l] + ebs[l] * x[l + 1];
}

/* physics/diffusion/wgsl/implicit.wgsl:28-47 */
static void diffuse_potential(const M* v, real* x, const real* k, const real* skap) {
  const int L = v->NLEV;
  real ebs[16];
  const real* ds = T(dsigma);
  ebs[0] = k[0] / (ds[0] + k[0] / skap[0]);
  x[0] = ds[0] * x[0] / (ds[0] + k[0] / skap[0]);
  for (int l = 1; l < L - 1; l++) {
    const real den = ds[l] + (k[l] + k[l - 1] * (1.0 - ebs[l - 1] / skap[l - 1])) / skap[l];
    ebs[l] = k[l] / den;
    x[l] = (x[l] * ds[l] + k[l - 1] / skap[l - 1] * x[l - 1]) / den;
  }
  const int last = L - 1;
  x[last] = (x[last] * ds[last] + k[L - 2] * x[L - 2] / skap[L - 2]) /
            (ds[last] + k[L - 2] / skap[last] * (1.0 - ebs[L - 2] / skap[L - 2]));
  for (int l = L - 2; l >= 0; l--) x[l] = x[l] + ebs[l] * x[l + 1] / skap[l + 1];
}

/* physics/surface-flux/wgsl/vdiff.wgsl:22-120 */
static void vdiff(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real DELTSEC2 = K(deltsec2), GA = K(ga), GASCON = K(gascon);
  const real moisture = 1.0 / K(rdbrv) - 1.0;
  real zt[16], zq[16], zu[16], zv[16], zabsu[16], zskap[16], zskaph[16], kh[16], km[16];
  real zun[16], zvn[16], zke[16], zken[16], zqn[16], ztn[16];
  for (int h = 0; h < NH; h++) {
    for (int l = 0; l < L; l++) {
      zt[l] = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
      zq[l] = LV(dq, l, h) + LV(dqdt, l, h) * DELTSEC2;
      zu[l] = LV(du, l, h) + LV(dudt, l, h) * DELTSEC2;
      zv[l] = LV(dv, l, h) + LV(dvdt, l, h) * DELTSEC2;
      zabsu[l] = sqrt(zu[l] * zu[l] + zv[l] * zv[l]);
      zskap[l] = T(vdiff_skap)[l];
      zskaph[l] = T(vdiff_skaph)[l];
    }
    const real zkonst1 = GA * DELTSEC2 / GASCON;
    const real zkonst2 = zkonst1 * GA / GASCON;
    const real zkonst3 = zkonst2 * GA / GASCON;
    for (int l = 0; l < L - 1; l++) {
      const int lp = l + 1;
      const real sigmah = T(sigmah)[l], dsigma = T(dsigma)[l], dsigmap = T(dsigma)[lp];
      const real zmixh2 = T(vdiff_mixh2)[l], zmixm2 = T(vdiff_mixm2)[l], zrdsig = T(vdiff_rdsig)[l];
      const real zfac = zkonst3 * T(vdiff_sigmah3)[l];
      const real zzkfac = T(vdiff_kfac)[l];
      const real zrthl = (dsigmap + dsigma) / (zt[l] * dsigmap + zt[lp] * dsigma);
      const real zdvds = MAX(K(zumin), fabs(zabsu[l] - zabsu[lp])) * zrdsig;
      const real zdthds =
          (zt[l] * (1.0 + moisture * zq[l]) / zskap[l] - zt[lp] * (1.0 + moisture * zq[lp]) / zskap[lp]) * zrdsig;
      const real zri = zskaph[l] * GASCON * zdthds / (sigmah * zdvds * zdvds);
      real zrifh, zrifm;
      if (zri <= 0) {
        const real zdh = 1.0 + 3.0 * K(vdiff_c) * K(vdiff_b) * sqrt(-1.0 * zri) * zzkfac * zmixh2;
        const real zdm = 1.0 + 3.0 * K(vdiff_c) * K(vdiff_b) * sqrt(-1.0 * zri) * zzkfac * zmixm2;
        zrifh = 1.0 - 3.0 * K(vdiff_b) * zri / zdh;
        zrifm = 1.0 - 2.0 * K(vdiff_b) * zri / zdm;
      } else {
        const real zdenom = sqrt(1.0 + K(vdiff_d) * zri);
        zrifh = 1.0 / (1.0 + 3.0 * K(vdiff_b) * zri * zdenom);
        zrifm = 1.0 / (1.0 + 2.0 * K(vdiff_b) * zri / zdenom);
      }
      kh[l] = zfac * zmixh2 * zrifh * zdvds * zrthl * zrthl * zrthl;
      km[l] = zfac * zmixm2 * zrifm * zdvds * zrthl * zrthl * zrthl;
    }
    for (int l = 0; l < L; l++) {
      zun[l] = zu[l];
      zvn[l] = zv[l];
    }
    diffuse(v, zun, km);
    diffuse(v, zvn, km);
    for (int l = 0; l < L; l++) {
      LV(dudt, l, h) = LV(dudt, l, h) + (zun[l] - zu[l]) / DELTSEC2;
      LV(dvdt, l, h) = LV(dvdt, l, h) + (zvn[l] - zv[l]) / DELTSEC2;
    }
    if (K(ndheat) != 0) {
      for (int l = 0; l < L; l++) zke[l] = zken[l] = 0.5 * zabsu[l] * zabsu[l];
      diffuse(v, zken, km);
      for (int l = 0; l < L; l++) {
        const real zdtdt = -((zun[l] * zun[l] - zu[l] * zu[l] + zvn[l] * zvn[l] - zv[l] * zv[l]) / DELTSEC2 -
                               (zken[l] - zke[l]) / DELTSEC2) *
                             0.5 / K(acpd) / (1.0 + K(adv) * LV(dq, l, h));
        LV(dtdt, l, h) = LV(dtdt, l, h) + zdtdt;
      }
    }
    for (int l = 0; l < L; l++) zqn[l] = zq[l];
    diffuse(v, zqn, kh);
    for (int l = 0; l < L; l++) LV(dqdt, l, h) = LV(dqdt, l, h) + (zqn[l] - zq[l]) / DELTSEC2;
    for (int l = 0; l < L - 1; l++) kh[l] = kh[l] * zskaph[l];
    for (int l = 0; l < L; l++) ztn[l] = zt[l];
    diffuse_potential(v, ztn, kh, zskap);
    for (int l = 0; l < L; l++) LV(dtdt, l, h) = LV(dtdt, l, h) + (ztn[l] - zt[l]) / DELTSEC2;
  }
}

/* ------------------------------------------------------------- radiation */

/* physics/radiation/wgsl/solang.wgsl:10-32 */
static void solang(const M* v) {
  const real PI = 3.14159265359, TWOPI = PI + PI;
  const real ZRLON = TWOPI / v->NLON;
  const real* rs = F(radiation_step);
  for (int h = 0; h < v->NHOR; h++) {
    const int lat = h / v->NLON, lon = h % v->NLON;
    real hangle = rs[2] + lon * ZRLON - PI;
    if (hangle < -PI) hangle = hangle + TWOPI;
    if (hangle > PI) hangle = hangle - TWOPI;
    if (K(rad_synchronous) != 0) hangle = hangle + PI;
    const real muz = rs[0] * T(sid)[lat] + T(cola)[lat] * rs[1] * cos(hangle);
    F(gmu0)[h] = muz > K(zdawn) ? muz : 0.0;
  }
}

/* physics/radiation/wgsl/mko3.wgsl:15-37 */
static void mko3(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real ZFO3 = 100.0 / 2.14, GA = K(ga), GASCON = K(gascon);
  for (int h = 0; h < NH; h++) {
    const real sid = T(sid)[h / v->NLON];
    const real za = K(a0o3) + K(a1o3) * fabs(sid) + K(aco3) * sid * F(radiation_step)[4];
    const real dp = F(dp)[h];
    real o3t = za, zh = 0;
    for (int l = L - 1; l >= 1; l--) {
      zh = zh - LV(dt, l, h) * GASCON / GA * T(rad_lnsig)[l];
      const real o3 = -(za + za * K(zconst_o3)) / (1.0 + exp((zh - K(bo3)) / K(co3))) + o3t;
      LV(dqo3, l, h) = o3 * GA / (ZFO3 * T(dsigma)[l] * dp);
      o3t = o3t - o3;
    }
    LV(dqo3, 0, h) = o3t * GA / (ZFO3 * T(dsigma)[0] * dp);
  }
}

static real ozone_absorption(const M* v, real o3) {
  return (0.02118 * o3 / (1.0 + 0.042 * o3 + 0.000323 * (o3 * o3)) + 1.082 * o3 / pow(1.0 + 138.6 * o3, 0.805) +
          0.0658 * o3 / (1.0 + (103.6 * o3) * (103.6 * o3) * (103.6 * o3))) /
         K(zsolar1);
}

static real vapour_absorption(const M* v, real wv) {
  return 2.9 * wv / (pow(1.0 + 141.5 * wv, 0.635) + 5.925 * wv) / K(zsolar2);
}

/* physics/radiation/wgsl/swr.wgsl:45-276 */
static void swr(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real ZERO = 1.0e-6, ZBETTA = 1.66, ZMBAR = 1.9, ZFO3 = 100.0 / 2.14, ZMU00 = 0.5;
  const real INV_LN10 = 0.43429448190325176, GA = K(ga);
  real rcl1[16], rcl1s[16], rcl2[16], rcl2s[16], tcl2[16], tcl2s[16];
  real o3l[16], xo3l[16], wvl[16], ywvl[16];
  real t1[17], t2[17], r1s[17], r2s[17], tb1[16], tb1u[16], rb1[16], rb1s[16];
  real tb2[16], tb2u[16], rb2[16], rb2s[16], rl1[17], rl2[17], rl1s[17], rl2s[17];
  for (int h = 0; h < NH; h++) {
    const real mu0 = F(gmu0)[h];
    const real gdist2 = F(radiation_step)[3];
    const real ftop1 = K(zsolar1) * K(gsol0) * gdist2 * mu0;
    const real ftop2 = K(zsolar2) * K(gsol0) * gdist2 * mu0;
    if (!(ftop1 + ftop2 > ZERO)) {
      for (int l = 0; l <= L; l++) {
        LV(dfu, l, h) = 0;
        LV(dfd, l, h) = 0;
        LV(dswfl, l, h) = 0;
      }
      continue;
    }
    const real dp = F(dp)[h];
    real cs = 1.0;
    for (int l = 0; l < L; l++) {
      const real cc = LV(dcc, l, h);
      rcl1[l] = rcl1s[l] = rcl2[l] = rcl2s[l] = 0;
      tcl2[l] = tcl2s[l] = 1;
      if (cc > 0) {
        const real lwp = MIN(1000.0, 1000.0 * LV(dql, l, h) * dp / GA * T(dsigma)[l]);
        const real tau = 2.0 * pow(log(lwp + 1.5) * INV_LN10, 3.9);
        const real zlog = log(1000.0 / tau);
        real b2 = K(zb4) / log(3.0 + 0.1 * tau);
        real om0 = MIN(0.9999, 1.0 - K(zb5) * zlog);
        real un = 1.0 - om0;
        real uz = un + 2.0 * b2 * om0;
        real u = sqrt(uz / un);
        real ex = exp(MIN(25.0, tau * sqrt(uz * un) / ZMU00));
        real r = (u + 1.0) * (u + 1.0) * ex - (u - 1.0) * (u - 1.0) / ex;
        rcl1s[l] = 1.0 - 1.0 / (1.0 + K(zb3) * tau);
        tcl2s[l] = 4.0 * u / r;
        rcl2s[l] = (u * u - 1.0) / r * (ex - 1.0 / ex);
        const real b1 = K(tswr1) * sqrt(mu0);
        b2 = K(tswr2) * sqrt(mu0) / log(3.0 + 0.1 * tau);
        om0 = MIN(0.9999, 1.0 - K(tswr3) * mu0 * mu0 * zlog);
        un = 1.0 - om0;
        uz = un + 2.0 * b2 * om0;
        u = sqrt(uz / un);
        ex = exp(MIN(25.0, tau * sqrt(uz * un) / mu0));
        r = (u + 1.0) * (u + 1.0) * ex - (u - 1.0) * (u - 1.0) / ex;
        const real direct1 = 1.0 - 1.0 / (1.0 + b1 * tau / mu0);
        const real directT = 4.0 * u / r;
        const real direct2 = (u * u - 1.0) / r * (ex - 1.0 / ex);
        rcl1[l] = cs * direct1 + (1.0 - cs) * rcl1s[l];
        tcl2[l] = cs * directT + (1.0 - cs) * tcl2s[l];
        rcl2[l] = cs * direct2 + (1.0 - cs) * rcl2s[l];
      }
      cs = cs * (1.0 - cc);
    }
    const real zm = 35.0 / sqrt(1.0 + 1224.0 * mu0 * mu0);
    real o3t = 0, xo3t = 0, wvt = 0, ywvt = 0;
    cs = 1.0;
    for (int l = 0; l < L; l++) {
      const real dsg = T(dsigma)[l];
      const real o3 = ZFO3 * dsg * dp * LV(dqo3, l, h) / GA;
      o3t = o3t + o3;
      xo3t = cs * (xo3t + zm * o3) + (1.0 - cs) * (xo3t + ZMBAR * o3);
      o3l[l] = o3t;
      xo3l[l] = xo3t;
      const real wv = 0.1 * dsg * LV(dq, l, h) * dp / GA * sqrt(273.0 / LV(dt, l, h)) * T(sigma)[l] * dp / 100000.0;
      wvt = wvt + wv;
      ywvt = cs * (ywvt + zm * wv) + (1.0 - cs) * (ywvt + ZBETTA * wv);
      wvl[l] = wvt;
      ywvl[l] = ywvt;
      cs = cs * (1.0 - LV(dcc, l, h));
    }
    real ta1 = 1, ta1s = 1, ra1 = 0, ra1s = 0, ta2 = 1, ta2s = 1, ra2 = 0, ra2s = 0;
    real to3t = 1.0;
    real to3tu = 1.0 - ozone_absorption(v, xo3t + ZMBAR * o3t);
    real twvt = 1.0;
    real twvtu = 1.0 - vapour_absorption(v, ywvt + ZBETTA * wvt);
    const real ccn = LV(dcc, L - 1, h);
    const real scf = K(rcoeff) * dp / 101100.0 * 9.80665 / GA;
    const real diffuseRayleigh = 1.0 - exp(scf * K(ln_diffuse_rayleigh));
    const real rcsuSurface = diffuseRayleigh * (1.0 - ccn);
    const real rcsSurface =
        (1.0 - exp(scf * log(1.0 - 0.219 / (1.0 + 0.816 * mu0)))) * cs + diffuseRayleigh * (1.0 - cs - ccn);
    for (int l = 0; l < L; l++) {
      const real cc = LV(dcc, l, h);
      t1[l] = ta1;
      t2[l] = ta2;
      r1s[l] = ra1s;
      r2s[l] = ra2s;
      const int surface = l == L - 1;
      rb1[l] = (surface ? rcsSurface : 0.0) + rcl1[l] * cc;
      rb1s[l] = (surface ? rcsuSurface : 0.0) + rcl1s[l] * cc;
      real o3 = xo3l[l];
      const real to3 = (1.0 - ozone_absorption(v, o3)) / to3t;
      to3t = to3t * to3;
      o3 = xo3t + ZMBAR * (o3t - o3l[l]);
      const real to3u = to3tu / (1.0 - ozone_absorption(v, o3));
      to3tu = to3tu / to3u;
      tb1[l] = 1.0 - (1.0 - to3) * (1.0 - cc) - rb1[l];
      tb1u[l] = 1.0 - (1.0 - to3u) * (1.0 - cc) - rb1s[l];
      const real m1 = 1.0 / (1.0 - ra1s * rb1s[l]);
      ra1 = ra1 + ta1 * rb1[l] * ta1s * m1;
      ta1 = ta1 * tb1[l] * m1;
      ra1s = rb1s[l] + tb1u[l] * ra1s * tb1[l] * m1;
      ta1s = tb1u[l] * ta1s * m1;
      rb2[l] = rcl2[l] * cc;
      rb2s[l] = rcl2s[l] * cc;
      real wv = ywvl[l];
      const real twv = (1.0 - vapour_absorption(v, wv)) / twvt;
      twvt = twvt * twv;
      wv = ywvt + ZBETTA * (wvt - wvl[l]);
      const real twvu = twvtu / (1.0 - vapour_absorption(v, wv));
      twvtu = twvtu / twvu;
      tb2[l] = 1.0 - (1.0 - twv) * (1.0 - cc) - (1.0 - tcl2[l]) * cc;
      tb2u[l] = 1.0 - (1.0 - twvu) * (1.0 - cc) - (1.0 - tcl2s[l]) * cc;
      const real m2 = 1.0 / (1.0 - ra2s * rb2s[l]);
      ra2 = ra2 + ta2 * rb2[l] * ta2s * m2;
      ta2 = ta2 * tb2[l] * m2;
      ra2s = rb2s[l] + tb2u[l] * ra2s * tb2[l] * m2;
      ta2s = tb2u[l] * ta2s * m2;
    }
    t1[L] = ta1;
    t2[L] = ta2;
    r1s[L] = ra1s;
    r2s[L] = ra2s;
    const real dalb = F(dalb)[h], salb1 = F(dsalb)[h], salb2 = F(dsalb)[NH + h];
    const int nst = K(nstartemp) != 0;
    ra1s = nst ? salb1 : dalb;
    ra2s = nst ? salb2 : dalb;
    const real dls = F(dls)[h], dicec = F(dicec)[h];
    const real water = (1.0 - dls) * (1.0 - dicec) * MIN(0.05 / (mu0 + 0.15), 0.15);
    const real nalb1 = dls * salb1 + (1.0 - dls) * dicec * salb1 + water;
    const real nalb2 = dls * salb2 + (1.0 - dls) * dicec * salb2 + water;
    F(dsalb)[h] = nalb1;
    F(dsalb)[NH + h] = nalb2;
    const real albedo = nst ? K(zsolars1) * nalb1 + K(zsolars2) * nalb2 : dls * dalb + (1.0 - dls) * dicec * dalb + water;
    F(dalb)[h] = albedo;
    ra1 = nst ? nalb1 : albedo;
    ra2 = nst ? nalb2 : albedo;
    for (int l = L - 1; l >= 0; l--) {
      rl1[l + 1] = ra1;
      rl2[l + 1] = ra2;
      rl1s[l + 1] = ra1s;
      rl2s[l + 1] = ra2s;
      const real d1 = 1.0 - ra1s * rb1s[l];
      ra1 = rb1[l] + tb1[l] * ra1 * tb1u[l] / d1;
      ra1s = rb1s[l] + tb1u[l] * ra1s * tb1u[l] / d1;
      const real d2 = 1.0 - ra2s * rb2s[l];
      ra2 = rb2[l] + tb2[l] * ra2 * tb2u[l] / d2;
      ra2s = rb2s[l] + tb2u[l] * ra2s * tb2u[l] / d2;
    }
    rl1[0] = ra1;
    rl2[0] = ra2;
    rl1s[0] = ra1s;
    rl2s[0] = ra2s;
    for (int l = 0; l <= L; l++) {
      const real m1 = 1.0 / (1.0 - r1s[l] * rl1s[l]);
      const real m2 = 1.0 / (1.0 - r2s[l] * rl2s[l]);
      const real fu = -t1[l] * rl1[l] * m1 * ftop1 + -t2[l] * rl2[l] * m2 * ftop2;
      const real fd = t1[l] * m1 * ftop1 + t2[l] * m2 * ftop2;
      LV(dfu, l, h) = fu;
      LV(dfd, l, h) = fd;
      LV(dswfl, l, h) = fu + fd;
    }
  }
}

/* lwr's band absorptivities of an absorber path u (lwr.wgsl:36-179): each
 * is a power law below its branch point and a logarithm above it, with the
 * offsets absorptivityOffsets() computes (physics/radiation/index.ts:
 * 108-125). */
static real lwr_ah2o(const M* v, real u) {
  real a;
  if (u <= 0.01)
    a = 0.846 * pow(u + 3.59e-5, 0.243) - K(zh2o0a);
  else
    a = 0.24 * log10(u + 0.01) + K(zah2oc);
  if (K(th2oc) > 0) a = MIN(a + (1.0 - exp(-K(th2oc) * u)), 1.0);
  return a;
}

static real lwr_aco2(const M* v, real u) {
  if (u <= 1.0) return 0.0676 * pow(u + 0.01022, 0.421) - K(zco20);
  return 0.0546 * log10(u) + K(zaco2c);
}

/* transmission of the water vapour overlapping the CO2 band */
static real lwr_th2o(const M* v, real u) {
  if (u <= 2.0) return 1.0 - (0.832 * pow(u + 0.0286, 0.26) - K(zh2o0));
  return MAX(0.0, K(zth2oc) - 0.1196 * log(u - 0.6931));
}

static real lwr_ao3(const M* v, real u) {
  if (u <= 0.01) return 0.209 * pow(u + 7.0e-5, 0.436) - K(zao30);
  return 0.0212 * log10(u) + K(zao3c);
}

/* physics/radiation/wgsl/lwr.wgsl:36-179 */
static void lwr(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real ZERO = 1.0e-6, ZFH2O = 0.1, ZFO3 = 100.0 / 2.14;
  real st4[17], st4h[17], bu[18], bd[18], q[16], qo3[16], qco2[16], taucc0[16];
  real dftu[17], dftd[17], taucs[16], tau[16], tausf[16];
  const real* sg = T(sigma);
  const real* sh = T(sigmah);
  for (int h = 0; h < NH; h++) {
    const real dp = F(dp)[h], ps2 = dp * dp;
    for (int l = 0; l <= L; l++) {
      const real t = LV(dt, l, h), t2 = t * t;
      st4[l] = K(sbk) * (t2 * t2);
    }
    st4h[0] = st4[0] - (st4[0] - st4[1]) * sg[0] / (sg[0] - sg[1]);
    for (int l = 1; l < L; l++)
      st4h[l] = (st4[l] * (sg[l - 1] - sh[l - 1]) + st4[l - 1] * (sh[l - 1] - sg[l])) / (sg[l - 1] - sg[l]);
    const int n = L - 1;
    if ((st4[n] - st4h[n]) * (st4[n] - st4[L]) > 0)
      st4h[L] = st4[L];
    else
      st4h[L] = st4[n] + (st4[n] - st4[n - 1]) * (1.0 - sg[n]) / (sg[n] - sg[n - 1]);
    bd[0] = 0;
    const real dls = F(dls)[h];
    const real eps = dls + 0.98 * (1.0 - dls);
    bu[L + 1] = eps * st4[L];
    for (int l = 0; l < L; l++) {
      const real sfac = T(lwr_zzf1)[l] * ps2;
      q[l] = ZFH2O * sfac * LV(dq, l, h);
      qo3[l] = ZFO3 * sfac * LV(dqo3, l, h);
      qco2[l] = K(zfco2_zzf2) * sfac * LV(dqco2, l, h);
      taucc0[l] = 1.0 - LV(dcc, l, h) * (1.0 - exp(T(lwr_zzf3)[l] * LV(dql, l, h) * dp));
    }
    for (int l = 0; l <= L; l++) dftu[l] = dftd[l] = 0;
    for (int j = 0; j < L; j++) {
      real taucc = 1.0, sumwv = 0, sumo3 = 0, sumco2 = 0;
      for (int k = j; k < L; k++) {
        sumwv = sumwv + q[k];
        sumo3 = sumo3 + qo3[k];
        sumco2 = sumco2 + qco2[k];
        const real ah2o = lwr_ah2o(v, sumwv), aco2 = lwr_aco2(v, sumco2);
        const real th2o = lwr_th2o(v, sumwv), ao3 = lwr_ao3(v, sumo3);
        const real cs = MIN(1.0 - ZERO, MAX(ZERO, 1.0 - ah2o - ao3 - aco2 * th2o));
        taucs[k] = cs;
        taucc = taucc * taucc0[k];
        tau[k] = cs * taucc;
      }
      if (j == 0) {
        real tau0 = 1.0 - ZERO;
        for (int k = 0; k < L; k++) {
          tau0 = taucs[k] / tau0 * taucc0[k] * K(tpofmt);
          tau0 = MIN(1.0 - ZERO, MAX(ZERO, tau0));
          if ((st4[k] - st4h[k]) * (st4[k] - st4h[k + 1]) > 0) {
            bd[k + 1] = 0.5 * st4[k] + 0.25 * (st4h[k] + st4h[k + 1]);
            bu[k + 1] = bd[k + 1];
          } else {
            bd[k + 1] = (st4h[k + 1] - tau0 * st4h[k]) / (1.0 - tau0) - (st4h[k] - st4h[k + 1]) / log(tau0);
            bu[k + 1] = st4h[k] + st4h[k + 1] - bd[k + 1];
          }
          tau0 = taucs[k];
        }
      }
      dftu[j] = dftu[j] - bu[j + 1];
      dftd[j] = dftd[j] + bd[j];
      const real bdl = bd[j] - bd[j + 1];
      for (int k = j; k < L; k++) {
        dftu[j] = dftu[j] - (bu[k + 2] - bu[k + 1]) * tau[k];
        dftd[k + 1] = dftd[k + 1] + bdl * tau[k];
      }
      tausf[j] = tau[L - 1];
    }
    dftu[L] = dftu[L] - bu[L + 1];
    dftd[L] = dftd[L] + bd[L];
    const real reflected = (1.0 - eps) * dftd[L];
    for (int l = 0; l < L; l++) dftu[l] = dftu[l] - tausf[l] * reflected;
    for (int l = 0; l <= L; l++) {
      LV(dftu, l, h) = dftu[l];
      LV(dftd, l, h) = dftd[l];
      LV(dlwfl, l, h) = dftu[l] + dftd[l];
    }
  }
}

/* physics/radiation/wgsl/radstep.wgsl:11-32 */
static void radstep(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real GA = K(ga), GASCON = K(gascon);
  for (int h = 0; h < NH; h++) {
    for (int l = 0; l <= L; l++) LV(dflux, l, h) = LV(dlwfl, l, h) + LV(dswfl, l, h);
    const real dp = F(dp)[h];
    for (int l = 0; l < L; l++) {
      const real zdh = -LV(dt, l, h) * GASCON / GA * T(rad_lnsig)[l];
      const real den = T(dsigma)[l] * dp * K(acpd) * (1.0 + K(adv) * LV(dq, l, h));
      const real flux = LV(dflux, l, h), below = LV(dflux, l + 1, h);
      LV(dtdt, l, h) = LV(dtdt, l, h) - GA * (below - flux) / den;
      LV(dtdtswr, l, h) = -GA * (LV(dswfl, l + 1, h) - LV(dswfl, l, h)) / den;
      LV(dtdtlwr, l, h) = -GA * (LV(dlwfl, l + 1, h) - LV(dlwfl, l, h)) / den;
      LV(dconv, l, h) = (flux - below) / (zdh + 1.0e-9);
    }
  }
}

/* --------------------------------------------------------- precipitation */

/* physics/precipitation/wgsl/saturation.wgsl:15-24 */
static real sat_capped(const M* v, real t, real p) {
  return MIN(K(rdbrv) * K(ra1) * exp(K(ra2) * (t - K(tmelt)) / (t - K(ra4))) / p, K(rdbrv));
}
static real sat_divided(const M* v, real t, real p) {
  const real q = sat_capped(v, t, p);
  return q / (1.0 - (1.0 / K(rdbrv) - 1.0) * q);
}

typedef struct {
  real t, q, cor;
  int saturated;
} Parcel;

/* kuo.wgsl:22-47 */
static Parcel condense(const M* v, real t0, real q0, real p) {
  const real MOIST = 1.0 / K(rdbrv) - 1.0;
  real s = sat_capped(v, t0, p);
  real cor = 1.0 / (1.0 - MOIST * s);
  s = s * cor;
  Parcel r = {t0, q0, cor, 0};
  if (!(s < q0)) return r;
  real t = t0, q = q0;
  const real lcp = (t0 < K(tmelt) ? K(als) : K(alv)) / (K(acpd) * (1.0 + K(adv) * q0));
  for (int i = 0; i < 2; i++) {
    if (i == 1) {
      s = sat_capped(v, t, p);
      cor = 1.0 / (1.0 - MOIST * s);
      s = s * cor;
    }
    const real dqsdt = K(ra2) * (K(tmelt) - K(ra4)) * s * cor / ((t - K(ra4)) * (t - K(ra4)));
    const real rain = (q - s) / (1.0 + lcp * dqsdt);
    t = t + lcp * rain;
    q = q - rain;
  }
  r.t = t;
  r.q = q;
  r.cor = cor;
  r.saturated = 1;
  return r;
}

/* physics/precipitation/wgsl/kuo.wgsl:53-298 */
static void kuo(const M* v) {
  const int NH = v->NHOR, L = v->NLEV, NLEM = L - 1;
  const real DELTSEC2 = K(deltsec2), ALV = K(alv), ACPD = K(acpd), ADV = K(adv);
  const real MOIST = 1.0 / K(rdbrv) - 1.0;
  const real *ds = T(dsigma), *sg = T(sigma), *skap = T(vdiff_skap), *zf = T(kuo_zfac);
  real zdqt[16], zqe[16], zte[16], zdqtot[16], zalcpe[16], zacpe[16], zdqdt[16], zdtdt[16];
  real zkdiff[16], zqn[16], ztn[16];
  for (int h = 0; h < NH; h++) {
    const real dp = F(dp)[h];
    for (int l = 0; l < L; l++) {
      zdqt[l] = LV(dqt, l, h) * K(psurf) * K(ww) / dp;
      LV(dqt, l, h) = zdqt[l];
      const real dq = LV(dq, l, h), dqdt = LV(dqdt, l, h);
      zdqtot[l] = (zdqt[l] + dqdt) * DELTSEC2;
      zqe[l] = dq + dqdt * DELTSEC2;
      zte[l] = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
      zacpe[l] = ACPD * (1.0 + ADV * dq);
      zalcpe[l] = ALV / zacpe[l];
      zdqdt[l] = 0;
      zdtdt[l] = 0;
    }
    int ilift = -1, itop = -1;
    real zi = 0, zpt = 0, zpq = 0, zbeta = 0, zpbeta = 0, ztp = 0, zqp = 0;
    const int n = L - 1;
    const real dsn = ds[n];
    const real ztsurf = LV(dt, L, h) * skap[n];
    if (zdqtot[n] > 0 && ztsurf > zte[n]) {
      const real p = dp * sg[n];
      const real zqsnl = sat_divided(v, zte[n], p);
      const real zqnew = LV(dq, L, h) * MIN(1.0, zqe[n] / zqsnl);
      const Parcel parcel = condense(v, ztsurf, zqnew, p);
      if (parcel.saturated) {
        ilift = L + 1;
        itop = L;
        ztp = parcel.t;
        zqp = parcel.q;
        const real zqsa = sat_divided(v, zte[n], p);
        zdqdt[n] = MAX(zqsa - zqe[n], 0.0);
        zdtdt[n] = ztp - zte[n];
        zpt = zdtdt[n] * dsn * zacpe[n];
        zpq = zdqdt[n] * dsn * ALV;
        zbeta = MIN(1.0, zqe[n] / zqsa) * dsn;
        zpbeta = dsn;
        zi = zdqtot[n] * dsn * ALV;
      }
    }
    for (int l = NLEM - 1; l >= 0; l--) {
      const int lp = l + 1;
      if (ilift == -1 && zdqtot[lp] > 0) {
        const real p = dp * sg[l];
        const Parcel parcel = condense(v, zte[lp] * zf[l], zqe[lp], p);
        if (parcel.saturated && parcel.t > zte[l]) {
          const real d = ds[l], dspp = ds[lp];
          ilift = l + 2;
          itop = l + 1;
          ztp = parcel.t;
          zqp = parcel.q;
          const real zqsa = sat_divided(v, zte[l], p);
          zdqdt[l] = MAX(zqsa - zqe[l], 0.0);
          zdtdt[l] = ztp - zte[l];
          zpt = zdtdt[l] * d * zacpe[l];
          zpq = zdqdt[l] * d * ALV;
          zbeta = MIN(1.0, zqe[l] / zqsa) * d;
          zpbeta = d;
          const real pp = dp * sg[lp];
          real zqsnl = sat_divided(v, zte[lp], pp);
          if (zqe[lp] > zqsnl) {
            real t = zte[lp], q = zqe[lp];
            const real lcp = (t < K(tmelt) ? K(als) : ALV) / (ACPD * (1.0 + ADV * q));
            real cor = parcel.cor;
            for (int i = 0; i < 2; i++) {
              if (i == 1) {
                zqsnl = sat_capped(v, t, pp);
                cor = 1.0 / (1.0 - MOIST * zqsnl);
                zqsnl = zqsnl * cor;
              }
              const real dqsdt = K(ra2) * (K(tmelt) - K(ra4)) * zqsnl * cor / ((t - K(ra4)) * (t - K(ra4)));
              const real rain = (q - zqsnl) / (1.0 + lcp * dqsdt);
              t = t + lcp * rain;
              q = q - rain;
            }
            zdqdt[lp] = MAX(q - zqe[lp], 0.0);
            zdtdt[lp] = t - zte[lp];
            zpt = zdtdt[lp] * dspp * zacpe[lp] + zpt;
            zpq = zdqdt[lp] * dspp * ALV + zpq;
            zbeta = MIN(1.0, zqe[lp] / zqsnl) * dspp + zbeta;
            zpbeta = dspp + zpbeta;
          }
          zi = zdqtot[l] * d * ALV + zdqtot[lp] * dspp * ALV;
        }
      }
    }
    for (int l = NLEM - 1; l >= 0; l--) {
      if (itop == l + 2) {
        const real p = dp * sg[l];
        const Parcel parcel = condense(v, ztp * zf[l], zqp, p);
        if (parcel.saturated && parcel.t > zte[l]) {
          const real d = ds[l];
          itop = l + 1;
          ztp = parcel.t;
          zqp = parcel.q;
          const real zqsa = sat_divided(v, zte[l], p);
          zdqdt[l] = MAX(zqsa - zqe[l], 0.0);
          zdtdt[l] = ztp - zte[l];
          zpt = zdtdt[l] * d * zacpe[l] + zpt;
          zpq = zdqdt[l] * d * ALV + zpq;
          zbeta = MIN(1.0, zqe[l] / zqsa) * d + zbeta;
          zpbeta = zpbeta + d;
          zi = zdqtot[l] * d * ALV + zi;
        }
      }
    }
    zi = MAX(0.0, zi);
    real zrhm = 0;
    if (zpbeta > 0) zrhm = zbeta / zpbeta;
    real beta = 1.0;
    if (zrhm >= K(rhbeta)) beta = powi((1.0 - zrhm) / (1.0 - K(rhbeta)), (int)K(nbeta));
    real zat = 0, zaq = 0;
    if (zpt > 0) zat = zi * (1.0 - beta) / zpt;
    if (zpq > 0) zaq = zi * beta / zpq;
    int shallow = 0;
    if (itop > 0 && (dp * sg[itop - 1] > K(pdeep) || zi <= 0)) shallow = 1;
    real icctot = 0;
    for (int l = 0; l < L; l++) {
      const int level = l + 1;
      real dprcl = 0, icclev = 0;
      if (ilift >= level && itop <= level && zi > 0 && !shallow) {
        LV(dqdt, l, h) = zaq * zdqdt[l] / DELTSEC2 - zdqt[l];
        LV(dtdt, l, h) = zat * zdtdt[l] / DELTSEC2 + LV(dtdt, l, h);
        dprcl = zat * zdtdt[l] * ds[l] / zalcpe[l] / DELTSEC2;
      }
      if (dprcl > 0) {
        icclev = 1;
        icctot = icctot + 1;
      }
      if (sg[l] < 0.4 && icclev == 1 && itop == level) icclev = 2;
      LV(dprcl, l, h) = dprcl;
      LV(dprscl, l, h) = 0;
      LV(icclev, l, h) = icclev;
    }
    F(icctot)[h] = icctot;
    if (!shallow) continue;
    const real zkonst1 = K(ga) * DELTSEC2 / K(gascon);
    const real zkonst2 = K(rkshallow) * zkonst1 * K(ga) / K(gascon);
    int kktop = 1;
    for (int level = 2; level <= NLEM; level++)
      if (dp * sg[level - 1] < K(pdeepth)) kktop = level;
    const int ktop = itop > kktop ? itop : kktop;
    for (int l = 0; l < L; l++) zkdiff[l] = 0;
    for (int l = 0; l < NLEM; l++) {
      const int lp = l + 1, level = l + 1;
      const real shh = T(sigmah)[l];
      const real zfac = zkonst2 * shh * shh * T(vdiff_rdsig)[l];
      const real d = ds[l], dspp = ds[lp];
      const real zrthl = (dspp + d) / (zte[l] * dspp + zte[lp] * d);
      if (level >= ktop && level < ilift) zkdiff[l] = zfac * zrthl * zrthl;
      if (level + 1 == ktop) {
        real zzqs = sat_capped(v, zte[l], dp * sg[l]);
        zzqs = zzqs * (1.0 / (1.0 - MOIST * zzqs));
        real zzqsp = sat_capped(v, zte[lp], dp * sg[lp]);
        zzqsp = zzqsp * (1.0 / (1.0 - MOIST * zzqsp));
        const real zrh = zqe[l] / zzqs, zrhp = zqe[lp] / zzqsp;
        const real zzf = fabs((zrh - 0.8) / 0.2 * (zrhp - zrh));
        zkdiff[l] = zfac * zrthl * zrthl * zzf;
      }
    }
    for (int l = 0; l < L; l++) zqn[l] = zqe[l];
    diffuse(v, zqn, zkdiff);
    for (int l = 0; l < NLEM; l++) zkdiff[l] = zkdiff[l] * T(vdiff_skaph)[l];
    for (int l = 0; l < L; l++) ztn[l] = zte[l];
    diffuse_potential(v, ztn, zkdiff, skap);
    for (int l = 0; l < L; l++) {
      LV(dtdt, l, h) = LV(dtdt, l, h) + (ztn[l] - zte[l]) / DELTSEC2;
      LV(dqdt, l, h) = LV(dqdt, l, h) + (zqn[l] - zqe[l]) / DELTSEC2;
    }
  }
}

/* physics/precipitation/wgsl/mkdca.wgsl:13-97 */
static void mkdca(const M* v) {
  const int NH = v->NHOR, L = v->NLEV, NLEM = L - 1;
  const real DELTSEC2 = K(deltsec2);
  real zt[16], zth[16], zq[16], zqn[16];
  int mixed[16];
  const real *skap = T(vdiff_skap), *ds = T(dsigma);
  for (int h = 0; h < NH; h++) {
    for (int l = 0; l < L; l++) {
      zt[l] = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
      zth[l] = zt[l] / skap[l];
      zq[l] = LV(dq, l, h) + LV(dqdt, l, h) * DELTSEC2;
      zqn[l] = zq[l];
      mixed[l] = 0;
    }
    real limited = 0;
    int passes = 0;
    for (;;) {
      int unstable = 0, ibase = 0, itop = L;
      real ztht = 0, zqt = 0, zsum1 = 0, zsum2 = 0, zsumq1 = 0, zsumq2 = 0;
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
          unstable = 1;
        }
      }
      if (!unstable) break;
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
      for (int l = 0; l < L; l++) {
        const int level = l + 1;
        if (level > itop && level <= ibase) {
          zth[l] = ztht;
          zqn[l] = zqt;
          mixed[l] = 1;
        }
      }
      passes = passes + 1;
      if (passes > 2 * L) {
        limited = 1;
        break;
      }
    }
    for (int l = 0; l < L; l++) {
      LV(dca_dtdt, l, h) = mixed[l] ? (zth[l] * skap[l] - zt[l]) / DELTSEC2 : 0.0;
      LV(dca_dqdt, l, h) = (zqn[l] - zq[l]) / DELTSEC2;
    }
    F(dca_limited)[h] = limited;
  }
}

/* physics/precipitation/wgsl/mkdca-limit.wgsl:5-22 */
static void mkdca_limit(const M* v) {
  real flag = 0;
  for (int h = 0; h < v->NHOR; h++) flag = MAX(flag, F(dca_limited)[h]);
  F(dca_skip)[0] = flag;
}

/* physics/precipitation/wgsl/mklsp.wgsl:3-39 */
static void mklsp(const M* v, int h) {
  const int NH = v->NHOR, L = v->NLEV;
  const real DELTSEC2 = K(deltsec2), MOIST = 1.0 / K(rdbrv) - 1.0;
  const real dp = F(dp)[h];
  for (int l = 0; l < L; l++) {
    LV(dprsll, l, h) = 0;
    real dprll = 0;
    const real dq = LV(dq, l, h);
    const real zq = dq + LV(dqdt, l, h) * DELTSEC2;
    const real zt = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
    const real zlcpe = K(alv) / (K(acpd) * (1.0 + K(adv) * dq));
    const real p = dp * T(sigma)[l];
    real zqsat = sat_capped(v, zt, p);
    real zcor = 1.0 / (1.0 - MOIST * zqsat);
    zqsat = zqsat * zcor;
    if (zqsat < 0) zqsat = 0;
    if (zq > zqsat) {
      const real c = K(ra2) * (K(tmelt) - K(ra4));
      real zdqdt = (zqsat - zq) / (1.0 + zlcpe * c * zqsat * zcor / ((zt - K(ra4)) * (zt - K(ra4))));
      real ztn = zt - zdqdt * zlcpe;
      real zqn = zq + zdqdt;
      zqsat = sat_capped(v, ztn, p);
      zcor = 1.0 / (1.0 - MOIST * zqsat);
      zqsat = zqsat * zcor;
      zdqdt = (zqsat - zqn) / (1.0 + zlcpe * c * zqsat * zcor / ((ztn - K(ra4)) * (ztn - K(ra4))));
      ztn = ztn - zdqdt * zlcpe;
      zqn = zqn + zdqdt;
      const real dqdt = (zqn - zq) / DELTSEC2;
      const real dtdt = (ztn - zt) / DELTSEC2;
      LV(dqdt, l, h) = LV(dqdt, l, h) + dqdt;
      LV(dtdt, l, h) = LV(dtdt, l, h) + dtdt;
      dprll = -MIN(dqdt, 0.0) * T(dsigma)[l];
    }
    LV(dprll, l, h) = dprll;
  }
}

/* mkrain.wgsl:8-15 */
static real evaporate(const M* v, real* flux, real deficit, real zlcpe, real zqsat, real zcor, real zt,
                        real ds) {
  const real zdqdt = MIN(K(gamma) * deficit * ds / K(deltsec2) /
                               (1.0 + zlcpe * K(ra2) * (K(tmelt) - K(ra4)) * zqsat * zcor / ((zt - K(ra4)) * (zt - K(ra4)))),
                           *flux);
  *flux = *flux - zdqdt;
  return zdqdt / ds;
}

/* physics/precipitation/wgsl/mkrain.wgsl:20-95 */
static void mkrain(const M* v, int h) {
  const int NH = v->NHOR, L = v->NLEV;
  const real DELTSEC2 = K(deltsec2), MOIST = 1.0 / K(rdbrv) - 1.0;
  const real ALV = K(alv), ALS = K(als), ACPD = K(acpd), ADV = K(adv), TMELT = K(tmelt);
  const real dp = F(dp)[h];
  real zprc = 0, zprl = 0, zprsc = 0, zprsl = 0;
  for (int l = 0; l < L; l++) {
    const real ds = T(dsigma)[l];
    const real dq = LV(dq, l, h);
    real zdtdt = 0, zdqdtl = 0, zdqdts = 0;
    zprc = zprc + LV(dprcl, l, h) - LV(dprscl, l, h);
    zprl = zprl + LV(dprll, l, h) - LV(dprsll, l, h);
    zprsc = zprsc + LV(dprscl, l, h);
    zprsl = zprsl + LV(dprsll, l, h);
    const int surface = l == L - 1 ? 2 : 1;
    for (int below = 0; below < surface; below++) {
      const real temperature = LV(dt, l + below, h);
      const real zlcp = (ALS - ALV) / (ACPD * (1.0 + ADV * dq));
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
    const real zq = dq + LV(dqdt, l, h) * DELTSEC2;
    const real zt = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
    real zqsat = sat_capped(v, zt, dp * T(sigma)[l]);
    const real zcor = 1.0 / (1.0 - MOIST * zqsat);
    zqsat = zqsat * zcor;
    if (zqsat < 0) zqsat = 0;
    if (zq < zqsat) {
      const real deficit = zqsat - zq;
      const real water = ALV / (ACPD * (1.0 + ADV * dq));
      const real ice = ALS / (ACPD * (1.0 + ADV * dq));
      if (zprl > 0) {
        const real e = evaporate(v, &zprl, deficit, water, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * water;
        zdqdtl = zdqdtl + e;
      }
      if (zprc > 0) {
        const real e = evaporate(v, &zprc, deficit, water, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * water;
        zdqdtl = zdqdtl + e;
      }
      if (zprsl > 0) {
        const real e = evaporate(v, &zprsl, deficit, ice, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * ice;
        zdqdts = zdqdts + e;
      }
      if (zprsc > 0) {
        const real e = evaporate(v, &zprsc, deficit, ice, zqsat, zcor, zt, ds);
        zdtdt = zdtdt - e * ice;
        zdqdts = zdqdts + e;
      }
    }
    LV(dqdt, l, h) = LV(dqdt, l, h) + zdqdtl + zdqdts;
    LV(dtdt, l, h) = LV(dtdt, l, h) + zdtdt;
  }
  F(dprc)[h] = (zprc + zprsc) * dp / K(ga) / 1000.0;
  F(dprl)[h] = (zprl + zprsl) * dp / K(ga) / 1000.0;
  F(dprs)[h] = (zprsc + zprsl) * dp / K(ga) / 1000.0;
}

/* physics/precipitation/wgsl/mkclouds.wgsl:18-86 */
static void mkclouds(const M* v, int h) {
  const int NH = v->NHOR, L = v->NLEV;
  const real ZCLMAX = 1.0, ZCCA = 0.245, ZCCB = 0.125, ZCCMAX = 0.8, ZCCMIN = 0.05;
  const real DELTSEC2 = K(deltsec2), GA = K(ga), GASCON = K(gascon);
  const real dp = F(dp)[h];
  real cover[16], zzf[16];
  real zcctot = ZCCMIN, zccconv = ZCCMIN;
  const real dprc = F(dprc)[h];
  if (dprc > 0) {
    zcctot = ZCCA + ZCCB * log(dprc * K(zrfac));
    zcctot = MIN(ZCCMAX, MAX(ZCCMIN, zcctot));
    const real levels = F(icctot)[h];
    zccconv = levels == 0.0 ? 1.0 : 1.0 - pow(1.0 - zcctot, 1.0 / levels);
  }
  for (int l = 0; l < L; l++) {
    const real zt = LV(dt, l, h) + LV(dtdt, l, h) * DELTSEC2;
    const real zq = LV(dq, l, h) + LV(dqdt, l, h) * DELTSEC2;
    const real zqsat = sat_divided(v, zt, dp * T(sigma)[l]);
    real zrh = zq / zqsat;
    const real icclev = LV(icclev, l, h);
    real dcc = 0;
    if (icclev > 0) dcc = zccconv;
    if (icclev == 2.0 && zcctot >= 0.3) dcc = MIN(dcc + 2.0 * (zcctot - 0.3), ZCLMAX);
    zrh = zrh * (1.0 - dcc);
    const real zrcrit = T(rcrit)[l];
    real zwfac = 1.0;
    if (T(sigma)[l] > 0.8 && K(clwfac) > 0) {
      const real dw = LV(dw, l, h);
      zwfac = K(clwfac) * MAX(0.0, K(clwcrit1) - dw);
      if (K(clwcrit2) > dw) zwfac = 1.0;
    }
    const real excess = MAX(0.0, (zrh - zrcrit) / (1.0 - zrcrit));
    cover[l] = MIN(dcc + zwfac * (excess * excess), ZCLMAX);
  }
  real dqvi = 0, zzh = 0;
  for (int l = L - 1; l >= 1; l--) {
    dqvi = LV(dq, l, h) * T(dsigma)[l] + dqvi;
    const real zdh = -LV(dt, l, h) * GASCON / GA * T(rad_lnsig)[l];
    zzf[l] = zzh + zdh * 0.5;
    zzh = zzh + zdh;
  }
  dqvi = LV(dq, 0, h) * T(dsigma)[0] + dqvi;
  zzf[0] = zzh - LV(dt, 0, h) * GASCON / GA * T(rad_lnsig)[0];
  dqvi = dqvi * dp / GA;
  F(dqvi)[h] = dqvi;
  zzh = 700.0 * log(1.0 + dqvi);
  real clear = 1.0;
  for (int l = 0; l < L; l++) {
    real dql = 0;
    if (zzh > 0 && cover[l] > 0) {
      dql = 0.00021 * exp(-zzf[l] / zzh) * GASCON * LV(dt, l, h) / (T(sigma)[l] * dp);
      dql = MAX(dql, 1.0e-9);
    }
    LV(dql, l, h) = dql;
    LV(dcc, l, h) = cover[l];
    clear = clear * (1.0 - cover[l]);
  }
  LV(dql, L, h) = 0;
  LV(dcc, L, h) = 1.0 - clear;
}

/* physics/precipitation/wgsl/column.wgsl:12-27 */
static void rain_column(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  for (int h = 0; h < NH; h++) {
    if (F(dca_skip)[0] == 0.0)
      for (int l = 0; l < L; l++) {
        LV(dtdt, l, h) = LV(dtdt, l, h) + LV(dca_dtdt, l, h);
        LV(dqdt, l, h) = LV(dqdt, l, h) + LV(dca_dqdt, l, h);
      }
    mklsp(v, h);
    mkrain(v, h);
    mkclouds(v, h);
  }
}

/* ----------------------------------------------------------------- land */

typedef struct {
  real ctop, soilz1, diff1, cap1;
} SnowLayer;

/* physics/land/wgsl/soil.wgsl:63-77 */
static SnowLayer snow_layer(const M* v, real snowz, real cap0, real diff0, real soilz0) {
  SnowLayer r;
  const real zsntop = MIN(snowz * 1000.0 / K(rhosnow), K(dztop));
  r.ctop = cap0 * K(snowcap) * K(dztop) / (cap0 * zsntop + K(snowcap) * (K(dztop) - zsntop));
  const real zsnowz = MIN(snowz * 1000.0 / K(rhosnow), 1.0) - zsntop;
  r.soilz1 = soilz0 + zsnowz;
  r.diff1 = diff0 * K(snowdiff) * r.soilz1 / (K(snowdiff) * soilz0 + diff0 * zsnowz);
  r.cap1 = cap0 * K(snowcap) * r.soilz1 / (cap0 * zsnowz + K(snowcap) * soilz0);
  return r;
}

/* soil.wgsl:84-91 */
static real surface_temperature(const M* v, SnowLayer layer, real dtsm, real flux, real soil1) {
  const real conduction = 2.0 * layer.diff1;
  return (layer.ctop * K(dztop) * dtsm / K(deltsec) + flux + conduction * soil1 / layer.soilz1) /
         (layer.ctop * K(dztop) / K(deltsec) + conduction / layer.soilz1);
}

/* soil.wgsl:23-55 */
static void mktsoil(const M* v, int h, real ftop, const real* soilz, const real* cap, const real* diff) {
  const int NH = v->NHOR, NS = 5;
  real zdiff[5], zcap[5], zebs[5], ztn[5];
  const int n = NS - 1;
  const real DELTSEC = K(deltsec);
  zdiff[n - 1] = 2.0 * diff[n] * diff[n - 1] / (diff[n] * soilz[n - 1] + diff[n - 1] * soilz[n]);
  zcap[n] = cap[n] * soilz[n] / DELTSEC;
  zebs[n] = 1.0 / (zcap[n] + zdiff[n - 1]);
  ztn[n] = zcap[n] * LV(dsoilt, n, h) * zebs[n];
  for (int l = n - 1; l >= 1; l--) {
    zdiff[l - 1] = 2.0 * diff[l] * diff[l - 1] / (diff[l] * soilz[l - 1] + diff[l - 1] * soilz[l]);
    zcap[l] = cap[l] * soilz[l] / DELTSEC;
    zebs[l] = 1.0 / (zcap[l] + zdiff[l - 1] + zdiff[l] * (1.0 - zdiff[l] * zebs[l + 1]));
    ztn[l] = zebs[l] * (zcap[l] * LV(dsoilt, l, h) + zdiff[l] * ztn[l + 1]);
  }
  zcap[0] = cap[0] * soilz[0] / DELTSEC;
  zebs[0] = 1.0 / (zcap[0] + zdiff[0] * (1.0 - zdiff[0] * zebs[1]));
  ztn[0] = zebs[0] * (zcap[0] * LV(dsoilt, 0, h) + zdiff[0] * ztn[1] + ftop);
  for (int l = 1; l < NS; l++) ztn[l] = ztn[l] + zebs[l] * zdiff[l - 1] * ztn[l - 1];
  const int glacier = F(dglac)[h] > 0.5;
  for (int l = 0; l < NS; l++) LV(dsoilt, l, h) = glacier ? MIN(ztn[l], K(tmelt)) : ztn[l];
}

/* physics/land/wgsl/soil.wgsl:95-185 */
static void soil(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real DELTSEC = K(deltsec), TMELT = K(tmelt), ALS = K(als), ALV = K(alv);
  for (int h = 0; h < NH; h++) {
    const real dls = F(dls)[h];
    const int land = dls > 0;
    const real snowStart = F(dsnowz)[h];
    real dwater = 0, dsmelt = 0;
    if (snowStart == 0.0) F(dsnowt)[h] = TMELT;
    if (land) {
      const real glac = F(dglac)[h];
      real soilz[5], cap[5], diff[5];
      for (int l = 0; l < 5; l++) {
        cap[l] = K(sicecap) * glac + K(soilcap) * (1.0 - glac);
        diff[l] = K(sicediff) * glac + K(soildiff) * (1.0 - glac);
        soilz[l] = T(dsoilz)[l];
      }
      const real zhfla = F(dshfl)[h] + F(dlhfl)[h] + LV(dflux, L, h);
      const real devap = F(devap)[h];
      real zdsnowz = 0;
      if (F(dprs)[h] > 0) zdsnowz = F(dprs)[h];
      if (snowStart > 0) zdsnowz = zdsnowz + devap;
      real zsnowz = MAX(0.0, snowStart + zdsnowz * DELTSEC);
      zdsnowz = (zsnowz - snowStart) / DELTSEC;
      dwater = devap + F(dprl)[h] + F(dprc)[h] - zdsnowz;
      real dsnowz = zsnowz;
      const real dtsm = F(land_dtsm)[h];
      const real soil1 = LV(dsoilt, 0, h);
      SnowLayer layer = snow_layer(v, dsnowz, cap[0], diff[0], soilz[0]);
      real dts = surface_temperature(v, layer, dtsm, zhfla, soil1);
      real zhfls = 2.0 * layer.diff1 / layer.soilz1 * (dts - soil1);
      real zhflm = 0;
      if (dsnowz > 0 && dts > TMELT) {
        dts = TMELT;
        zhflm = MAX(0.0, zhfla - layer.ctop * K(dztop) * (dts - dtsm) / DELTSEC +
                             2.0 * layer.diff1 * (soil1 - dts) / layer.soilz1);
        zsnowz = MAX(0.0, dsnowz - zhflm * DELTSEC / ((ALS - ALV) * 1000.0));
        zdsnowz = (zsnowz - dsnowz) / DELTSEC;
        zhflm = -zdsnowz * 1000.0 * (ALS - ALV);
        dsnowz = zsnowz;
        dsmelt = -zdsnowz;
        zhfls = 2.0 * layer.diff1 / layer.soilz1 * (dts - soil1);
        dwater = dwater - zdsnowz;
        layer = snow_layer(v, dsnowz, cap[0], diff[0], soilz[0]);
      }
      if (dsnowz == 0.0 && zhflm > 0) {
        dts = surface_temperature(v, layer, dtsm, zhfla - zhflm, soil1);
        zhfls = 2.0 * layer.diff1 / layer.soilz1 * (dts - soil1);
      }
      soilz[0] = layer.soilz1;
      cap[0] = layer.cap1;
      diff[0] = layer.diff1;
      mktsoil(v, h, zhfls, soilz, cap, diff);
      if (dsnowz > 0) F(dsnowt)[h] = dts;
      if (K(dsmax) > 0 && dsnowz > K(dsmax)) {
        zdsnowz = (K(dsmax) - dsnowz) / DELTSEC;
        dsnowz = K(dsmax);
        dwater = dwater - zdsnowz;
        dts = dts + zdsnowz * 1000.0 * (ALS - ALV) * DELTSEC / layer.ctop / K(dztop);
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
    real drunoff = 0;
    if (land) {
      const real dwatc = F(dwatc)[h] + DELTSEC * dwater;
      const real dwmax = F(dwmax)[h];
      drunoff = MAX(0.0, dwatc - dwmax) / DELTSEC;
      F(dwatc)[h] = MAX(MIN(dwmax, dwatc), 0.0);
    }
    F(drunoff)[h] = drunoff;
  }
}

/* physics/land/wgsl/mkradv.wgsl:1-59 */
static real river(const M* v, int h) { return F(driver)[h] * F(darea)[h] * F(dls)[h]; }

static void eastward(const M* v, int row, int lon, real* out, real* into) {
  const int h = row + lon;
  const int next = row + (lon + 1) % v->NLON;
  const real u = F(duroff)[h];
  const real water = u > 0 ? river(v, h) : river(v, next);
  *out = u * water / F(darea)[h];
  *into = u * water / F(darea)[next];
}

static void mkradv(const M* v) {
  const int NH = v->NHOR, NLON = v->NLON;
  real* out = (real*)malloc(sizeof(real) * NH);
  for (int h = 0; h < NH; h++) {
    const int lat = h / NLON, lon = h % NLON, row = lat * NLON;
    real runoff = F(drunoff)[h], o, in;
    if (lon == 0) {
      eastward(v, row, 0, &o, &in);
      runoff = runoff - o;
      eastward(v, row, NLON - 1, &o, &in);
      runoff = runoff + in;
    } else {
      eastward(v, row, lon - 1, &o, &in);
      runoff = runoff + in;
      eastward(v, row, lon, &o, &in);
      runoff = runoff - o;
    }
    const real area = F(darea)[h];
    const real vv = F(dvroff)[h];
    const real below = lat + 1 < v->NLAT ? river(v, h + NLON) : 0.0;
    if (vv < 0)
      runoff = runoff + vv * river(v, h) / area;
    else
      runoff = runoff + vv * below / area;
    real fromAbove = 0;
    if (lat > 0) {
      const int above = h - NLON;
      const real va = F(dvroff)[above];
      fromAbove = va < 0 ? -va * river(v, above) : -va * river(v, h);
    }
    out[h] = runoff + fromAbove / area;
  }
  for (int h = 0; h < NH; h++) F(drunoff)[h] = out[h];
  free(out);
}

/* land/wgsl/surface.wgsl:43-49: momint over a 0..13 climatology at nstep+1 */
static real clim_next(const M* v, const real* field, int h) {
  const real* ss = F(surface_step);
  const real w = ss[5];
  return (1.0 - w) * field[(long)ss[3] * v->NHOR + h] + w * field[(long)ss[4] * v->NHOR + h];
}

static real snow_albedo(const M* v, real forest, real maxF, real maxB, real minF, real minB, real ts) {
  const real zalbmax = forest * maxF + (1.0 - forest) * maxB;
  const real zalbmin = forest * minF + (1.0 - forest) * minB;
  const real zdalb = (zalbmax - zalbmin) * (ts - 263.16) / (K(tmelt) - 263.16);
  return MAX(zalbmin, MIN(zalbmax, zalbmax - zdalb));
}

static real glacier_albedo(const M* v, real mx, real mn, real ts) {
  return MAX(mn, MIN(mx, mx - (mx - mn) * (ts - 263.16) / (K(tmelt) - 263.16)));
}

/* simba's vegstep, one land column below 90 % glacier
 * (physics/vegetation/wgsl/vegstep.wgsl:31-99). */
static void vegstep(const M* v, int h) {
  const int NH = v->NHOR, L = v->NLEV;
  if (!(F(dls)[h] > 0 && F(dglac)[h] < 0.90)) return;
  const real VWS_CRIT = 0.25, Q10 = 2, ZLAIMAX = 9, ZLAIMIN = 0.1, VALB_MIN = 0.12, VALB_MAX = 0.3;
  const real VSALB_MIN = 0.2, VWMAX_MAX = 0.5, VWMAX_MIN = 0.05, VZ0_MIN = 0.05, CVEG_A = 2.9;
  const real CVEG_C = 1.570796326794897, CVEG_F = 3, CVEG_K = 0.5, CVEG_L = 0.25, CT_CRIT = 5;
  const real CO2_REF = 360, CO2_COMP = 0, CO2_SENS = 0.3;
  const real co2 = F(co2)[0];
  real zbeta = 0;
  if (!(co2 < 1.0e-8)) zbeta = MAX(0.0, 1.0 + CO2_SENS * log((co2 - CO2_COMP) / (CO2_REF - CO2_COMP)));
  const real ts = LV(dt, L, h), dp = F(dp)[h], dwatc = F(dwatc)[h], dsnow = F(dsnow)[h];
  real zwmax = F(dwmax)[h];
  const real zglacfree = 1.0 - F(dglac)[h];
  const real zsvp = K(ra1) * exp(K(ra2) * (ts - K(tmelt)) / (ts - K(ra4)));
  const real zvpd = MAX(0.01, zsvp - dp * LV(dq, L - 1, h) / 0.622);
  real dcveg = F(dcveg)[h], dcsoil = F(dcsoil)[h];
  const real zlaim = ZLAIMIN + 0.63661977236758 * ZLAIMAX * atan(CVEG_L * 2.0 * F(dagg)[h] * dcveg);
  const real zvegm = 1.0 - exp(-CVEG_K * zlaim);
  real zveg = 0;
  if (zwmax > 0) zveg = MIN(MIN(1.0, zvegm), MAX(0.0, dwatc / (zwmax * VWS_CRIT)));
  const real zft = MIN(1.0, MAX(0.0, (ts - K(tmelt)) / CT_CRIT));
  const real zgppl = K(rlue) * zbeta * zft * zveg * LV(dfd, L, h);
  real zgppw = 0;
  if (!(co2 < 1.0e-8)) zgppw = K(co2conv) * dp * zveg * (-F(devap)[h]) / zvpd * (0.3 * co2);
  zgppw = MIN(1.0, MAX(0.0, zgppw));
  const real zgpp = MIN(zgppl, zgppw);
  const real znpp = F(dgrow)[h] * 0.5 * zgpp * zglacfree;
  F(dnpp)[h] = znpp;
  const real zlitter = dcveg / K(tau_veg);
  const real zres = pow(Q10, (ts - K(tmelt) - 10.0) / 10.0) * dcsoil / K(tau_soil);
  dcveg = dcveg + (znpp - zlitter) * K(deltsec) * K(ncveg);
  dcsoil = dcsoil + (zlitter - zres) * K(deltsec) * K(ncveg);
  const real zforest = MIN(1.0, MAX(0.0, (atan(dcveg - CVEG_A) - K(cveg_e)) / (CVEG_C - K(cveg_e))));
  const real dvsoil = MIN(1.0, MAX(0.0, (atan(dcveg - CVEG_F) - K(cveg_g)) / (CVEG_C - K(cveg_g))));
  F(dcveg)[h] = dcveg;
  F(dcsoil)[h] = dcsoil;
  F(dvsoil)[h] = dvsoil;
  F(dlai)[h] = -log(1.0 - zveg) / CVEG_K;
  const real zvz0 = F(dmr)[h] * zforest + VZ0_MIN * (1.0 - zforest);
  real zvalb = VALB_MIN * zveg + VALB_MAX * (1.0 - zveg);
  zwmax = VWMAX_MAX * dvsoil + VWMAX_MIN * (1.0 - dvsoil);
  real zvrhs = F(dsc)[h] * MIN(1.0, MAX(0.0, dwatc / (zwmax * VWS_CRIT)));
  if (dsnow > 0) {
    real zalbsn = (K(albsmax) - K(albsmin)) * (ts - 263.16) / (K(tmelt) - 263.16);
    zalbsn = MAX(K(albsmin), MIN(K(albsmax), K(albsmax) - zalbsn));
    zvalb = zvalb + (zalbsn - zvalb) * dsnow / (dsnow + 0.01);
    zvalb = VSALB_MIN * zforest + zvalb * (1.0 - zforest);
    zvrhs = 1.0;
  }
  if (K(nveg) == 2) {
    const real z0clim = F(dz0climo)[h];
    F(dz0)[h] = sqrt(zvz0 * zvz0 + z0clim * z0clim);
    F(dwmax)[h] = zwmax;
    F(drhs)[h] = zvrhs;
    F(dalb)[h] = zvalb;
    F(dforest)[h] = zforest;
  }
}

/* glacierstep and glacierstop's yearly promotion
 * (physics/glacier/wgsl/glacierstep.wgsl:1-25). */
static void glacierstep(const M* v, int h) {
  if (K(nglacier) == 0) return;
  const int land = F(dls)[h] > 0.5;
  if (land && F(dsnowz)[h] < K(glacelim)) F(dpersist)[h] = 0;
  if (land && F(dsnowz)[h] < 0.1) F(dglac)[h] = 0;
  if (F(surface_step)[6] > 0.5) {
    if (F(dpersist)[h] > 0.5) F(dglac)[h] = 1;
    F(dpersist)[h] = F(dls)[h] >= 0.5 ? 1 : 0;
  }
}

/* physics/land/wgsl/surface.wgsl:67-142 */
static void land_surface(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  for (int h = 0; h < NH; h++) {
    const real dls = F(dls)[h];
    if (dls > 0) F(driver)[h] = F(driver)[h] + F(drunoff)[h] * K(deltsec);
    const real dts = F(land_dts)[h];
    real albclim = clim_next(v, F(dalbcl), h);
    real albclim1 = clim_next(v, F(dalbcl1), h);
    real albclim2 = clim_next(v, F(dalbcl2), h);
    if (K(nwetsoil) != 0 && dls > 0.5) {
      if (dts > 273.15) {
        const real bf = F(dwater)[h] / K(wsmax);
        const real al = 1.0 / 5.2 * pow(pow(MAX(bf, 1.0e-3), (1.0 - 4.0) / 4.0) - 1.0, 1.0 / 4.0);
        const real dry = exp(-powi(bf * 25.0, 6));
        const real wet = exp(-powi((1.0 - bf) * 30.0, 9));
        albclim = K(alblandmax) * dry + al * (1.0 - dry - wet) + K(doceanalb1) * wet;
        albclim1 = albclim;
        albclim2 = K(alblandmax) * dry + al * (1.0 - dry - wet) + K(doceanalb2) * wet;
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
      real dqs = K(rdbrv) * K(ra1) * exp(K(ra2) * (dts - K(tmelt)) / (dts - K(ra4))) / F(dp)[h];
      dqs = dqs / (1.0 - (1.0 / K(rdbrv) - 1.0) * dqs);
      F(land_dqs)[h] = dqs;
      const real dsnow = F(dsnowz)[h];
      F(dsnow)[h] = dsnow;
      if (dsnow > 0) {
        const real forest = F(dforest)[h];
        const real s0 = snow_albedo(v, forest, K(albsmaxf), K(albsmax), K(albsminf), K(albsmin), dts);
        const real s1 = snow_albedo(v, forest, K(albsmaxf1), K(albsmax1), K(albsminf1), K(albsmin1), dts);
        const real s2 = snow_albedo(v, forest, K(albsmaxf2), K(albsmax2), K(albsminf2), K(albsmin2), dts);
        F(dalb)[h] = albclim + (s0 - albclim) * dsnow / (dsnow + 0.01);
        F(dsalb)[h] = albclim1 + (s1 - albclim1) * dsnow / (dsnow + 0.01);
        F(dsalb)[NH + h] = albclim2 + (s2 - albclim2) * dsnow / (dsnow + 0.01);
        F(drhs)[h] = 1.0;
      } else {
        F(dalb)[h] = albclim;
        F(dsalb)[h] = albclim1;
        F(dsalb)[NH + h] = albclim2;
        const real dwmax = F(dwmax)[h];
        if (dwmax > 0) F(drhs)[h] = MIN(1.0, F(dwatc)[h] / (K(drhsfull) * dwmax));
      }
      F(dtsoil)[h] = LV(dsoilt, 0, h);
      F(dtd2)[h] = LV(dsoilt, 1, h);
      F(dtd3)[h] = LV(dsoilt, 2, h);
      F(dtd4)[h] = LV(dsoilt, 3, h);
      F(dtd5)[h] = LV(dsoilt, 4, h);
      LV(dt, L, h) = dts;
      LV(dq, L, h) = dqs;
    }
    if (K(nveg) > 0) vegstep(v, h);
    if (F(dglac)[h] > 0.5 && dls > 0) {
      F(dalb)[h] = glacier_albedo(v, K(albgmax), K(albgmin), dts);
      F(dsalb)[h] = glacier_albedo(v, K(albgmax1), K(albgmin1), dts);
      F(dsalb)[NH + h] = glacier_albedo(v, K(albgmax2), K(albgmin2), dts);
      F(drhs)[h] = 1.0;
    }
    glacierstep(v, h);
  }
}

/* -------------------------------------------------------- sea ice, ocean */

/* sea-ice/wgsl/ice.wgsl:21-28: offset 0 for nstep, 3 for nstep + 1 */
static real clim_at(const M* v, const real* field, int h, int offset) {
  const real* ss = F(surface_step);
  const real w = ss[offset + 2];
  return (1.0 - w) * field[(long)ss[offset] * v->NHOR + h] + w * field[(long)ss[offset + 1] * v->NHOR + h];
}

/* physics/sea-ice/wgsl/ice.wgsl:36-225 */
static void sea_ice(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real XDT = K(xdt), XMIND = K(xmind), XMAXD = K(xmaxd), CICEMIN = K(cicemin);
  const real TFREEZE = K(tfreeze), TMELT = K(tmelt), CRHOI = K(crhoi), CRHOS = K(crhos);
  const real CRHOSN = K(crhosn), CPI = K(cpi), CKAPI = K(ckapi), CKAPSN = K(ckapsn);
  const real CLFI = K(clfi), CLFSN = K(clfsn);
  for (int h = 0; h < NH; h++) {
    const int sea = F(dls)[h] < 0.5;
    real xheat = 0, xprs = 0;
    if (sea) {
      const real naccua = F(sea_naccua)[h] + 1.0;
      xheat = (F(dshfl)[h] + LV(dswfl, L, h) + LV(dlwfl, L, h) + F(dlhfl)[h] + F(sea_cheata)[h]) / naccua;
      xprs = (F(dprs)[h] + F(sea_cprsa)[h]) / naccua;
    }
    F(sea_cheata)[h] = 0;
    F(sea_cprsa)[h] = 0;
    F(sea_naccua)[h] = 0;
    const real zrhoilfdt = CRHOI * CLFI / XDT;
    F(xclssto)[h] = clim_at(v, F(xclsst), h, 0);
    F(xclsst2)[h] = clim_at(v, F(xclsst), h, 3);
    const real cliced = MAX(clim_at(v, F(xcliced), h, 3), 0.0);
    real clicec = MAX(clim_at(v, F(xclicec), h, 3), 0.0);
    if (cliced <= 0) clicec = 0;
    F(xcliced2)[h] = cliced;
    F(xclicec2)[h] = clicec;
    const real xls = F(xls)[h], xsst = F(xsst)[h];
    real xiced = F(xiced)[h], xsnow = F(xsnow)[h], xicecc = F(xicecc)[h];
    real xcflux = F(xcflux)[h], xcfluxf = F(xcfluxf)[h];
    F(ice_snowold)[h] = xsnow;
    const real zicedold = xls < 0.5 ? xiced : 0.0;
    real ximelt = 0, xsmelt = 0, xsmflx = 0, xscflx = 0, xsndch = 0, xcfluxr = 0;
    if (F(xicec)[h] >= CICEMIN) {
      xsnow = xsnow + xprs * XDT;
      xprs = 0;
    }
    real xts = F(xts)[h], xqmelt = 0;
    const real ztso = MIN(TMELT, xts);
    const real zhsnow = 1.0e3 / CRHOSN * xsnow;
    if (xiced >= XMIND) {
      const real zcpdt = XMIND * CRHOI * CPI / XDT;
      const real zckap = (CKAPSN * zhsnow + CKAPI * xiced) / (zhsnow + xiced);
      real zkapz = zckap / (xiced + zhsnow);
      if (xiced >= XMAXD && XMAXD >= 0) zkapz = 0;
      const real zflx = xheat + zkapz * xsst;
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
      const real zhsnowc = 1.0e3 / CRHOSN * xsnow;
      real zckap = (CKAPSN * zhsnowc + CKAPI * xiced) / (zhsnowc + xiced);
      if (XMAXD >= 0 && xiced >= XMAXD) zckap = 0;
      xcflux = zckap * (xts - xsst) / (zhsnowc + xiced);
      xcfluxf = zckap * (TFREEZE - xsst) / (zhsnowc + xiced);
    }
    if (F(xicec)[h] < CICEMIN && xprs > 0) {
      xcflux = xcflux - xprs * 1000.0 * CLFSN;
      xsmelt = xprs + xsmelt;
      xsmflx = xsmflx + xprs * 1000.0 * CLFSN;
      xprs = 0;
    }
    real zqmelt = xqmelt;
    if (xiced > 0) {
      if (xqmelt > xsnow * 1000.0 * CLFSN / XDT) {
        xsmelt = xsmelt + xsnow / XDT;
        xsmflx = xsmflx + xsnow * 1000.0 * CLFSN / XDT;
        zqmelt = xqmelt - xsnow * 1000.0 * CLFSN / XDT;
        xsnow = 0;
      } else if (xqmelt > 0 && xsnow > 0) {
        xsmelt = xsmelt + xqmelt / CLFSN / 1000.0;
        xsnow = xsnow - xqmelt * XDT / CLFSN / 1000.0;
        xsmflx = xsmflx + xqmelt;
        zqmelt = 0;
      }
      xcflux = xcflux + zqmelt;
    } else {
      xsmelt = xsmelt + xsnow / XDT;
      xcflux = xcflux - xsnow * 1000.0 * CLFSN / XDT;
      xsmflx = xsmflx + xsnow * 1000.0 * CLFSN / XDT;
      xsnow = 0;
    }
    if (xls < 0.5 && 1.0e3 * xsnow > (CRHOS - CRHOI) * xiced) {
      const real zhice = (1.0e3 * xsnow + CRHOI * xiced) / CRHOS;
      const real zdice = zhice - xiced;
      const real zdsnow = -zdice * CRHOI / 1.0e3;
      xsnow = xsnow + zdsnow;
      xscflx = zdsnow * 1.0e3 * CLFSN / XDT;
      xsmelt = xsmelt - zdsnow / XDT;
      xsmflx = xsmflx - zdsnow * 1.0e3 * CLFSN / XDT;
    }
    if (xls < 0.5) xcflux = xcflux + F(xoflux)[h];
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
    if (fabs(xiced) < 1.0e-9) xiced = 0;
    if (xiced > zicedold) xicecc = MIN(xicecc + (1.0 - xicecc) * (xiced - zicedold) / 0.5, 1.0);
    if (xiced < zicedold) xicecc = MAX(xicecc + xicecc * (xiced - zicedold) / (2.0 * zicedold), 0.0);
    if (xiced <= 0) xicecc = 0;
    if (XMAXD >= 0 && xiced > XMAXD) {
      const real zcflux = (xiced - XMAXD) * zrhoilfdt;
      xiced = XMAXD;
      xcfluxr = xcfluxr + zcflux;
      ximelt = ximelt + zcflux;
      xsndch = xsndch + zcflux * 1000.0 / CRHOI / zrhoilfdt;
    }
    real zflx = F(ice_zflx)[h];
    if (xls < 1.0) zflx = MAX(0.0, XMAXD - xiced) * zrhoilfdt;
    F(ice_zflx)[h] = zflx;
    const real xgw = F(xgw)[h];
    F(ice_part1)[h] = xls < 1.0 ? xcfluxr * xgw : 0.0;
    F(ice_part2)[h] = xls < 1.0 ? zflx * xgw : 0.0;
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
}

/* physics/ocean/wgsl/oceanstep.wgsl:11-46 */
static void oceanstep(const M* v, int h, real heat) {
  F(yclsst2)[h] = clim_at(v, F(yclsst), h, 3);
  const real yls = F(yls)[h], yiced = F(xiced)[h], ymld = F(ymld)[h], yfldo = F(yfldo)[h];
  const real zcpsdt = K(dtmix) / (K(crhos) * K(cps));
  const real ysst = F(ysst)[h];
  real zsst = ysst, yiflux = 0;
  if (yls < 1.0 && yiced <= 0) zsst = zsst + (heat + yfldo) * zcpsdt / ymld;
  if (yls < 1.0 && yiced > 0) yiflux = yiflux - yfldo - heat;
  if (ysst < K(tfreeze) && yls < 1.0) {
    yiflux = yiflux + (K(tfreeze) - zsst) * ymld / zcpsdt;
    zsst = K(tfreeze);
  }
  if (yiced > 0 && yls < 1.0) {
    const real zflx = (K(tfreeze) - zsst) * ymld / zcpsdt;
    zsst = zsst + zflx * zcpsdt / ymld;
    yiflux = yiflux + zflx;
  }
  F(ysst)[h] = zsst;
  F(xsst)[h] = zsst;
  F(xmld)[h] = ymld;
  F(xoheat)[h] = -yiflux;
}

/* physics/sea-ice/wgsl/finish.wgsl:8-46 */
static void ice_finish(const M* v, int h) {
  const real xls = F(xls)[h];
  real xcflux = F(xcflux)[h], xcfluxr = F(xcfluxr)[h];
  const real s1 = F(ice_sums)[0], s2 = F(ice_sums)[1];
  if (s1 > 0 && s2 > 0 && xls < 1.0) {
    real zflx = F(ice_zflx)[h];
    const real zfac = s1 / s2;
    if (zfac <= 1.0) zflx = zflx * zfac;
    xcfluxr = xcfluxr - zflx;
    xcflux = xcflux - zflx;
  }
  F(xcfluxr)[h] = xcfluxr;
  const real xiced = F(xiced)[h];
  const real xicec = F(xicec)[h] >= K(thicec) ? 1.0 : 0.0;
  real xsnow = F(xsnow)[h], xsmelt = F(xsmelt)[h];
  if (xiced <= 0 && xsnow > 0) {
    const real zcflux = xsnow * 1000.0 * K(clfsn) / K(xdt);
    xsmelt = xsmelt + xsnow / K(xdt);
    F(xsmflx)[h] = F(xsmflx)[h] + zcflux;
    xcflux = xcflux - zcflux;
    xsnow = 0;
  }
  F(xicec)[h] = xicec;
  F(xsnow)[h] = xsnow;
  F(xsmelt)[h] = xsmelt;
  F(xcflux)[h] = xcflux;
  const int sea = xls < 0.5;
  oceanstep(v, h, sea ? xcflux : 0.0);
  F(xoflux)[h] = F(xoheat)[h];
  if (sea) F(xsndch)[h] = F(xsndch)[h] + (xsnow - F(ice_snowold)[h]);
}

/* physics/sea/wgsl/seastep.wgsl:32-75 */
static void seastep(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const int coupled = K(ice_coupled) != 0;
  for (int h = 0; h < NH; h++) {
    if (coupled)
      ice_finish(v, h);
    else
      F(sea_naccua)[h] = F(sea_naccua)[h] + 1.0;
    if (!(F(dls)[h] < 0.5)) continue;
    const real dts = F(xts)[h];
    const real dicec = F(xicec)[h];
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
      F(sea_cheata)[h] = F(dshfl)[h] + LV(dswfl, L, h) + LV(dlwfl, L, h) + F(dlhfl)[h] + F(sea_cheata)[h];
      F(sea_cprsa)[h] = F(dprs)[h] + F(sea_cprsa)[h];
    }
    const real dp = F(dp)[h];
    LV(dt, L, h) = dts;
    real dqs = K(rdbrv) * K(ra1) * exp(K(ra2) * (dts - K(tmelt)) / (dts - K(ra4))) / dp;
    dqs = dqs / (1.0 - (1.0 / K(rdbrv) - 1.0) * dqs);
    F(sea_dqs)[h] = dqs;
    LV(dq, L, h) = dqs;
    const real taux = F(dtaux)[h], tauy = F(dtauy)[h];
    const real zz0 = MAX(K(charnock) * sqrt(taux * taux + tauy * tauy) * K(gascon) * dts / (K(ga) * dp), K(dz0sea));
    F(drhs)[h] = K(drhssea) * (1.0 - dicec) + K(drhsice) * dicec;
    F(dalb)[h] = K(albsea) * (1.0 - dicec) + dicec * MIN(K(albice), 0.5 + 0.025 * (273.0 - dts));
    F(dsalb)[h] = K(doceanalb1) * (1.0 - dicec) + dicec * MIN(K(dicealbmx1), K(dicealbmn1) + 0.025 * (273.0 - dts));
    F(dsalb)[NH + h] = K(doceanalb2) * (1.0 - dicec) + dicec * MIN(K(dicealbmx2), K(dicealbmn2) + 0.025 * (273.0 - dts));
    F(dz0)[h] = zz0 * (1.0 - dicec) + K(dz0ice) * dicec;
  }
}


/* ------------------------------------------------------------- storms */

/* physics/storm/wgsl/storm.wgsl:1-208 (hurricanemod.f90 at ExoPlaSim
 * 5f8dff4): CAPE of the lowest-level parcel, the maximum potential
 * intensity by fixed-point iteration on the central pressure, the entropy
 * deficit, ventilation and genesis potential indices. */
#define ST_CPD ((real)1005.7)
#define ST_CL ((real)2500.0)
#define ST_RVP ((real)461.5)
#define ST_ALV0 ((real)2.501e6)
#define ST_CPVMCL ((real)-630.0)

typedef struct {
  real energy, temperature, pressure;
} st_cape;

/* storm.wgsl:26-28: hPa over water of a Celsius temperature */
static real st_es(real tc) { return (real)6.112 * exp((real)17.67 * tc / ((real)243.5 + tc)); }
/* storm.wgsl:31-38 */
static real st_vapor(real r, real p, real eps) { return r * p / (eps + r); }
static real st_mixing(real e, real p, real eps) { return eps * e / (p - e); }
/* storm.wgsl:41-43 */
static real st_density(real t, real rt, real r, real eps) { return t * (1 + r / eps) / (1 + rt); }
/* storm.wgsl:46-54 */
static real st_entropy(real t, real r, real p, real RD, real eps) {
  const real e = st_vapor(r, p, eps);
  const real rh = MIN(e / st_es(t - (real)273.15), 1);
  const real lv = ST_ALV0 + ST_CPVMCL * (t - (real)273.15);
  const real dry = (ST_CPD + r * ST_CL) * log(t) - RD * log(p - e);
  if (r == 0) return dry;
  return dry + lv * r / t - r * ST_RVP * log(rh);
}

/* storm.wgsl:56-131 */
static st_cape st_cape_of(int L, const real* t, const real* r, const real* p, real tp, real rp, real pp, real RD) {
  const real eps = RD / ST_RVP;
  st_cape out = {0, -1, -1};
  if (rp < (real)1e-6 || tp < 200 || pp <= 0) return out;
  const real rh = MIN(st_vapor(rp, pp, eps) / st_es(tp - (real)273.15), 1);
  const real parcel_entropy = st_entropy(tp, rp, pp, RD, eps);
  const real plcl = pp * pow(rh, tp / ((real)1669.0 - (real)122.0 * rh - tp));
  real buoyancy[16];
  for (int l = L - 1; l >= 0; l--) { /* the topmost level is the last one tried */
    real tg, rg = rp;
    if (p[l] >= plcl) {
      tg = tp * pow(p[l] / pp, RD / ST_CPD);
    } else {
      real next = t[l];
      unsigned count = 0;
      tg = 0;
      for (;;) {
        if (fabs(next - tg) <= (real)0.001) break;
        tg = next;
        const real tc = tg - (real)273.15;
        const real enew = st_es(tc);
        rg = st_mixing(enew, p[l], eps);
        count++;
        const real lv = ST_ALV0 + ST_CPVMCL * tc;
        const real sl = (ST_CPD + rp * ST_CL + lv * lv * rg / (ST_RVP * tg * tg)) / tg;
        const real em = st_vapor(rg, p[l], eps);
        if (p[l] - em <= 0 || tg <= 0) {
          out.energy = 0, out.temperature = t[l], out.pressure = -1;
          return out;
        }
        const real sg = (ST_CPD + rp * ST_CL) * log(tg) - RD * log(p[l] - em) + lv * rg / tg;
        next = tg + (count < 3 ? (real)0.3 : (real)1.0) * (parcel_entropy - sg) / sl;
        if (count > 500 || enew > p[l] - 1 || !(next > 0)) {
          out.energy = 0, out.temperature = t[L - 1], out.pressure = p[L - 1];
          return out;
        }
      }
    }
    buoyancy[l] = st_density(tg, rp, rg, eps) - st_density(t[l], r[l], r[l], eps);
  }
  int inb = L - 1;
  for (int l = 0; l < L; l++)
    if (buoyancy[l] > 0 && l < inb) inb = l;
  if (inb == L - 1) {
    out.energy = 0, out.temperature = t[L - 1], out.pressure = 0;
    return out;
  }
  real area = 0;
  for (int j = L - 2; j >= inb; j--)
    area += RD * (buoyancy[j] + buoyancy[j + 1]) * (p[j + 1] - p[j]) / (p[j] + p[j + 1]);
  area += RD * (pp - p[L - 1]) / (pp + p[L - 1]) * buoyancy[L - 1];
  real tob = t[inb], lnb = p[inb];
  if (inb > 0) {
    lnb = (p[inb - 1] * buoyancy[inb] - p[inb] * buoyancy[inb - 1]) / (buoyancy[inb] - buoyancy[inb - 1]);
    area += RD * buoyancy[inb] * (p[inb] - lnb) / (p[inb] + lnb);
    tob = (t[inb] * (lnb - p[inb - 1]) + t[inb - 1] * (p[inb] - lnb)) / (p[inb] - p[inb - 1]);
  }
  out.energy = MAX(area, 0), out.temperature = tob, out.pressure = lnb;
  return out;
}

/* The maximum potential intensity's iteration (storm.wgsl:140-208): the
 * central pressure pm is iterated until it moves by at most 0.5 hPa; -1
 * when it fails. parts (may be NULL) receives the last iteration's outflow
 * ratio sst / T_out, T_out (the environment's temperature at the saturated
 * parcel's level of neutral buoyancy), that level's pressure, and the CAPE
 * of the saturated and of the ambient parcel at pm. */
static real st_potential(int L, const real* t, const real* r, const real* p, real sst, real msl, real RD,
                         real CKCD, real environmental, real* parts) {
  const real eps = RD / ST_RVP;
  const int bottom = L - 1;
  real pm = 970, old = pm, next = 0, capem = 0, capems = 0, ratio = 1, tout = -1, pout = -1;
  unsigned count = 0;
  for (;;) {
    if (fabs(next - old) <= (real)0.5) break;
    const real pp = MIN(pm, 1000);
    const real rp = eps * r[bottom] * msl / (pp * (eps + r[bottom]) - r[bottom] * msl);
    capem = st_cape_of(L, t, r, p, t[bottom], rp, pp, RD).energy;
    const real rs = st_mixing(st_es(sst - (real)273.15), pp, eps);
    const st_cape saturated = st_cape_of(L, t, r, p, sst, rs, pp, RD);
    capems = saturated.energy;
    tout = saturated.temperature;
    pout = saturated.pressure;
    ratio = sst / saturated.temperature;
    const real tvav = (real)0.5 * (st_density(t[bottom], r[bottom], r[bottom], eps) + st_density(sst, rs, rs, eps));
    const real cat = MAX((capem - environmental) + (real)0.5 * CKCD * ratio * (capems - capem), 0);
    next = msl * exp(-cat / (RD * tvav));
    old = pm;
    pm = next;
    count++;
    if (count > 200 || pm < 400 || !(pm > 0)) return -1;
  }
  if (parts) {
    parts[0] = ratio;
    parts[1] = tout;
    parts[2] = pout;
    parts[3] = capems;
    parts[4] = capem;
  }
  return sqrt(MAX(0, CKCD * ratio * MAX(0, capems - capem)));
}

/* storm.wgsl:133-138 */
static real st_reduced(real vi, real vithresh) {
  if (vi == 0 || fabs(vi) >= vithresh) return 0;
  return 2 / sqrt((real)3) * cos(acos(-vi / vithresh) / 3);
}

/* storm.wgsl:140-208 */
static void storm(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  const real RD = K(gascon), OMEGA = K(ww), CKCD = K(ckcd), eps = RD / ST_RVP;
  const real* sigma = T(sigma);
  for (int h = 0; h < NH; h++) {
    real t[16], r[16], p[16];
    const real msl = F(dp)[h] * (real)0.01;
    const real sst = LV(dt, L, h);
    int i600 = 0, i850 = 0, i200 = 0, cold = 0;
    for (int l = 0; l < L; l++) {
      t[l] = LV(dt, l, h);
      r[l] = LV(dq, l, h);
      p[l] = msl * sigma[l];
      cold = cold || t[l] <= 100;
      if (fabs(sigma[l] - (real)0.6) < fabs(sigma[i600] - (real)0.6)) i600 = l;
      if (fabs(sigma[l] - (real)0.85) < fabs(sigma[i850] - (real)0.85)) i850 = l;
      if (fabs(sigma[l] - (real)0.2) < fabs(sigma[i200] - (real)0.2)) i200 = l;
    }
    const int bottom = L - 1;
    const real u850 = LV(du, i850, h), v850 = LV(dv, i850, h), u200 = LV(du, i200, h), v200 = LV(dv, i200, h);
    const real lowwind = sqrt(u850 * u850 + v850 * v850), highwind = sqrt(u200 * u200 + v200 * v200);
    const real shear = fabs(highwind - lowwind);
    const real vort = LV(gz, i850, h) * OMEGA;
    real deficit = 1;
    if (r[i600] > 0 && r[bottom] > 0) {
      const real smr = st_mixing(st_es(t[i600] - (real)273.15), p[i600], eps);
      const real sms = st_entropy(t[i600], smr, p[i600], RD, eps);
      const real sme = st_entropy(t[i600], MIN(r[i600], smr), p[i600], RD, eps);
      const real smb = st_entropy(sst, LV(dq, L, h), msl, RD, eps);
      const real smo = st_entropy(sst, st_mixing(st_es(sst - (real)273.15), msl, eps), msl, RD, eps);
      if (smo != smb) deficit = (sms - sme) / (smo - smb);
    }
    st_cape environmental = {-1, -1, -1};
    real vmax = -1;
    if (sst > (real)278.15 && !cold && msl > 0) {
      environmental = st_cape_of(L, t, r, p, t[bottom], r[bottom], p[bottom], RD);
      vmax = st_potential(L, t, r, p, sst, msl, RD, CKCD, environmental.energy, NULL);
    }
    real vi = 1;
    if (vmax > 0) vi = shear * deficit / vmax;
    const real rh = st_vapor(r[i600], p[i600], eps) / st_es(t[i600] - (real)273.15);
    const real v70 = vmax / 70, vent = 1 + (real)0.1 * shear;
    F(capen)[h] = environmental.energy;
    F(lnb)[h] = environmental.pressure;
    F(mpoti)[h] = vmax;
    F(chim)[h] = deficit;
    F(laav)[h] = vort;
    F(venti)[h] = vi;
    F(vrmpi)[h] = st_reduced(vi, K(vithresh));
    F(gpi)[h] = pow(fabs(vort * (real)1e5), (real)1.5) * pow(rh * 2, 3) * v70 * v70 * v70 / (vent * vent);
    F(shear)[h] = shear;
  }
}

/* physics/carbon/wgsl/carbonstep.wgsl:22-50 */
static void carbonstep(const M* v) {
  const int NH = v->NHOR, L = v->NLEV;
  for (int h = 0; h < NH; h++) {
    const real precip = (F(dprc)[h] + F(dprl)[h]) * 3.154e9;
    F(localprecip)[h] = precip;
    F(localavgprecip)[h] = F(localavgprecip)[h] + precip * K(timeweight);
    if (F(carbon_step)[0] == 0.0) continue;
    const real tsurf = LV(dt, L, h);
    if (K(ncarbon) != 0) {
      real weathering = 0;
      if (F(dls)[h] > 0.5 && tsurf >= 273.15) {
        const real pco2 = F(co2)[0] * F(dp)[h] * 1.0e-5;
        weathering = pow(pco2 / K(co2earth), K(cbeta)) * exp(K(kact) * (tsurf - 288.0)) * pow(precip / K(pearth), 0.65);
        if (K(nsupply) != 0) {
          const real wmaxp = MAX(K(wmax) / K(vearth), 2.0e-15);
          weathering = wmaxp * (1.0 - exp(-weathering / wmaxp));
        }
      }
      F(localweathering)[h] = weathering * (K(plarad) * K(plarad)) / (K(rad_earthsq) * 0.29);
    }
    F(localavgtemps)[h] = F(localavgtemps)[h] + tsurf * K(timeweight);
    F(localavgweather)[h] = F(localavgweather)[h] + F(localweathering)[h] * K(timeweight);
  }
}

/* ------------------------------------------------------------- snapshot */

/* snapshot/wgsl/pack.wgsl:14-107. Plane k of the snapshot is published
 * field k of GCM_FIELDS ORDER (fields/index.ts:81-128); plane 26 (sea-level
 * pressure) and the unset layer planes stay as they are. The tracking rows
 * hold the running minimum and maximum of the surface temperature (0, 1),
 * the last closed window's minimum, maximum and range (2-4), the sum and
 * count of closed ranges since the last publish (5, 6), whether a window
 * is open (7) and whether one has closed (8). */
void oracle_pack(const oracle_member* m) {
  M vv = view(m);
  const M* v = &vv;
  const int NH = v->NHOR, L = v->NLEV;
  /* gcm_publish.f90's jet level: the full level nearest sigma 0.25
   * (snapshot/index.ts:34-38) */
  int jet = 0;
  for (int l = 1; l < L; l++)
    if (fabs(T(sigma)[l] - 0.25) < fabs(T(sigma)[jet] - 0.25)) jet = l;
  const real* ss = F(snapshot_step);
  const unsigned step = (unsigned)ss[0], window = (unsigned)ss[1];
  real* snap = F(snapshot);
  real* tr = F(snapshot_tracking);
#define PUT(k, value) (snap[(long)(k) * NH + h] = (value))
#define TR(k) tr[(long)(k) * NH + h]
  for (int h = 0; h < NH; h++) {
    const real ts = LV(dt, L, h);
    if (TR(7) == 0) {
      TR(0) = ts;
      TR(1) = ts;
      TR(7) = 1;
    }
    TR(0) = MIN(TR(0), ts);
    TR(1) = MAX(TR(1), ts);
    if (step > 0 && window > 0 && step % window == 0) {
      const real range = TR(1) - TR(0);
      TR(2) = TR(0);
      TR(3) = TR(1);
      TR(4) = range;
      TR(5) = TR(5) + range;
      TR(6) = TR(6) + 1;
      TR(7) = 0;
      TR(8) = 1;
    }
    if (ss[2] == 0) continue;
    const real sw_top = LV(dswfl, 0, h), lw_top = LV(dlwfl, 0, h);
    const real sw_sfc = LV(dswfl, L, h), lw_sfc = LV(dlwfl, L, h);
    PUT(0, ts);                                   /* surfaceTemperature */
    PUT(1, F(dtsa)[h]);                           /* airTemperature */
    PUT(2, LV(du, L - 1, h));                     /* windU, lowest level */
    PUT(3, LV(dv, L - 1, h));
    PUT(4, LV(dcc, L, h));                        /* total cloud cover */
    PUT(5, F(dprc)[h] + F(dprl)[h]);
    PUT(6, F(dp)[h]);
    PUT(7, F(dqvi)[h]);
    PUT(8, F(dicec)[h]);
    PUT(9, F(dsnow)[h]);
    PUT(10, F(dwatc)[h]);
    PUT(11, F(dalb)[h]);
    PUT(12, LV(du, jet, h));
    PUT(13, LV(dv, jet, h));
    PUT(14, -lw_top);                             /* outgoing longwave */
    PUT(15, F(dls)[h]);
    PUT(16, sw_top + lw_top);                     /* TOA net */
    /* hfns = snow melt x L_f rho_water + rss + rls + hfss + hfls */
    PUT(17, F(dsmelt)[h] * (real)-333700000.0 + sw_sfc + lw_sfc + F(dshfl)[h] + F(dlhfl)[h]);
    PUT(18, -F(devap)[h]);
    PUT(19, F(drunoff)[h]);
    PUT(20, F(dforest)[h]);
    PUT(21, F(dlai)[h]);
    PUT(22, F(dprs)[h]);
    PUT(23, F(dsmelt)[h]);
    PUT(24, F(diced)[h]);
    PUT(25, F(dglac)[h]);
    PUT(27, sw_sfc);
    PUT(28, lw_sfc);
    PUT(29, F(dshfl)[h]);
    PUT(30, F(dlhfl)[h]);
    PUT(31, F(dprc)[h]);
    PUT(32, F(dprl)[h]);
    PUT(33, sw_top);
    PUT(34, lw_top);
    const int closed = TR(8) != 0;
    PUT(35, closed ? TR(2) : TR(0));
    PUT(36, closed ? TR(3) : TR(1));
    PUT(37, closed ? TR(4) : TR(1) - TR(0));
    PUT(38, TR(5));
    PUT(39, TR(6));
    PUT(40, F(dnpp)[h]);
    PUT(41, F(gmu0)[h]);
    TR(5) = 0;
    TR(6) = 0;
    PUT(42, F(capen)[h]);
    PUT(43, F(lnb)[h]);
    PUT(44, F(mpoti)[h]);
    PUT(45, F(chim)[h]);
    PUT(46, F(laav)[h]);
    PUT(47, F(venti)[h]);
    PUT(48, F(vrmpi)[h]);
    PUT(49, F(gpi)[h]);
    PUT(50, F(shear)[h]);
    for (int l = 0; l < L; l++) {
      const int k = 51 + 5 * l;
      PUT(k, LV(dt, l, h));
      PUT(k + 1, LV(du, l, h));
      PUT(k + 2, LV(dv, l, h));
      PUT(k + 4, LV(dcc, l, h));
    }
    /* relative humidity of the lowest level only */
    const real qs = sat_divided(v, LV(dt, L - 1, h), F(dp)[h] * T(sigma)[L - 1]);
    PUT(51 + 5 * (L - 1) + 3, MAX(0.0, LV(dq, L - 1, h)) / MAX((real)1.17549435e-38, qs));
  }
#undef PUT
#undef TR
}

/* ------------------------------------------------------------- sequence */

/* gridpointd's parameterizations in the reference's dispatch order
 * (stepper/index.ts:148-161; fixer/index.ts:39-73; surface-flux/index.ts:
 * 72-113; radiation/index.ts:151-234; precipitation/index.ts:71-160;
 * surface/index.ts:62-144; carbon/index.ts:32-45). */
/* One parameterization by name, for the invariant checks of the oracle's
 * routines; -1 for an unknown name. "fixer" runs the column, row, global
 * and apply passes; "rain" runs mklsp, mkrain and mkclouds per column. */
/* The fits of lwr (which: 0 H2O absorptivity, 1 CO2 absorptivity, 2 H2O
 * transmission in the CO2 band, 3 O3 absorptivity) or of swr (4 ozone
 * absorption, 5 water vapour absorption) at n absorber paths; -1 for an
 * unknown fit. For the pins of the fits alone. */
int oracle_radiation_fit(const oracle_member* m, int which, const real* u, int n, real* out) {
  M vv = view(m);
  const M* v = &vv;
  for (int i = 0; i < n; i++) {
    switch (which) {
      case 0: out[i] = lwr_ah2o(v, u[i]); break;
      case 1: out[i] = lwr_aco2(v, u[i]); break;
      case 2: out[i] = lwr_th2o(v, u[i]); break;
      case 3: out[i] = lwr_ao3(v, u[i]); break;
      case 4: out[i] = ozone_absorption(v, u[i]); break;
      case 5: out[i] = vapour_absorption(v, u[i]); break;
      default: return -1;
    }
  }
  return 0;
}

/* storm's potential-intensity iteration for column h of the arena (the
 * column storm() builds); out[0..4] as st_potential's parts, out[5] the
 * intensity. -1 when the column is skipped (cool water, cold levels) or
 * the iteration fails. */
int oracle_storm_potential(const oracle_member* m, int h, real* out) {
  M vv = view(m);
  const M* v = &vv;
  const int L = v->NLEV, NH = v->NHOR;
  const real RD = K(gascon);
  real t[16], r[16], p[16];
  const real msl = F(dp)[h] * (real)0.01, sst = LV(dt, L, h);
  int cold = 0;
  for (int l = 0; l < L; l++) {
    t[l] = LV(dt, l, h);
    r[l] = LV(dq, l, h);
    p[l] = msl * T(sigma)[l];
    cold = cold || t[l] <= 100;
  }
  (void)NH;
  if (!(sst > (real)278.15 && !cold && msl > 0)) return -1;
  const st_cape env = st_cape_of(L, t, r, p, t[L - 1], r[L - 1], p[L - 1], RD);
  out[5] = st_potential(L, t, r, p, sst, msl, RD, K(ckcd), env.energy, out);
  return out[5] < 0 ? -1 : 0;
}

int oracle_routine(const oracle_member* m, const char* name) {
  M vv = view(m);
  const M* v = &vv;
#define IS(s) (strcmp(name, s) == 0)
  if (IS("fixer")) {
    fixer_column(v);
    fixer_row(v);
    global_sum(v, F(fixneg), F(fixpos), F(fixsum));
    fixer_apply(v);
  } else if (IS("surflx")) {
    surflx(v);
  } else if (IS("vdiff")) {
    vdiff(v);
  } else if (IS("solang")) {
    solang(v);
  } else if (IS("mko3")) {
    mko3(v);
  } else if (IS("swr")) {
    swr(v);
  } else if (IS("lwr")) {
    lwr(v);
  } else if (IS("radstep")) {
    radstep(v);
  } else if (IS("kuo")) {
    kuo(v);
  } else if (IS("mkdca")) {
    mkdca(v);
  } else if (IS("mklsp")) {
    for (int h = 0; h < v->NHOR; h++) mklsp(v, h);
  } else if (IS("mkrain")) {
    for (int h = 0; h < v->NHOR; h++) mkrain(v, h);
  } else if (IS("soil")) {
    soil(v);
  } else if (IS("mkradv")) {
    mkradv(v);
  } else if (IS("mkclouds")) {
    for (int h = 0; h < v->NHOR; h++) mkclouds(v, h);
  } else if (IS("land")) {
    land_surface(v);
  } else if (IS("sea_ice")) {
    sea_ice(v);
  } else if (IS("carbon")) {
    carbonstep(v);
  } else if (IS("storm")) {
    storm(v);
  } else if (IS("seastep")) {
    seastep(v);
  } else {
    return -1;
  }
#undef IS
  return 0;
}

void oracle_physics(const oracle_member* m) {
  M vv = view(m);
  const M* v = &vv;
  const int L = v->NLEV;
  fixer_column(v);
  fixer_row(v);
  global_sum(v, F(fixneg), F(fixpos), F(fixsum));
  fixer_apply(v);
  surflx(v);
  vdiff(v);
  solang(v);
  if (K(ozone) != 0) mko3(v);
  swr(v);
  lwr(v);
  radstep(v);
  /* mkdqtgp: sqt to the grid (precipitation/index.ts:103-113) */
  oracle_transform_inverse(m, F(sqt), F(dqt), L, 0);
  kuo(v);
  mkdca(v);
  mkdca_limit(v);
  rain_column(v);
  soil(v);
  mkradv(v);
  land_surface(v);
  if (K(ice_coupled) != 0) {
    sea_ice(v);
    global_sum(v, F(ice_part1), F(ice_part2), F(ice_sums));
  }
  seastep(v);
  carbonstep(v);
  if (K(nstormdiag) > 0) storm(v);
}
