/*
 * FP64 CPU oracle of the ExoPlaSim model step: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement, in real precision and in the reference's own
 * dispatch order, of every WGSL kernel of one model step
 * (src/model/climate/gcm/stepper/index.ts:125-162): the Legendre and
 * Fourier transforms with their Fourier intermediates, gridpointa,
 * spectrala, gridpointd's frame and parameterizations, spectrald. It works
 * on an arena of doubles laid out like the engine's FP32 arena; the caller
 * resolves every slot, table and constant by name (oracle_slot_names() and
 * friends), so this file shares no code or header with the engine.
 *
 * Only the tests, the smoke check and the CPU baseline of the benchmark
 * call it, as the checker; nothing in the shipped library links it.
 */
#define EXO_ORACLE_INTERNAL
#include "exo_oracle.h"


#include <tgmath.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- names */

const char* const* oracle_slot_names(void) { return oracle_slot_names_; }
const char* const* oracle_table_names(void) { return oracle_table_names_; }
const char* const* oracle_constant_names(void) { return oracle_const_names_; }


/* ------------------------------------------------------------ transforms */

/* sp2fc / sp2fcdmu: legendre/wgsl/inverse.wgsl:6-27 */
static void sp2fc(const M* v, const real* sp, real* fc, int planes, const real* q) {
  for (int p = 0; p < planes; p++)
    for (int lat = 0; lat < v->NLAT; lat++)
      for (int m = 0; m < v->NTP1; m++) {
        const int first = m * (2 * v->NTP1 - m + 1) / 2;
        real re = 0, im = 0;
        for (int w = first; w < first + v->NTP1 - m; w++) {
          const real qq = q[lat * v->NCSP + w];
          re += qq * sp[p * v->NRSP + 2 * w];
          im += qq * sp[p * v->NRSP + 2 * w + 1];
        }
        const long row = (long)(p * v->NLAT + lat) * v->NFC;
        fc[row + 2 * m] = re;
        fc[row + 2 * m + 1] = im;
      }
}

static void mode_of(int w, int ntp1, int* mo, int* no) {
  int m = 0, start = 0;
  while (w >= start + ntp1 - m) {
    start += ntp1 - m;
    m++;
  }
  *mo = m;
  *no = m + w - start;
}

/* fc2sp: legendre/wgsl/forward.wgsl:7-38 */
static void fc2sp(const M* v, const real* fc, real* sp, int planes, const real* q) {
  for (int p = 0; p < planes; p++)
    for (int w = 0; w < v->NCSP; w++) {
      int m, n;
      mode_of(w, v->NTP1, &m, &n);
      const real sign = (m + n) % 2 == 0 ? 1.0 : -1.0;
      const real* f = fc + (long)p * v->NLAT * v->NFC + 2 * m;
      real re = 0, im = 0;
      for (int l = 0; l < v->NLAT / 2; l++) {
        const real qq = q[l * v->NCSP + w];
        const real* no = f + l * v->NFC;
        const real* so = f + (v->NLAT - 1 - l) * v->NFC;
        re += qq * (no[0] + sign * so[0]);
        im += qq * (no[1] + sign * so[1]);
      }
      sp[p * v->NRSP + 2 * w] = re;
      sp[p * v->NRSP + 2 * w + 1] = im;
    }
}

/* fc2gp: fourier/wgsl/inverse.wgsl:4-18 */
static void fc2gp(const M* v, const real* fc, real* gp, int planes) {
  const real* inv = TB(*v, fourier_inverse);
  for (int row = 0; row < planes * v->NLAT; row++)
    for (int j = 0; j < v->NLON; j++) {
      real sum = 0;
      for (int k = 0; k < v->NFC; k++) sum += inv[k * v->NLON + j] * fc[(long)row * v->NFC + k];
      gp[(long)row * v->NLON + j] = sum;
    }
}

/* gp2fc: fourier/wgsl/forward.wgsl:4-18 */
static void gp2fc(const M* v, const real* gp, real* fc, int planes) {
  const real* fwd = TB(*v, fourier_forward);
  for (int row = 0; row < planes * v->NLAT; row++)
    for (int k = 0; k < v->NFC; k++) {
      real sum = 0;
      for (int j = 0; j < v->NLON; j++) sum += fwd[j * v->NFC + k] * gp[(long)row * v->NLON + j];
      fc[(long)row * v->NFC + k] = sum;
    }
}

/* dlambda: fourier/wgsl/dlambda.wgsl:5-17 */
static void dlambda(const M* v, const real* src, real* dst, int planes) {
  for (int row = 0; row < planes * v->NLAT; row++)
    for (int m = 0; m < v->NTP1; m++) {
      const long at = (long)row * v->NFC + 2 * m;
      const real re = src[at], im = src[at + 1];
      dst[at] = -im * m;
      dst[at + 1] = re * m;
    }
}

/* dv2uv: legendre/wgsl/dv2uv.wgsl:9-43 */
static void dv2uv(const M* v, const real* d, const real* z, real* u, real* vv, real plavor) {
  const real* qu = TB(*v, legendre_qu);
  const real* qv = TB(*v, legendre_qv);
  for (int p = 0; p < v->NLEV; p++)
    for (int lat = 0; lat < v->NLAT; lat++)
      for (int m = 0; m < v->NTP1; m++) {
        const int first = m * (2 * v->NTP1 - m + 1) / 2;
        const real* dd = d + p * v->NRSP;
        const real* zz = z + p * v->NRSP;
        real ure = 0, uim = 0, vre = 0, vim = 0;
        for (int w = first; w < first + v->NTP1 - m; w++) {
          const real a = qu[lat * v->NCSP + w], b = qv[lat * v->NCSP + w];
          const real zre = w == 1 ? zz[2 * w] - plavor : zz[2 * w];
          const real zim = zz[2 * w + 1];
          const real dre = dd[2 * w], dim = dd[2 * w + 1];
          ure = ure + b * zre + a * dim;
          uim = uim + b * zim - a * dre;
          vre = vre + a * zim - b * dre;
          vim = vim - a * zre - b * dim;
        }
        const long at = (long)(p * v->NLAT + lat) * v->NFC + 2 * m;
        u[at] = ure;
        u[at + 1] = uim;
        vv[at] = vre;
        vv[at + 1] = vim;
      }
}

/* uv2dv / mktend: legendre/wgsl/uv2dv.wgsl:11-63 */
static void uv2dv(const M* v, const real* U, const real* V, const real* K, real* D, real* Z) {
  const real* qe = TB(*v, legendre_qe);
  const real* qm = TB(*v, legendre_qm);
  const real* qq = TB(*v, legendre_qq);
  for (int p = 0; p < v->NLEV; p++)
    for (int w = 0; w < v->NCSP; w++) {
      int m, n;
      mode_of(w, v->NTP1, &m, &n);
      const real s = (m + n) % 2 == 0 ? 1.0 : -1.0;
      const long col = (long)p * v->NLAT * v->NFC + 2 * m;
      real dre = 0, dim = 0, zre = 0, zim = 0;
      for (int l = 0; l < v->NLAT / 2; l++) {
        const long no = col + l * v->NFC, so = col + (v->NLAT - 1 - l) * v->NFC;
        const real e = qe[l * v->NCSP + w], mm = qm[l * v->NCSP + w];
        const real dure = U[no] - s * U[so], duim = U[no + 1] - s * U[so + 1];
        const real dvre = V[no] - s * V[so], dvim = V[no + 1] - s * V[so + 1];
        const real sure = U[no] + s * U[so], suim = U[no + 1] + s * U[so + 1];
        const real svre = V[no] + s * V[so], svim = V[no + 1] + s * V[so + 1];
        if (K) {
          const real c = qq[l * v->NCSP + w];
          dre = dre + c * (K[no] + s * K[so]);
          dim = dim + c * (K[no + 1] + s * K[so + 1]);
        }
        dre = dre - e * dvre - mm * suim;
        dim = dim - e * dvim + mm * sure;
        zre = zre + e * dure - mm * svim;
        zim = zim + e * duim + mm * svre;
      }
      const long at = (long)p * v->NRSP + 2 * w;
      D[at] = dre;
      D[at + 1] = dim;
      Z[at] = zre;
      Z[at + 1] = zim;
    }
}

/* qtend: legendre/wgsl/qtend.wgsl:9-46 */
static void qtend(const M* v, const real* U, const real* V, const real* S, real* out) {
  const real* qe = TB(*v, legendre_qe);
  const real* qc = TB(*v, legendre_qc);
  const real* qm = TB(*v, legendre_qm);
  for (int p = 0; p < v->NLEV; p++)
    for (int w = 0; w < v->NCSP; w++) {
      int m, n;
      mode_of(w, v->NTP1, &m, &n);
      const real s = (m + n) % 2 == 0 ? 1.0 : -1.0;
      const long col = (long)p * v->NLAT * v->NFC + 2 * m;
      real re = 0, im = 0;
      for (int l = 0; l < v->NLAT / 2; l++) {
        const long no = col + l * v->NFC, so = col + (v->NLAT - 1 - l) * v->NFC;
        const real e = qe[l * v->NCSP + w], c = qc[l * v->NCSP + w], mm = qm[l * v->NCSP + w];
        re = re + e * (V[no] - s * V[so]) + c * (S[no] + s * S[so]) + mm * (U[no + 1] + s * U[so + 1]);
        im = im + e * (V[no + 1] - s * V[so + 1]) + c * (S[no + 1] + s * S[so + 1]) - mm * (U[no] + s * U[so]);
      }
      out[(long)p * v->NRSP + 2 * w] = re;
      out[(long)p * v->NRSP + 2 * w + 1] = im;
    }
}

/* --------------------------------------------------------------- columns */

/* calcgp: dynamics/wgsl/calcgp.wgsl:14-147 */
static void calcgp(const M* v) {
  const int L = v->NLEV, NH = v->NHOR;
  const real PSURF = KK(*v, psurf), WW = KK(*v, ww), CV = KK(*v, cv);
  const real AKAP = KK(*v, akap), RDBRV = KK(*v, rdbrv), ADV = KK(*v, adv);
  const real *ds = TB(*v, dsigma), *sh = TB(*v, sigmah), *rd = TB(*v, rdsig);
  const real *t01 = TB(*v, t01s2), *tkp = TB(*v, tkp), *c = TB(*v, c), *g = TB(*v, g), *t0 = TB(*v, t0);
  real gu[16], gv[16], gt[16], gd[16], gz[16], gq[16], zvgpg[16], ztv1[16], ztv2[16];
  real zsdotp[16], gtd[16], gqm[16], gud[16], gvd[16], gtn[16], gqn[16], guz[16], gvz[16], dw[16];
  for (int h = 0; h < NH; h++) {
    const real rcsq = TB(*v, rcsq)[h];
    const real ps = exp(A(*v, gp)[h]);
    A(*v, gp)[h] = ps;
    const real gpm = A(*v, gpmt)[h], gpj = A(*v, gpj)[h];
    for (int l = 0; l < L; l++) {
      gu[l] = A(*v, gu)[l * NH + h];
      gv[l] = A(*v, gv)[l * NH + h];
      gt[l] = A(*v, gt)[l * NH + h];
      gd[l] = A(*v, gd)[l * NH + h];
      gz[l] = A(*v, gz)[l * NH + h];
      gq[l] = A(*v, gq)[l * NH + h];
      zvgpg[l] = rcsq * (gu[l] * gpm + gv[l] * gpj);
      const real zq = fmax(gq[l] / ps, 0.0);
      const real tvirt = (gt[l] + t0[l]) * (1.0 + (1.0 / RDBRV - 1.0) * zq);
      ztv1[l] = tvirt - t0[l];
      ztv2[l] = tvirt / (1.0 + ADV * zq) - t0[l];
    }
    real zsumd = ds[0] * gd[0];
    real gvp = ds[0] * zvgpg[0];
    zsdotp[0] = zsumd + gvp;
    for (int l = 1; l < L - 1; l++) {
      zsumd = zsumd + ds[l] * gd[l];
      gvp = gvp + ds[l] * zvgpg[l];
      zsdotp[l] = zsumd + gvp;
    }
    zsumd = zsumd + ds[L - 1] * gd[L - 1];
    gvp = gvp + ds[L - 1] * zvgpg[L - 1];
    for (int l = 0; l < L - 1; l++) {
      zsdotp[l] = sh[l] * (zsumd + gvp) - zsdotp[l];
      gtd[l] = zsdotp[l] * (gt[l + 1] - gt[l]);
      gqm[l] = zsdotp[l] * (gq[l + 1] + gq[l]);
      gud[l] = zsdotp[l] * (gu[l + 1] - gu[l]);
      gvd[l] = zsdotp[l] * (gv[l + 1] - gv[l]);
    }
    zsumd = zvgpg[0] * ds[0];
    gtn[0] = gt[0] * gd[0] - AKAP * ztv2[0] * gd[0] - rd[0] * (gtd[0] + t01[0] * (sh[0] * gvp - zsumd));
    gqn[0] = -rd[0] * gqm[0];
    guz[0] = -gu[0] * gz[0] - gpj * ztv1[0] - rd[0] * gvd[0];
    gvz[0] = gv[0] * gz[0] - gpm * ztv1[0] - rd[0] * gud[0];
    dw[0] = gd[0] * ps * PSURF * WW;
    for (int l = 1; l < L; l++) {
      const int column = l * L;
      real ztpta = c[column] * zvgpg[0];
      real ztptb = c[column] * (zvgpg[0] + gd[0]);
      for (int j = 1; j <= l; j++) {
        ztpta = ztpta + c[column + j] * zvgpg[j];
        ztptb = ztptb + c[column + j] * (zvgpg[j] + gd[j]);
      }
      if (l < L - 1) {
        zsumd = zsumd + zvgpg[l] * ds[l];
        gtn[l] = gt[l] * gd[l] + AKAP * ztv2[l] * (zvgpg[l] - ztptb) + tkp[l] * (zvgpg[l] - ztpta) -
                 rd[l] * (gtd[l] + gtd[l - 1] + gvp * (t01[l] * sh[l] + t01[l - 1] * sh[l - 1]) -
                          zsumd * (t01[l - 1] + t01[l]) + zvgpg[l] * ds[l] * t01[l - 1]);
        gqn[l] = -rd[l] * (gqm[l] - gqm[l - 1]);
        guz[l] = -gu[l] * gz[l] - gpj * ztv1[l] - rd[l] * (gvd[l] + gvd[l - 1]);
        gvz[l] = gv[l] * gz[l] - gpm * ztv1[l] - rd[l] * (gud[l] + gud[l - 1]);
      } else {
        gtn[l] = gt[l] * gd[l] + AKAP * ztv2[l] * (zvgpg[l] - ztptb) + tkp[l] * (zvgpg[l] - ztpta) -
                 rd[l] * (gtd[L - 2] + t01[l - 1] * (sh[l - 1] * gvp - zsumd));
        gqn[l] = rd[l] * gqm[L - 2];
        guz[l] = -gu[l] * gz[l] - gpj * ztv1[l] - rd[l] * gvd[L - 2];
        gvz[l] = gv[l] * gz[l] - gpm * ztv1[l] - rd[l] * gud[L - 2];
      }
      dw[l] = (zvgpg[l] - ztptb) * ps * PSURF * WW;
    }
    const real scale = sqrt(rcsq);
    for (int l = 0; l < L; l++) {
      real gsum = 0;
      for (int k = 0; k < L; k++) gsum = gsum + g[l * L + k] * (ztv1[k] - gt[k]);
      const real gphi = gsum * 2.0 / rcsq;
      const long at = (long)l * NH + h;
      A(*v, gtn)[at] = gtn[l];
      A(*v, gqn)[at] = gqn[l];
      A(*v, guz)[at] = guz[l];
      A(*v, gvz)[at] = gvz[l];
      A(*v, dw)[at] = dw[l];
      A(*v, gut)[at] = gu[l] * gt[l];
      A(*v, gvt)[at] = gv[l] * gt[l];
      A(*v, gke)[at] = gu[l] * gu[l] + gv[l] * gv[l] + gphi;
      A(*v, guq)[at] = gu[l] * gq[l];
      A(*v, gvq)[at] = gv[l] * gq[l];
      A(*v, du0)[at] = CV * gu[l] * scale;
      A(*v, dv0)[at] = CV * gv[l] * scale;
    }
    A(*v, gvpp)[h] = gvp;
    A(*v, dp0)[h] = PSURF * ps;
  }
}

/* spectrala: dynamics/wgsl/spectrala.wgsl:15-109 */
static void spectrala(const M* v) {
  const int L = v->NLEV, NR = v->NRSP;
  const real DELT = KK(*v, delt), DELT2 = KK(*v, delt2), PNU = KK(*v, pnu), PNU21 = KK(*v, pnu21);
  const int FILTER = KK(*v, filter) != 0;
  const real *g = TB(*v, g), *tau = TB(*v, tau), *bm1 = TB(*v, bm1), *t0 = TB(*v, t0), *ds = TB(*v, dsigma);
  real adm[16], azm[16], atm[16], aqm[16], adt[16], stt[16], zgt[16], zgm[16], sdt[16];
  for (int s = 0; s < NR; s++) {
    const real apm = A(*v, spm)[s];
    for (int l = 0; l < L; l++) {
      adm[l] = A(*v, sdm)[l * NR + s];
      azm[l] = A(*v, szm)[l * NR + s];
      atm[l] = A(*v, stm)[l * NR + s];
      aqm[l] = A(*v, sqm)[l * NR + s];
      adt[l] = A(*v, sdt)[l * NR + s];
      stt[l] = A(*v, stt)[l * NR + s];
    }
    for (int l = 0; l < L; l++) {
      real gt = 0, gm = 0;
      for (int k = 0; k < L; k++) {
        gt = gt + g[l * L + k] * stt[k];
        gm = gm + g[l * L + k] * atm[k];
      }
      zgt[l] = gt;
      zgm[l] = gm;
    }
    real spt = A(*v, spt)[s];
    const real spm = apm;
    const real sop = TB(*v, sop)[s];
    const int jn = (int)TB(*v, nindex)[s];
    for (int l = 0; l < L; l++) {
      real zsum = 0;
      if (jn > 0) {
        const real zq = 1.0 / (real)(jn * (jn + 1));
        for (int l2 = 0; l2 < L; l2++) {
          const real zt = zgt[l2] - t0[l2] * spt;
          const real zm = zgm[l2] + t0[l2] * spm;
          const real za = adt[l2] * zq + sop;
          const real zb = adm[l2] * zq;
          const real zz = zb + DELT * (zm + za + DELT * zt);
          zsum = zsum + zz * bm1[((jn - 1) * L + l) * L + l2];
        }
      }
      sdt[l] = s < 2 ? 0.0 : zsum;
    }
    for (int l = 0; l < L; l++) spt = spt + ds[l] * sdt[l];
    for (int l = 0; l < L; l++) {
      real t = 0;
      for (int k = 0; k < L; k++) t = t + tau[l * L + k] * sdt[k];
      stt[l] = stt[l] - t;
    }
    const real spp = A(*v, spp)[s];
    if (FILTER) A(*v, spm)[s] = PNU21 * spp + PNU * apm;
    const real sppn = s < 2 ? 0.0 : apm - DELT2 * spt;
    A(*v, spp)[s] = sppn;
    A(*v, sp)[s] = sppn;
    A(*v, spt)[s] = spt;
    for (int l = 0; l < L; l++) {
      const long at = (long)l * NR + s;
      if (FILTER) {
        A(*v, sdm)[at] = PNU21 * A(*v, sdp)[at] + PNU * adm[l];
        A(*v, szm)[at] = PNU21 * A(*v, szp)[at] + PNU * azm[l];
        A(*v, stm)[at] = PNU21 * A(*v, stp)[at] + PNU * atm[l];
        A(*v, sqm)[at] = PNU21 * A(*v, sqp)[at] + PNU * aqm[l];
      }
      const real sdp = 2.0 * sdt[l] - adm[l];
      const real stp = DELT2 * stt[l] + atm[l];
      const real szp = DELT2 * A(*v, szt)[at] + azm[l];
      const real sqp = DELT2 * A(*v, sqt)[at] + aqm[l];
      A(*v, sdt)[at] = sdt[l];
      A(*v, stt)[at] = stt[l];
      A(*v, sdp)[at] = sdp;
      A(*v, stp)[at] = stp;
      A(*v, szp)[at] = szp;
      A(*v, sqp)[at] = sqp;
      A(*v, sd)[at] = sdp;
      A(*v, st)[at] = stp;
      A(*v, sz)[at] = szp;
      A(*v, sq)[at] = sqp;
    }
  }
}

/* spectrald up to mkdheat: dynamics/wgsl/spectrald-diffusion.wgsl:12-60 */
static void spectrald_diffusion(const M* v) {
  const int L = v->NLEV, NR = v->NRSP;
  const real DELT2 = KK(*v, delt2), PLAVOR = KK(*v, plavor);
  for (int s = 0; s < NR; s++) {
    if (s < 2) A(*v, spp)[s] = 0.0;
    for (int l = 0; l < L; l++) {
      const long at = (long)l * NR + s;
      const real szp = A(*v, szp)[at] + DELT2 * A(*v, szt)[at];
      const real stp = A(*v, stp)[at] + DELT2 * A(*v, stt)[at];
      const real sdp = A(*v, sdp)[at] + DELT2 * A(*v, sdt)[at];
      const real sqp = A(*v, sqp)[at] + DELT2 * A(*v, sqt)[at];
      const real sakpp = TB(*v, sakpp)[at];
      const real damp = TB(*v, damp)[l], tdisst = TB(*v, tdisst)[l], tdissq = TB(*v, tdissq)[l];
      const real tdissd = TB(*v, tdissd)[l], tdissz = TB(*v, tdissz)[l], tfrc = TB(*v, tfrc)[l];
      const real stt = ((-damp - tdisst * sakpp) * stp + damp * TB(*v, srp)[at]) /
                         (1.0 + DELT2 * (damp + tdisst * sakpp));
      const real sqt = -tdissq * sakpp * sqp / (1.0 + DELT2 * tdissq * sakpp);
      const real zsdt1 = -tfrc * sdp / (1.0 + DELT2 * tfrc);
      const real zsdt2 = -tdissd * sakpp * sdp / (1.0 + DELT2 * tdissd * sakpp);
      real zszt1 = -tfrc * szp / (1.0 + DELT2 * tfrc);
      real zszt2 = -tdissz * sakpp * szp / (1.0 + DELT2 * tdissz * sakpp);
      real szt = zszt1 + zszt2;
      if (s == 2) {
        zszt1 = zszt1 + PLAVOR * tfrc / (1.0 + DELT2 * tfrc);
        zszt2 = zszt2 + PLAVOR * tdissz * sakpp / (1.0 + DELT2 * tdissz * sakpp);
        szt = szt + PLAVOR * (tfrc / (1.0 + DELT2 * tfrc) + tdissz * sakpp / (1.0 + DELT2 * tdissz * sakpp));
      }
      A(*v, szp)[at] = szp;
      A(*v, stp)[at] = stp;
      A(*v, sdp)[at] = sdp;
      A(*v, sqp)[at] = sqp;
      A(*v, stt)[at] = stt;
      A(*v, sqt)[at] = sqt;
      A(*v, sdt)[at] = zsdt1 + zsdt2;
      A(*v, szt)[at] = szt;
      A(*v, zdsdp1)[at] = zsdt1 * DELT2;
      A(*v, zdszp1)[at] = zszt1 * DELT2;
      A(*v, zdsdp2)[at] = zsdt2 * DELT2;
      A(*v, zdszp2)[at] = zszt2 * DELT2;
    }
  }
}

/* mkdheat grid-point stages: dynamics/wgsl/mkdheat.wgsl:17-44 */
static void mkdheat(const M* v, int stage) {
  const int NH = v->NHOR;
  const real CV = KK(*v, cv), CT = KK(*v, ct), WW = KK(*v, ww), GA = KK(*v, ga);
  const real ACPD = KK(*v, acpd), ADV = KK(*v, adv), DELTSEC2 = KK(*v, deltsec2);
  for (long i = 0; i < (long)v->NLEV * NH; i++) {
    const int l = (int)(i / NH), h = (int)(i % NH);
    const real heat = 1.0 + ADV * A(*v, dq)[i];
    const real dsig = TB(*v, dsigma)[l];
    if (stage == 3) {
      const real dtdt = -A(*v, zdekin)[i] * 0.5 / ACPD / heat / A(*v, dp)[h] * GA / dsig;
      A(*v, zdtdt)[i] = dtdt / CT / WW;
      continue;
    }
    const real scale = CV * sqrt(TB(*v, rcsq)[h]);
    const real u = A(*v, zu)[i] * scale, vv = A(*v, zv)[i] * scale;
    const real du = A(*v, zdu)[i] * scale, dv = A(*v, zdv)[i] * scale;
    const real change = du * (2.0 * u + du) + dv * (2.0 * vv + dv);
    if (stage == 1)
      A(*v, zdtdt)[i] = (-change / DELTSEC2 * 0.5 / ACPD / heat) / CT / WW;
    else
      A(*v, zdekin)[i] = change / DELTSEC2 * A(*v, dp)[h] / GA * dsig;
  }
}

/* spectrald after mkdheat: dynamics/wgsl/spectrald-update.wgsl:12-45 */
static void spectrald_update(const M* v) {
  const int L = v->NLEV, NR = v->NRSP;
  const real DELT2 = KK(*v, delt2), PNU = KK(*v, pnu);
  const int FILTER = KK(*v, filter) != 0;
  for (int s = 0; s < NR; s++) {
    const real spp = A(*v, spp)[s];
    if (FILTER) A(*v, spm)[s] = A(*v, spm)[s] + PNU * spp;
    A(*v, sp)[s] = spp;
    for (int l = 0; l < L; l++) {
      const long at = (long)l * NR + s;
      const real stt = A(*v, stt)[at] + A(*v, zstt1)[at] + A(*v, zstt2)[at];
      A(*v, stt)[at] = stt;
      const real szp = A(*v, szp)[at] + DELT2 * A(*v, szt)[at];
      const real stp = A(*v, stp)[at] + DELT2 * stt;
      const real sdp = A(*v, sdp)[at] + DELT2 * A(*v, sdt)[at];
      const real sqp = A(*v, sqp)[at] + DELT2 * A(*v, sqt)[at];
      A(*v, szp)[at] = szp;
      A(*v, stp)[at] = stp;
      A(*v, sdp)[at] = sdp;
      A(*v, sqp)[at] = sqp;
      if (FILTER) {
        A(*v, szm)[at] = A(*v, szm)[at] + PNU * szp;
        A(*v, stm)[at] = A(*v, stm)[at] + PNU * stp;
        A(*v, sdm)[at] = A(*v, sdm)[at] + PNU * sdp;
        A(*v, sqm)[at] = A(*v, sqm)[at] + PNU * sqp;
      }
      A(*v, sz)[at] = szp;
      A(*v, st)[at] = stp;
      A(*v, sd)[at] = sdp;
      A(*v, sq)[at] = sqp;
    }
  }
}

/* physics frame in: dynamics/wgsl/physics-in.wgsl:8-34 */
static void physics_in(const M* v) {
  const int L = v->NLEV, NH = v->NHOR;
  const real PSURF = KK(*v, psurf), CV = KK(*v, cv), CT = KK(*v, ct);
  for (int h = 0; h < NH; h++) {
    const real dp = PSURF * exp(A(*v, gp)[h]);
    A(*v, dp)[h] = dp;
    const real scale = sqrt(TB(*v, rcsq)[h]);
    for (int l = 0; l <= L; l++) {
      const long at = (long)l * NH + h;
      A(*v, dudt)[at] = 0;
      A(*v, dvdt)[at] = 0;
      A(*v, dtdt)[at] = 0;
      A(*v, dqdt)[at] = 0;
      if (l < L) {
        A(*v, gudt)[at] = 0;
        A(*v, gvdt)[at] = 0;
        A(*v, gtdt)[at] = 0;
        A(*v, gqdt)[at] = 0;
        A(*v, du)[at] = CV * A(*v, gu)[at] * scale;
        A(*v, dv)[at] = CV * A(*v, gv)[at] * scale;
        A(*v, dt)[at] = CT * (A(*v, gt)[at] + TB(*v, t0)[l]);
        A(*v, dq)[at] = A(*v, gq)[at] * PSURF / dp;
      }
    }
  }
}

/* physics frame out: dynamics/wgsl/physics-out.wgsl:9-21 */
static void physics_out(const M* v) {
  const int NH = v->NHOR;
  const real PSURF = KK(*v, psurf), CV = KK(*v, cv), CT = KK(*v, ct), WW = KK(*v, ww);
  for (long i = 0; i < (long)v->NLEV * NH; i++) {
    const int h = (int)(i % NH);
    const real scale = sqrt(TB(*v, rcsq)[h]);
    A(*v, gudt)[i] = A(*v, dudt)[i] / scale / CV / WW + A(*v, gudt)[i];
    A(*v, gvdt)[i] = A(*v, dvdt)[i] / scale / CV / WW + A(*v, gvdt)[i];
    A(*v, gtdt)[i] = A(*v, dtdt)[i] / CT / WW + A(*v, gtdt)[i];
    A(*v, gqdt)[i] = A(*v, dqdt)[i] * A(*v, dp)[h] / WW / PSURF + A(*v, gqdt)[i];
  }
}

/* ----------------------------------------------------------------- step */

void oracle_transform_inverse(const oracle_member* m, const real* sp, real* grid, int planes, int dmu) {
  M v = view(m);
  const real* q = dmu ? TB(v, legendre_qj) : TB(v, legendre_qi);
  real* fc = (real*)malloc(sizeof(real) * (size_t)planes * v.NLAT * v.NFC);
  sp2fc(&v, sp, fc, planes, q);
  fc2gp(&v, fc, grid, planes);
  free(fc);
}

void oracle_transform_forward(const oracle_member* m, const real* grid, real* sp, int planes) {
  M v = view(m);
  real* fc = (real*)malloc(sizeof(real) * (size_t)planes * v.NLAT * v.NFC);
  gp2fc(&v, grid, fc, planes);
  fc2sp(&v, fc, sp, planes, TB(v, legendre_qc));
  free(fc);
}

/* gridpointa: dynamics/index.ts:171-230 */
static void gridpointa(const M* v) {
  const int L = v->NLEV;
  const real *qi = TB(*v, legendre_qi), *qj = TB(*v, legendre_qj), *qc = TB(*v, legendre_qc);
  dv2uv(v, A(*v, sd), A(*v, sz), A(*v, fgu), A(*v, fgv), KK(*v, plavor));
  sp2fc(v, A(*v, sd), A(*v, fgd), L, qi);
  sp2fc(v, A(*v, st), A(*v, fgt), L, qi);
  sp2fc(v, A(*v, sz), A(*v, fgz), L, qi);
  sp2fc(v, A(*v, sq), A(*v, fgq), L, qi);
  sp2fc(v, A(*v, sp), A(*v, fgp), 1, qi);
  sp2fc(v, A(*v, sp), A(*v, fgpj), 1, qj);
  dlambda(v, A(*v, fgp), A(*v, fgpmt), 1);
  fc2gp(v, A(*v, fgu), A(*v, gu), L);
  fc2gp(v, A(*v, fgv), A(*v, gv), L);
  fc2gp(v, A(*v, fgt), A(*v, gt), L);
  fc2gp(v, A(*v, fgd), A(*v, gd), L);
  fc2gp(v, A(*v, fgz), A(*v, gz), L);
  fc2gp(v, A(*v, fgpj), A(*v, gpj), 1);
  fc2gp(v, A(*v, fgpmt), A(*v, gpmt), 1);
  fc2gp(v, A(*v, fgp), A(*v, gp), 1);
  fc2gp(v, A(*v, fgq), A(*v, gq), L);
  calcgp(v);
  gp2fc(v, A(*v, gtn), A(*v, fgtn), L);
  gp2fc(v, A(*v, gqn), A(*v, fgqn), L);
  gp2fc(v, A(*v, gut), A(*v, fgut), L);
  gp2fc(v, A(*v, gvt), A(*v, fgvt), L);
  gp2fc(v, A(*v, guz), A(*v, fguz), L);
  gp2fc(v, A(*v, gvz), A(*v, fgvz), L);
  gp2fc(v, A(*v, gke), A(*v, fgke), L);
  gp2fc(v, A(*v, gvpp), A(*v, fgvpp), 1);
  gp2fc(v, A(*v, guq), A(*v, fguq), L);
  gp2fc(v, A(*v, gvq), A(*v, fgvq), L);
  fc2sp(v, A(*v, fgvpp), A(*v, spt), 1, qc);
  uv2dv(v, A(*v, fgvz), A(*v, fguz), A(*v, fgke), A(*v, sdt), A(*v, szt));
  qtend(v, A(*v, fgut), A(*v, fgvt), A(*v, fgtn), A(*v, stt));
  qtend(v, A(*v, fguq), A(*v, fgvq), A(*v, fgqn), A(*v, sqt));
}

/* spectrald: dynamics/index.ts:239-296 */
static void spectrald(const M* v) {
  const int L = v->NLEV;
  const real *qi = TB(*v, legendre_qi), *qc = TB(*v, legendre_qc);
  spectrald_diffusion(v);
  dv2uv(v, A(*v, sdp), A(*v, szp), A(*v, fzu), A(*v, fzv), KK(*v, plavor));
  fc2gp(v, A(*v, fzu), A(*v, zu), L);
  fc2gp(v, A(*v, fzv), A(*v, zv), L);
  dv2uv(v, A(*v, zdsdp1), A(*v, zdszp1), A(*v, fzdu), A(*v, fzdv), 0.0);
  fc2gp(v, A(*v, fzdu), A(*v, zdu), L);
  fc2gp(v, A(*v, fzdv), A(*v, zdv), L);
  mkdheat(v, 1);
  gp2fc(v, A(*v, zdtdt), A(*v, fzdtdt), L);
  fc2sp(v, A(*v, fzdtdt), A(*v, zstt1), L, qc);
  dv2uv(v, A(*v, zdsdp2), A(*v, zdszp2), A(*v, fzdu), A(*v, fzdv), 0.0);
  fc2gp(v, A(*v, fzdu), A(*v, zdu), L);
  fc2gp(v, A(*v, fzdv), A(*v, zdv), L);
  mkdheat(v, 2);
  gp2fc(v, A(*v, zdekin), A(*v, fzdekin), L);
  fc2sp(v, A(*v, fzdekin), A(*v, zsdef), L, qc);
  for (long i = 0; i < (long)L * v->NRSP; i++) /* mean-mode.wgsl:4-13 */
    if (i % v->NRSP != 0) A(*v, zsdef)[i] = 0.0;
  sp2fc(v, A(*v, zsdef), A(*v, fzdekin), L, qi);
  fc2gp(v, A(*v, fzdekin), A(*v, zdekin), L);
  mkdheat(v, 3);
  gp2fc(v, A(*v, zdtdt), A(*v, fzdtdt), L);
  fc2sp(v, A(*v, fzdtdt), A(*v, zstt2), L, qc);
  spectrald_update(v);
}

void oracle_step(const oracle_member* m, int with_physics) {
  M v = view(m);
  const int L = v.NLEV;
  const real *qi = TB(v, legendre_qi), *qc = TB(v, legendre_qc);
  gridpointa(&v);
  spectrala(&v);
  /* gridpointd's frame in: dynamics/index.ts:299-320 */
  dv2uv(&v, A(v, sd), A(v, sz), A(v, fgu), A(v, fgv), KK(v, plavor));
  sp2fc(&v, A(v, sq), A(v, fgq), L, qi);
  sp2fc(&v, A(v, st), A(v, fgt), L, qi);
  sp2fc(&v, A(v, sp), A(v, fgp), 1, qi);
  fc2gp(&v, A(v, fgu), A(v, gu), L);
  fc2gp(&v, A(v, fgv), A(v, gv), L);
  fc2gp(&v, A(v, fgt), A(v, gt), L);
  fc2gp(&v, A(v, fgp), A(v, gp), 1);
  fc2gp(&v, A(v, fgq), A(v, gq), L);
  physics_in(&v);
  if (with_physics) oracle_physics(m);
  physics_out(&v);
  gp2fc(&v, A(v, gtdt), A(v, fgtdt), L);
  gp2fc(&v, A(v, gudt), A(v, fgudt), L);
  gp2fc(&v, A(v, gvdt), A(v, fgvdt), L);
  gp2fc(&v, A(v, gqdt), A(v, fgqdt), L);
  fc2sp(&v, A(v, fgtdt), A(v, stt), L, qc);
  fc2sp(&v, A(v, fgqdt), A(v, sqt), L, qc);
  uv2dv(&v, A(v, fgudt), A(v, fgvdt), NULL, A(v, sdt), A(v, szt));
  spectrald(&v);
}
