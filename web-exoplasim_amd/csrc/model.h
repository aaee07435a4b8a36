// Dimensions, table registry and per-step constants of one ExoPlaSim member
// (one planet). Shared by the host set-up, the engine and the device kernels.
#pragma once
#include "fields.h"

// Table registry. X(name, count) with counts in terms of the dimensions:
// T truncation, NLAT, NLON, L levels, NCSP complex modes, NRSP real
// coefficients, NFC Fourier floats per row, NHOR grid points.
// Names and contents follow STEPPER.tables (stepper/index.ts:52-90) and
// DYNAMICS.TABLES (dynamics/index.ts:70-103).
#define EXO_TABLES(X)                                                          \
  X(fourier_forward, NLON * NFC) X(fourier_inverse, NFC * NLON)                \
  X(legendre_qc, NLAT * NCSP) X(legendre_qe, NLAT * NCSP)                      \
  X(legendre_qm, NLAT * NCSP) X(legendre_qq, NLAT * NCSP)                      \
  X(legendre_qi, NLAT * NCSP) X(legendre_qj, NLAT * NCSP)                      \
  X(legendre_qu, NLAT * NCSP) X(legendre_qv, NLAT * NCSP)                      \
  X(zgw, NHOR)                                                                 \
  X(flux_sigfac, 1) X(flux_lnsig, 1) X(vdiff_skap, L) X(vdiff_skaph, L)        \
  X(vdiff_mixh2, L) X(vdiff_mixm2, L) X(vdiff_rdsig, L) X(vdiff_sigmah3, L)    \
  X(vdiff_kfac, L)                                                             \
  X(sid, NLAT) X(cola, NLAT) X(rad_lnsig, L) X(lwr_zzf1, L) X(lwr_zzf3, L)     \
  X(rcrit, L) X(kuo_zfac, L)                                                   \
  X(dsoilz, 5)                                                                 \
  X(t0, L) X(dsigma, L) X(sigma, L) X(sigmah, L) X(rdsig, L) X(t01s2, L)       \
  X(tkp, L) X(damp, L) X(tdisst, L) X(tdissq, L) X(tdissd, L) X(tdissz, L)     \
  X(tfrc, L) X(c, L * L) X(g, L * L) X(tau, L * L) X(bm1, L * L * T)           \
  X(rcsq, NHOR) X(nindex, NRSP) X(sop, NRSP) X(sakpp, NRSP * L)                \
  X(srp, NRSP * L)

enum ExoTable {
#define EXO_TABLE_ENUM(name, count) TAB_##name,
  EXO_TABLES(EXO_TABLE_ENUM)
#undef EXO_TABLE_ENUM
  TAB_COUNT
};

// Scalar constants of one start-up phase or of the main loop: the
// pipeline override constants the reference bakes into its shaders
// (setup/index.ts:55-379), here data so that members with different
// planets share one kernel.
#define EXO_CONSTANTS(X)                                                       \
  /* pumamod */                                                                \
  X(psurf) X(ww) X(cv) X(ct) X(akap) X(rdbrv) X(adv) X(acpd) X(ga)             \
  X(gascon) X(plavor) X(delt) X(delt2) X(deltsec) X(deltsec2) X(pnu)           \
  X(pnu21) X(filter) X(ra1) X(ra2) X(ra4) X(tmelt) X(alv) X(als)               \
  X(solar_day)                                                                 \
  /* fluxmod */                                                                \
  X(zumin) X(vdiff_b) X(vdiff_c) X(vdiff_d) X(ndheat)                          \
  /* radmod */                                                                 \
  X(rad_synchronous) X(gsol0) X(zsolar1) X(zsolar2) X(zsolars1) X(zsolars2)    \
  X(tswr1) X(tswr2) X(tswr3) X(zb3) X(zb4) X(zb5) X(rcoeff)                    \
  X(ln_diffuse_rayleigh) X(nstartemp) X(ozone) X(a0o3) X(a1o3) X(aco3)         \
  X(bo3) X(co3) X(zconst_o3) X(zdawn) X(sbk) X(th2oc) X(tpofmt)                \
  X(zao30) X(zco20) X(zh2o0a) X(zh2o0) X(zao3c) X(zaco2c) X(zah2oc)            \
  X(zth2oc) X(zfco2_zzf2)                                                      \
  /* rainmod */                                                                \
  X(rhbeta) X(nbeta) X(pdeep) X(pdeepth) X(rkshallow) X(gamma) X(zrfac)        \
  X(clwfac) X(clwcrit1) X(clwcrit2)                                            \
  /* landmod soil */                                                           \
  X(sicecap) X(soilcap) X(sicediff) X(soildiff) X(snowcap) X(snowdiff)         \
  X(rhosnow) X(dztop) X(dsmax)                                                 \
  /* landmod surface */                                                        \
  X(albsmax) X(albsmin) X(albsmaxf) X(albsminf) X(albsmax1) X(albsmin1)        \
  X(albsmaxf1) X(albsminf1) X(albsmax2) X(albsmin2) X(albsmaxf2)               \
  X(albsminf2) X(albgmax) X(albgmin) X(albgmax1) X(albgmin1) X(albgmax2)       \
  X(albgmin2) X(drhsfull) X(nwetsoil) X(wsmax) X(albland) X(alblandmax)        \
  X(dgroundalb1) X(dgroundalb2) X(doceanalb1) X(doceanalb2)                    \
  /* vegetation and glacier */                                                 \
  X(nveg) X(nglacier) X(glacelim) X(rlue) X(co2conv) X(tau_veg) X(tau_soil)   \
  X(ncveg) X(cveg_e) X(cveg_g)                                                 \
  /* icemod */                                                                 \
  X(xdt) X(xmind) X(xmaxd) X(cicemin) X(thicec) X(tfreeze) X(crhoi)            \
  X(crhos) X(crhosn) X(cpi) X(ckapi) X(ckapsn) X(clfi) X(clfsn)                \
  X(ice_coupled)                                                               \
  /* oceanmod */                                                               \
  X(dtmix) X(cps)                                                              \
  /* seamod */                                                                 \
  X(charnock) X(dz0sea) X(dz0ice) X(drhssea) X(drhsice) X(albsea) X(albice)    \
  X(dicealbmx1) X(dicealbmx2) X(dicealbmn1) X(dicealbmn2)                      \
  /* carbonmod */                                                              \
  X(ncarbon) X(nsupply) X(timeweight) X(co2earth) X(cbeta) X(kact)             \
  X(pearth) X(wmax) X(vearth) X(plarad) X(rad_earthsq)                         \
  /* hurricanemod */                                                           \
  X(nstormdiag) X(ckcd) X(vithresh)                                            \
  /* snapshot */                                                               \
  X(jet_level) X(sigma_bottom)

enum ExoConst {
#define EXO_CONST_ENUM(name) CON_##name,
  EXO_CONSTANTS(EXO_CONST_ENUM)
#undef EXO_CONST_ENUM
  CON_COUNT
};

struct ExoDims {
  int ntru, nlat, nlon, nlev;
};

static inline int exo_ncsp(int t) { return (t + 1) * (t + 2) / 2; }
static inline int exo_nrsp(int t) { return (t + 1) * (t + 2); }
static inline int exo_nfc(int t) { return 2 * (t + 1); }

// Per-step values the host computes from the calendar (stepper/index.ts:
// 281-334): 5 radiation, 7 surface, 1 carbon, 3 snapshot floats.
enum { EXO_STEP_VALUES = 16 };
