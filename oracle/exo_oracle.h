/*
 * FP64 CPU oracle of the ExoPlaSim model step: TEST INFRASTRUCTURE ONLY.
 * See exo_oracle.c. Used by the tests, the smoke check and the benchmark's
 * CPU baseline as the checker; never linked into the shipped library.
 */
#ifndef EXO_ORACLE_H
#define EXO_ORACLE_H

/* The oracle's arithmetic type: double for the FP64 oracle (liboracle.so);
 * liboracle32.so builds the same source with real = float and
 * single-precision constants, a sequential FP32 emulation of the
 * reference's step used to size parity tolerances. */
#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
typedef ORACLE_REAL real;

/* One member: dimensions, a real arena, real tables and the current
 * phase's constants. slot_offsets / table_offsets / constants are indexed
 * by the position of each name in oracle_slot_names() /
 * oracle_table_names() / oracle_constant_names(); the caller resolves the
 * names against its own layout. */
typedef struct {
  int ntru, nlat, nlon, nlev;
  real* arena;
  const real* tables;
  const real* constants;
  const long* slot_offsets;
  const long* table_offsets;
} oracle_member;

const char* const* oracle_slot_names(void);
const char* const* oracle_table_names(void);
const char* const* oracle_constant_names(void);

/* One model step in the reference's dispatch order; the parameterizations
 * run when with_physics is non-zero. */
void oracle_step(const oracle_member* m, int with_physics);
/* gridpointd's parameterizations alone (fixer .. carbon). */
void oracle_physics(const oracle_member* m);
/* One parameterization by name (see exo_oracle_physics.c); -1 if unknown. */
int oracle_routine(const oracle_member* m, const char* name);
/* The radiation's absorber fits at n paths, and the storm's potential
 * intensity parts of one column (see exo_oracle_physics.c). */
int oracle_radiation_fit(const oracle_member* m, int which, const real* u, int n, real* out);
int oracle_storm_potential(const oracle_member* m, int h, real* out);
/* The snapshot pack of one main-loop step (snapshot/wgsl/pack.wgsl:14-107):
 * surface-temperature extrema tracking and, when snapshot_step[2] is set,
 * the published planes. */
void oracle_pack(const oracle_member* m);
/* Set-up transforms over `planes` planes (qi or qj when dmu). */
void oracle_transform_inverse(const oracle_member* m, const real* sp, real* grid, int planes, int dmu);
void oracle_transform_forward(const oracle_member* m, const real* grid, real* sp, int planes);

#ifdef EXO_ORACLE_INTERNAL
#define SLOTS(X)                                                             \
  X(sd) X(st) X(sz) X(sq) X(sdp) X(stp) X(szp) X(sqp) X(sdt) X(stt) X(szt)   \
  X(sqt) X(sdm) X(stm) X(szm) X(sqm) X(zdsdp1) X(zdszp1) X(zdsdp2)           \
  X(zdszp2) X(zstt1) X(zstt2) X(zsdef) X(sp) X(spp) X(spt) X(spm)            \
  X(gd) X(gt) X(gz) X(gq) X(gu) X(gv) X(dw) X(gtn) X(gqn) X(guz) X(gvz)      \
  X(gke) X(gut) X(gvt) X(guq) X(gvq) X(zu) X(zv) X(zdu) X(zdv) X(zdtdt)      \
  X(zdekin) X(gudt) X(gvdt) X(gtdt) X(gqdt) X(gp) X(gpj) X(gpmt) X(gvpp)     \
  X(dp0) X(dp) X(du0) X(dv0) X(dq) X(du) X(dv) X(dt) X(dudt) X(dvdt)         \
  X(dtdt) X(dqdt) X(fgd) X(fgt) X(fgz) X(fgq) X(fgu) X(fgv) X(fgtn) X(fgqn)  \
  X(fguz) X(fgvz) X(fgke) X(fgut) X(fgvt) X(fguq) X(fgvq) X(fzu) X(fzv)      \
  X(fzdu) X(fzdv) X(fzdtdt) X(fzdekin) X(fgudt) X(fgvdt) X(fgtdt) X(fgqdt)   \
  X(fgp) X(fgpj) X(fgpmt) X(fgvpp)                                           \
  X(fixq) X(fixneg) X(fixpos) X(fixsum) X(dz0) X(dls) X(drhs) X(dwatc)       \
  X(dtsa) X(dtransh) X(dtransm) X(dtaux) X(dtauy) X(dust3) X(dshfl)         \
  X(dshdt) X(devap) X(dlhfl) X(dlhdt) X(dcc) X(dql) X(dfu) X(dfd) X(dftu)   \
  X(dftd) X(dswfl) X(dlwfl) X(dflux) X(dqo3) X(dqco2) X(dtdtswr)            \
  X(dtdtlwr) X(dconv) X(gmu0) X(dalb) X(dicec) X(dsalb) X(radiation_step)   \
  X(dqt) X(fdqt) X(icclev) X(dprcl) X(dprll) X(dprscl) X(dprsll)            \
  X(dca_dtdt) X(dca_dqdt) X(dprc) X(dprl) X(dprs) X(dqvi) X(icctot)         \
  X(dca_limited) X(dca_skip) X(dforest) X(dglac) X(dwmax) X(drunoff)        \
  X(dsnow) X(dsmelt) X(dsndch) X(dtsoil) X(dtd2) X(dtd3) X(dtd4) X(dtd5)    \
  X(diced) X(dmld) X(surface_step) X(land_dts) X(land_dtsm) X(land_dqs)     \
  X(driver) X(duroff) X(dvroff) X(darea) X(dsnowt) X(dsnowz) X(dwater)      \
  X(dz0climo) X(dalbclim) X(dalbclim1) X(dalbclim2) X(dsoilt) X(dalbcl)     \
  X(dalbcl1)                                                                 \
  X(dalbcl2) X(xts) X(xsst) X(xmld) X(xls) X(xiced) X(xicec) X(xsnow)       \
  X(xqmelt) X(xicecc) X(xcflux) X(xcfluxf) X(xoheat) X(xoflux) X(xgw)       \
  X(xclsst2) X(xclicec2) X(xcliced2) X(xclssto) X(xcfluxr) X(ximelt)        \
  X(xsmelt) X(xsmflx) X(xsndch) X(ice_zflx) X(ice_snowold) X(ice_part1)     \
  X(ice_part2) X(xclsst) X(xclicec) X(xcliced) X(ice_sums) X(ysst) X(ymld)  \
  X(yls) X(yfldo) X(yclsst2) X(yclsst) X(sea_dts) X(sea_dqs) X(sea_dsst)    \
  X(sea_cheata) X(sea_cprsa) X(sea_naccua) X(localprecip)                   \
  X(localavgprecip) X(localweathering) X(localavgweather) X(localavgtemps)  \
  X(co2) X(carbon_step) X(capen) X(lnb) X(mpoti) X(chim) X(laav) X(venti)    \
  X(vrmpi) X(gpi) X(shear) X(dpersist) X(dcveg) X(dcsoil) X(dlai) X(dvsoil) \
  X(dagg) X(dgrow) X(dsc) X(dmr) X(dnpp) X(snapshot) X(snapshot_tracking)    \
  X(snapshot_step)

#define TABLES(X)                                                            \
  X(fourier_forward) X(fourier_inverse) X(legendre_qc) X(legendre_qe)        \
  X(legendre_qm) X(legendre_qq) X(legendre_qi) X(legendre_qj)                \
  X(legendre_qu) X(legendre_qv) X(t0) X(dsigma) X(sigma) X(sigmah) X(rdsig)  \
  X(t01s2) X(tkp) X(damp) X(tdisst) X(tdissq) X(tdissd) X(tdissz) X(tfrc)    \
  X(c) X(g) X(tau) X(bm1) X(rcsq) X(nindex) X(sop) X(sakpp) X(srp)         \
  X(zgw) X(flux_sigfac) X(flux_lnsig) X(vdiff_skap) X(vdiff_skaph)          \
  X(vdiff_mixh2) X(vdiff_mixm2) X(vdiff_rdsig) X(vdiff_sigmah3)             \
  X(vdiff_kfac) X(sid) X(cola) X(rad_lnsig) X(lwr_zzf1) X(lwr_zzf3)         \
  X(rcrit) X(kuo_zfac) X(dsoilz)

#define CONSTS(X)                                                            \
  X(psurf) X(ww) X(cv) X(ct) X(akap) X(rdbrv) X(adv) X(acpd) X(ga)           \
  X(plavor) X(delt) X(delt2) X(deltsec2) X(pnu) X(pnu21) X(filter)          \
  X(gascon) X(deltsec) X(ra1) X(ra2) X(ra4) X(tmelt) X(alv) X(als)           \
  X(zumin) X(vdiff_b) X(vdiff_c) X(vdiff_d) X(ndheat) X(rad_synchronous)     \
  X(gsol0) X(zsolar1) X(zsolar2) X(zsolars1) X(zsolars2) X(tswr1) X(tswr2)   \
  X(tswr3) X(zb3) X(zb4) X(zb5) X(rcoeff) X(ln_diffuse_rayleigh)             \
  X(nstartemp) X(ozone) X(a0o3) X(a1o3) X(aco3) X(bo3) X(co3) X(zconst_o3)   \
  X(zdawn) X(sbk) X(th2oc) X(tpofmt) X(zao30) X(zco20) X(zh2o0a) X(zh2o0)    \
  X(zao3c) X(zaco2c) X(zah2oc) X(zth2oc) X(zfco2_zzf2) X(rhbeta) X(nbeta)    \
  X(pdeep) X(pdeepth) X(rkshallow) X(gamma) X(zrfac) X(clwfac) X(clwcrit1)   \
  X(clwcrit2) X(sicecap) X(soilcap) X(sicediff) X(soildiff) X(snowcap)       \
  X(snowdiff) X(rhosnow) X(dztop) X(dsmax) X(albsmax) X(albsmin) X(albsmaxf) \
  X(albsminf) X(albsmax1) X(albsmin1) X(albsmaxf1) X(albsminf1) X(albsmax2)  \
  X(albsmin2) X(albsmaxf2) X(albsminf2) X(albgmax) X(albgmin) X(albgmax1)    \
  X(albgmin1) X(albgmax2) X(albgmin2) X(drhsfull) X(nwetsoil) X(wsmax)       \
  X(albland) X(alblandmax) X(dgroundalb1) X(dgroundalb2) X(doceanalb1)       \
  X(doceanalb2) X(nveg) X(nglacier) X(glacelim) X(rlue) X(co2conv)           \
  X(tau_veg) X(tau_soil) X(ncveg) X(cveg_e) X(cveg_g) X(xdt) X(xmind) X(xmaxd) \
  X(cicemin) X(thicec) X(tfreeze) X(crhoi) X(crhos) X(crhosn) X(cpi)         \
  X(ckapi) X(ckapsn) X(clfi) X(clfsn) X(ice_coupled) X(dtmix) X(cps)         \
  X(charnock) X(dz0sea) X(dz0ice) X(drhssea) X(drhsice) X(albsea) X(albice)  \
  X(dicealbmx1) X(dicealbmx2) X(dicealbmn1) X(dicealbmn2) X(ncarbon)         \
  X(nsupply) X(timeweight) X(co2earth) X(cbeta) X(kact) X(pearth) X(wmax)   \
  X(vearth) X(plarad) X(rad_earthsq) X(nstormdiag) X(ckcd) X(vithresh)

enum {
#define E(n) S_##n,
  SLOTS(E)
#undef E
  S_COUNT
};
enum {
#define E(n) T_##n,
  TABLES(E)
#undef E
  T_COUNT
};
enum {
#define E(n) K_##n,
  CONSTS(E)
#undef E
  K_COUNT
};

static const char* const oracle_slot_names_[] = {
#define E(n) #n,
    SLOTS(E)
#undef E
        0};
static const char* const oracle_table_names_[] = {
#define E(n) #n,
    TABLES(E)
#undef E
        0};
static const char* const oracle_const_names_[] = {
#define E(n) #n,
    CONSTS(E)
#undef E
        0};


typedef struct {
  int T, NTP1, NLAT, NLON, NLEV, NHOR, NCSP, NRSP, NFC;
  real* a;
  const real* t;
  const real* k;
  const long* so;
  const long* to;
} M;

static inline M view(const oracle_member* m) {
  M v;
  v.T = m->ntru;
  v.NTP1 = m->ntru + 1;
  v.NLAT = m->nlat;
  v.NLON = m->nlon;
  v.NLEV = m->nlev;
  v.NHOR = m->nlat * m->nlon;
  v.NCSP = (m->ntru + 1) * (m->ntru + 2) / 2;
  v.NRSP = (m->ntru + 1) * (m->ntru + 2);
  v.NFC = 2 * (m->ntru + 1);
  v.a = m->arena;
  v.t = m->tables;
  v.k = m->constants;
  v.so = m->slot_offsets;
  v.to = m->table_offsets;
  return v;
}

#define A(v, s) ((v).a + (v).so[S_##s])
#define TB(v, s) ((v).t + (v).to[T_##s])
#define KK(v, s) ((v).k[K_##s])

static inline int planes_of(const M* v, int per_level) { return per_level ? v->NLEV : 1; }

#endif
#endif
