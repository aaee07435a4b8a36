/*
 * FP64 CPU oracle of the ExoPlaSim model step: TEST INFRASTRUCTURE ONLY.
 * See exo_oracle.c. Used by the tests, the smoke check and the benchmark's
 * CPU baseline as the checker; never linked into the shipped library.
 */
#ifndef EXO_ORACLE_H
#define EXO_ORACLE_H

/* One member: dimensions, a double arena, double tables and the current
 * phase's constants. slot_offsets / table_offsets / constants are indexed
 * by the position of each name in oracle_slot_names() /
 * oracle_table_names() / oracle_constant_names(); the caller resolves the
 * names against its own layout. */
typedef struct {
  int ntru, nlat, nlon, nlev;
  double* arena;
  const double* tables;
  const double* constants;
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
/* Set-up transforms over `planes` planes (qi or qj when dmu). */
void oracle_transform_inverse(const oracle_member* m, const double* sp, double* grid, int planes, int dmu);
void oracle_transform_forward(const oracle_member* m, const double* grid, double* sp, int planes);

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
  X(fgp) X(fgpj) X(fgpmt) X(fgvpp)

#define TABLES(X)                                                            \
  X(fourier_forward) X(fourier_inverse) X(legendre_qc) X(legendre_qe)        \
  X(legendre_qm) X(legendre_qq) X(legendre_qi) X(legendre_qj)                \
  X(legendre_qu) X(legendre_qv) X(t0) X(dsigma) X(sigma) X(sigmah) X(rdsig)  \
  X(t01s2) X(tkp) X(damp) X(tdisst) X(tdissq) X(tdissd) X(tdissz) X(tfrc)    \
  X(c) X(g) X(tau) X(bm1) X(rcsq) X(nindex) X(sop) X(sakpp) X(srp)

#define CONSTS(X)                                                            \
  X(psurf) X(ww) X(cv) X(ct) X(akap) X(rdbrv) X(adv) X(acpd) X(ga)           \
  X(plavor) X(delt) X(delt2) X(deltsec2) X(pnu) X(pnu21) X(filter)

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
  double* a;
  const double* t;
  const double* k;
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
