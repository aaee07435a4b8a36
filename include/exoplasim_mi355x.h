/*
 * exoplasim_mi355x: C-ABI of the MI355X-native ExoPlaSim time step.
 *
 * One handle is one engine on one GPU holding one or more members
 * (planets). Every member's prognostic state, tables and scratch live in
 * HBM; the host passes plain pointers and sizes only.
 *
 * The engine-internal seam this library replaces is the reference WebGPU
 * worker's use of SETUP / STEPPER / LAYOUT / RESTART
 * (src/model/climate/gcm/runner/webgpu/gcm.worker.ts:53-388): each entry
 * point below cites the reference call it stands in for.
 *
 * Errors: every call returns 0 on success and a negative EXO_E_* code on
 * failure; exo_last_error() returns the message of the last failure on the
 * calling thread, worded like the reference's worker errors
 * (gcm.worker.ts:47-51, :317-323; checkpoint/restart/index.ts:50-66).
 * Calls on one handle are not re-entrant; different handles are
 * independent.
 */
#ifndef EXOPLASIM_MI355X_H
#define EXOPLASIM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXO_OK 0
#define EXO_E_ARG -1      /* invalid argument or configuration */
#define EXO_E_DEVICE -2   /* HIP runtime failure */
#define EXO_E_STATE -3    /* call out of order (e.g. step before init) */
#define EXO_E_UNSTABLE -4 /* non-finite value in the published state */
#define EXO_E_FORMAT -5   /* restart bytes do not match this layout */

/* Precision of the Legendre contractions. */
#define EXO_PRECISION_FP32 0       /* plain form: FP32 throughout */
#define EXO_PRECISION_SPLIT 1 /* Legendre on FP16 hi/lo split operands, FP32 acc. */

/*
 * Planet and run configuration: the subset of the reference's GcmConfig
 * (src/model/climate/gcm/types.ts:28-62) that the synthetic workloads
 * vary, plus the seeded terrain and initial-state generator.
 */
typedef struct exo_config {
  int truncation;              /* 21, 42 or 63 */
  int levels;                  /* 5 or 10 */
  double timestep_minutes;     /* GcmConfig.timestepMinutes */
  double radius_earth;         /* GcmConfig.radiusEarth (x 6371220 m) */
  double gravity;              /* GcmConfig.gravityMs2 */
  double rotation_period_days; /* GcmConfig.rotationPeriodDays */
  double pressure_bar;         /* surface pressure (atmosphere resolve) */
  double gas_constant;         /* J/(kg K) (atmosphere resolve) */
  double co2_ppm;              /* atmosphere resolve */
  double flux_wm2;             /* GcmConfig.fluxWm2 */
  double obliquity_deg;        /* GcmConfig.obliquityDeg */
  double eccentricity;         /* GcmConfig.eccentricity */
  double year_days;            /* GcmConfig.yearDays */
  double ocean_depth_m;        /* GcmConfig.oceanDepthM */
  int ozone;                   /* GcmConfig.atmosphere.ozone */
  /* Seeded synthetic planet (no reference counterpart: the reference
   * reads terrain rasters and .sra climatologies). */
  double land_fraction;        /* fraction of grid points that are land */
  double terrain_max_m;        /* highest land, metres */
  double initial_temperature;  /* isothermal atmosphere, K */
  double initial_humidity;     /* kg/kg */
  double noise_kelvin;         /* std of grid-point temperature noise, K */
  double sea_surface_k;        /* uniform SST climatology (code 169), K; NaN: none, and the
                                  ocean starts as the reference's does without one, at the
                                  freezing point under sea ice */
  uint64_t terrain_seed;
  uint64_t noise_seed;
  /* GcmConfig.stormclim (types.ts:30): 1 runs hurricanemod's storm
   * diagnostics every step (hurricane_namelist NSTORMDIAG, namelist/
   * index.ts:175; physics/storm/index.ts:22-45) into the snapshot's
   * storm fields; 0 (default) leaves them at zero. */
  int storm_diagnostics;
  /* The rest of GcmConfig (types.ts:28-62), as GCM_NAMELIST.edits writes
   * them into the namelists (namelist/index.ts:40-182). A real the
   * reference allows to be null takes NaN for null. */
  int vegetation;                  /* vegetation 0, 1 or 2: NVEG */
  double initial_growth;           /* initialGrowth: RINIDAGG */
  int glaciers;                    /* glaciers.enabled: NGLACIER */
  double glacier_min_depth_m;      /* glaciers.minDepthM: GLACELIM */
  double glacier_initial_height_m; /* glaciers.initialHeightM: ICESHEETH (< 0: none) */
  double star_temperature_k;       /* starTemperatureK: > 0 sets NSTARTEMP 1, STARBBTEMP */
  int synchronous;                 /* synchronous: NFIXED */
  double substellar_longitude_deg; /* substellarLongitudeDeg: FIXEDLON */
  int fixed_orbit;                 /* fixedOrbit: NFIXORB (0, Berger's orbit, is not built) */
  int keplerian;                   /* keplerian: NGENKEPLERIAN */
  double periapse_longitude_deg;   /* periapseLongitudeDeg: MVELP (NaN: null) */
  double initial_mean_anomaly_deg; /* initialMeanAnomalyDeg: MEANANOM0 (NaN: null) */
  double soil_depth_scale;         /* soilDepthScale: DSOILZ scale */
  double soil_water_capacity_m;    /* soilWaterCapacityM: WSMAX (NaN: null) */
  double soil_heat_capacity;       /* soilHeatCapacity: SOILCAP (NaN: null) */
  double land_albedo;              /* landAlbedo: ALBLAND, DGROUNDALB (NaN: null) */
  double ocean_albedo;             /* oceanAlbedo: ALBSEA, DOCEANALB (NaN: null) */
  double snow_ice_albedo;          /* snowIceAlbedo: ALBICE, DICEALBMN/MX, DSNOWALBMN, DGLACALBMN,
                                      DSNOWALB (NaN: null) */
  int wet_soil;                    /* wetSoil: NWETSOIL */
  int filter;                      /* physicsFilter's NFILTER: 0 none, 1 cesaro, 2 exp, 3 lh
                                      (NGPTFILTER and NSPVFILTER stay 1, as in the base
                                      namelist) */
  /* Fresh-start state (prolog's fresh start, setup/index.ts:434-449).
   * 0 (default): the seeded synthetic state above: an isothermal
   * atmosphere at initial_temperature with noise_kelvin grid-point noise,
   * initial_humidity, ln(ps) from the orography (no reference counterpart).
   * 1: the reference's initfd (setup/state/index.ts:60-86): temperature at
   * setzt's restoration profile (its global mean; the whole field under a
   * tidally locked gradient), solid-body planetary vorticity, no
   * divergence or humidity, ln(ps) from noise(KICK) drawn from gfortran's
   * random_number after random_seed(put=SEED), or from the orography when
   * nspinit > 0. */
  int initial_state;
  int kick;          /* plasim_nl KICK: 0 none, 1 every non-zonal coefficient,
                        2 the equatorially symmetric ones, 3 mode (1,1) alone */
  int32_t seed[8];   /* plasim_nl SEED (the base namelist's 14963991,2,...,8);
                        seed[0] == 0 takes the words from clock_ms (initrandom) */
  int64_t clock_ms;  /* the clock for a seed from the clock (Date.now()) */
  int nspinit;       /* surfmod_nl NSPINIT (default 0) */
} exo_config;

/* Fills *cfg with the defaults of a T21 L10 Earth-like synthetic planet:
 * 1 bar N2 with 360 ppm CO2, 1367 W/m^2 from a 5780 K star, obliquity
 * 23.4 deg, a circular orbit, and every other GcmConfig field at
 * GCM.DEFAULT_CONFIG's value (index.ts:4-44). */
void exo_config_default(exo_config *cfg);

/* Last error message of the calling thread ("" if none). */
const char *exo_last_error(void);

typedef struct exo_engine exo_engine;

/*
 * Creates an engine on HIP device `device` with room for `max_members`
 * members. Stands in for GCM_DEVICE.request + KERNEL.context
 * (device/index.ts:6-35, kernel/index.ts:17-87).
 */
int exo_create(int device, int max_members, exo_engine **out);
void exo_destroy(exo_engine *e);

/* Evaluates the column physics' transcendental routines on the device
 * over n host values (a, and b for the powers), writing n results: the
 * counterparts of the reference's accurate f32 math (GPU_MATH,
 * shared/gpu/math/wgsl/precise.wgsl:1-103), for callers that check their
 * accuracy. function: 0 log2 and 1 pow(a, b) of positive normal a on the
 * hardware's v_log_f32/v_exp_f32 (the longwave and shortwave absorber
 * paths, the surface layer), 2 log10 and 3 ln likewise, 4 e^x for x <= 0
 * (exp2 of x log2 e), 5 expf, 6 logf and 7 powf of the device library. */
int exo_math_eval(exo_engine *e, int function, const float *a, const float *b, float *out, size_t n);

/*
 * Adds a member configured from *cfg (FP64 set-up on the host: pumamod,
 * Gaussian grid, Legendre and Fourier tables, initpm/initsi/makebm/setzt,
 * physics tables, seeded initial state) and uploads it; the member then
 * runs the reference's nkits start-up steps on its first exo_step calls.
 * Stands in for SETUP.state + SETUP.schedule + STEPPER.context
 * (setup/index.ts:434-477, stepper/index.ts:94-119).
 * `precision` is EXO_PRECISION_FP32 or EXO_PRECISION_SPLIT.
 * Writes the member id to *member.
 */
int exo_add_member(exo_engine *e, const exo_config *cfg, int precision,
                   int *member);

/* exo_add_member with the planet's land and orography given instead of the
 * seeded terrain: land_fraction and geopotential (m^2/s^2) on the member's
 * nlat x 2 nlat grid (exo_boundary_from_terrain's output, or the values
 * of .sra files 172 and 129 read by exo_boundary_read_sra). As landini
 * (setup/surface/index.ts:293-297), land is where the fraction exceeds
 * 0.5. */
int exo_add_member_boundary(exo_engine *e, const exo_config *cfg,
                            int precision, const double *land_fraction,
                            const double *geopotential, int *member);

/* One surface field of a planet's .sra files (surface_ini's codes,
 * setup/surface/index.ts:49-70: 172 land-sea mask, 129 orography
 * (geopotential, m^2/s^2), 169 surface temperature, 173 roughness, 1730
 * orographic roughness, 174/175/176 albedos, 209 soil temperature, 210 ice
 * cover, 211 ice thickness, 212 forest, 229 field capacity, 232 glacier,
 * 1604-1607 vegetation): `months` records (1, 12 or 14) of nlat x 2 nlat
 * values each, north first. */
typedef struct {
  int code;
  int months;
  const double *values;
} exo_surface_field;

/* exo_add_member with the planet's surface from .sra fields, as the
 * reference's surface set-up reads them (SURFACE_SETUP.initial and its
 * read, setup/surface/index.ts:76-130, :267-643): each code absent takes
 * the reference's fallback; a 12-month climatology gains December before
 * and January after, a single record fills all 14 months, any other count
 * fails with the reference's "N032_surf_0169.sra has 3 fields, expected
 * 14". Without both 172 and 129 the seeded terrain of *cfg gives the land
 * and orography; cfg->sea_surface_k stands for a missing 169. */
int exo_add_member_surface(exo_engine *e, const exo_config *cfg, int precision,
                           const exo_surface_field *fields, int count, int *member);

/* GCM_ATMOSPHERE.resolve (atmosphere/index.ts:84-160): surface pressure
 * (bar), gas constant (J/(kg K)) and CO2 (ppm) from partial pressures
 * (bar; NaN: the reference's null) of H2, He, N2, O2, Ar, Ne, Kr, CO2 and
 * H2O, in that order, in the mode "partialPressures" (pressure_bar and
 * gas_constant NaN) or "pressureAndGasConstant" (either may be NaN for
 * the native value). EXO_E_ARG with the reference's message when the
 * atmosphere is invalid. */
int exo_atmosphere_resolve(const double *partial_pressures_bar, int pressure_and_gas_constant,
                           double pressure_bar, double gas_constant, double *out_pressure_bar,
                           double *out_gas_constant, double *out_co2_ppm);

/* GCM_NAMELIST.read (namelist/index.ts:196-233): reads `nkeys` keys from
 * the text of one namelist file. `values` holds, key after key, each key's
 * `lengths[k]` default numbers on entry and the read values on return: a
 * "KEY = v1,v2,..." line (key case ignored, the file's last line for a key
 * winning) writes its comma-separated numbers over the defaults from the
 * front, each as JavaScript's Number() reads it with a 'd' exponent taken
 * for 'e' (an empty part reads 0, a malformed one NaN); numbers beyond a
 * key's length are dropped. A NULL text reads as an empty file. */
int exo_namelist_read(const char *text, const char *const *keys, const int *lengths, int nkeys,
                      double *values);

/* The set-up's reading of a run's namelist files (setup/pumamod/
 * index.ts:103-232, setup/modules/index.ts:41-305, setup/index.ts:
 * 107-112), as the inverse of GCM_NAMELIST.edits (namelist/index.ts:
 * 40-182): `count` files by name ("planet_namelist", "plasim_namelist",
 * "radmod_namelist", ...; other names are ignored, an absent file reads
 * as empty) and text. Every exo_config field that edits writes is set from
 * the files, a key absent from them taking the module's default; the
 * fields no namelist carries (the synthetic planet's, initial_state,
 * clock_ms) keep the values *cfg holds on entry. A key the set-up reads
 * that no field carries must hold the value this engine runs with (the
 * base namelists' or the module's); any other value fails with EXO_E_ARG
 * and a message naming the file and key. */
int exo_config_from_namelists(int truncation, int levels, const char *const *names,
                              const char *const *texts, int count, exo_config *cfg);

/* Removes a member and frees its device memory. */
int exo_remove_member(exo_engine *e, int member);

/*
 * Advances every active member by `steps` model steps in one enqueue
 * (start-up phases included); returns when the work is queued. Stands in
 * for STEPPER.values + STEPPER.batch + queue.submit (stepper/index.ts:
 * 203-334).
 */
int exo_step(exo_engine *e, int steps);

/* exo_step for callers that cannot allocate (code that must not call
 * hipMalloc: tight serving loops, harnesses that forbid allocation): no
 * counterpart in the reference. It enqueues the same kernels with the same
 * results, launched one by one (no step graph, whose instantiation
 * allocates), and keeps the steps' calendar rows in a static device array
 * of EXO_NOALLOC_ROWS member-steps (1 MiB per device, held until the
 * process ends) staged through a pinned block exo_create allocated; it
 * calls no hipMalloc, hipFree or other allocating routine. active members
 * x steps above EXO_NOALLOC_ROWS fails with EXO_E_ARG before any work.
 * Every engine on one device shares the array: calls that overlap on one
 * device from different host threads race. */
#define EXO_NOALLOC_ROWS 16384
int exo_step_noalloc(exo_engine *e, int steps);

/* Publish a snapshot every `every` main-loop steps (0: never), as
 * GcmRunRequest.publishEvery (runner/types.ts:47-67). */
int exo_set_publish_every(exo_engine *e, int every);

/* Tile shape of the grouped transforms: how many members of one
 * (truncation, levels) class share a workgroup of the Legendre and zonal
 * transforms, 1..8 (default 4). Members compute bit-identically for every
 * size. The reference compiles one pipeline per planet and has no such
 * choice (kernel/index.ts:17-122); these calls and exo_autotune serve the
 * batched engine. */
int exo_set_transform_group(exo_engine *e, int truncation, int levels,
                            int members);
int exo_transform_group(exo_engine *e, int truncation, int levels,
                        int *members);

/* Picks each class's transform group size for the active batch by timing
 * `trial_steps` steps per candidate (coordinate descent over the classes
 * that hold two or more members), then restores every member's state,
 * step counters and phase as they were: the next exo_step continues
 * exactly as without the call. Nothing is published during the trials.
 * It then times the other lane count (see exo_set_step_lanes) and the
 * other launch mode (see exo_set_step_graphs), keeping the faster of
 * each; a later exo_set_step_lanes or exo_set_step_graphs overrides that
 * choice.
 * tuned_ms (may be null) receives the best trial's device ms per step. */
int exo_autotune(exo_engine *e, int trial_steps, double *tuned_ms);

/* Replay each step of exo_step from a captured HIP graph (enable != 0,
 * the default) or launch its kernels one by one (0). The graph is captured
 * when two consecutive steps of a call have the same active members and
 * start-up phases, and is kept while they last; results are bit-identical
 * either way. The reference submits one WebGPU command buffer per batch of
 * steps (stepper/index.ts:175-276); this is that batching's launch side. */
int exo_set_step_graphs(exo_engine *e, int enable);
int exo_step_graphs(exo_engine *e, int *enabled);

/* Lanes of a step, 1 to 4: n > 1 (2 by default) splits the active
 * members into n sets of about equal work (grid points x levels) whose
 * launches run on n streams, forked from and joined back to the engine's
 * stream (inside the step graph when one is replayed), so that one set's
 * latency-bound launches overlap the others'; 1 runs them all on one
 * stream. Results are bit-identical for every n (each member's launches
 * keep their order and arithmetic); exo_autotune keeps the fastest count.
 * While exo_probe_kernel times a kernel, steps run on one stream. No
 * reference counterpart: the reference's one command buffer per batch has
 * no concurrent branches. */
int exo_set_step_lanes(exo_engine *e, int lanes);
int exo_step_lanes(exo_engine *e, int *lanes);

/* Blocks until queued steps have finished. Stands in for mapAsync. */
int exo_sync(exo_engine *e);

/* Model step counter of a member (nstep, counted from the start step). */
int exo_member_step(exo_engine *e, int member, int64_t *step);
/* Calendar date {year, month, day, hour, minute} of the member's step
   counter, as GcmSnapshot.date (CALENDAR date, calendar/date/index.ts:
   46-82). */
int exo_member_date(exo_engine *e, int member, int *date5);
/* The same date at any model step of the member's calendar (GCM_DATE.date,
   calendar/date/index.ts:46-82, as the worker dates its progress reports,
   gcm.worker.ts:307-315): the 360-day calendar, or the Gregorian one for a
   365-day year. */
int exo_member_date_at(exo_engine *e, int member, int64_t step, int *date5);

/* Arena layout: total floats and one slot's offset/length by name.
 * LAYOUT.create / LAYOUT.constants (layout/index.ts:22-53). */
int exo_arena_floats(exo_engine *e, int member, size_t *floats);
int exo_slot(exo_engine *e, int member, const char *name, size_t *offset,
             size_t *floats);
/* Name of slot `index` in arena order, or NULL past the end. */
const char *exo_slot_name(int index);

/* Field readback / write (LAYOUT.read / LAYOUT.write, layout/index.ts:
 * 66-108). `floats` must equal the slot's length. Fourier slots read and
 * write as the reference's rows (row r's 2(T+1) coefficients together),
 * whatever their arrangement on the device; so does a checkpoint.
 * Rows that are scratch inside a step can hold other intermediates at a
 * step's end than the reference's: the dissipation tail borrows the
 * physics frame's gu, gv, fgu, fgv, gtn and fgtn, and fzdekin is not
 * formed. No step reads them before writing them. */
int exo_read_field(exo_engine *e, int member, const char *name, float *out,
                   size_t floats);
int exo_write_field(exo_engine *e, int member, const char *name,
                    const float *values, size_t floats);

/* FP64 set-up tables and constants of a member, as computed on the host
 * (for checkers and diagnostics). */
int exo_read_table(exo_engine *e, int member, const char *name, double *out,
                   size_t count);
int exo_table_count(exo_engine *e, int member, const char *name,
                    size_t *count);
int exo_read_constants(exo_engine *e, int member, int phase, double *out,
                       size_t count);
const char *exo_constant_name(int index);
/* Number of constant phases (start-up steps + main loop) and of start-up
 * steps (SETUP.schedule, setup/index.ts:455-477). */
int exo_phase_count(exo_engine *e, int member, int *phases, int *startup);
int exo_dims(exo_engine *e, int member, int *dims4);

/* Per-step calendar values (16 doubles) for model step `step`
 * (STEPPER.values, stepper/index.ts:281-334). */
int exo_step_values(exo_engine *e, int member, int64_t step, double *out16);

/*
 * Snapshot readback: packs the reference's published field set (101
 * fields at L10, 76 at L5, in fields/index.ts:81-132 order) and copies it
 * to `out` (fields x NHOR floats). Fails with EXO_E_UNSTABLE and the
 * reference's "model became unstable at step N" message when any value
 * is not finite (gcm.worker.ts:317-323).
 */
int exo_read_snapshot(exo_engine *e, int member, float *out, size_t floats);

/* Id of snapshot plane `index` for a member of `levels` levels, as
 * GCM_FIELDS orders them (fields/index.ts:81-128, fieldCount :416-418);
 * NULL past the last plane. The string stays valid until the calling
 * thread's next call. */
const char *exo_snapshot_field_name(int levels, int index);

/* Area-weighted global means (GCM_GRID.globalMean, grid/index.ts:99-111)
 * of surface temperature (K) and TOA net flux (W/m^2). */
int exo_global_means(exo_engine *e, int member, double *tsurf,
                     double *toa_net);

/*
 * Checkpoint in the reference's GPU restart layout
 * (checkpoint/restart/index.ts:11-79): "GCM-WEBGPU-STATE\n" magic, a JSON
 * index of [name, floats] pairs in arena offset order, then the f32
 * payload. exo_checkpoint_size gives the byte count; exo_checkpoint_save
 * writes it; exo_checkpoint_restore rejects a mismatched layout with the
 * reference's decode errors.
 */
int exo_checkpoint_size(exo_engine *e, int member, size_t *bytes);
int exo_checkpoint_save(exo_engine *e, int member, void *out, size_t bytes);
int exo_checkpoint_restore(exo_engine *e, int member, const void *bytes,
                           size_t size);
/* A restart under a changed configuration: exo_checkpoint_restore, then
 * the fields that GCM_CONTINUATION.fields (continuation/index.ts:60-85;
 * gcm.worker.ts:117-124) resets take the member's own (new) initial
 * values: CO2 (co2, dqco2) when co2_ppm differs from *previous, the mixed
 * layer (ymld, xmld, dmld) when ocean_depth_m does. previous = NULL is a
 * plain restore. The reference's other resets (land albedo, star
 * temperature, soil water capacity) have no exo_config counterpart. */
int exo_checkpoint_continue(exo_engine *e, int member, const void *bytes,
                            size_t size, const exo_config *previous);

/*
 * Serving loop (continuous batching). Runs arrive over time
 * (arrival_s, seconds after the call starts); host threads run each
 * arrived run's FP64 set-up; runs are admitted first-fit in arrival order
 * while a descriptor is free and the resident runs' device memory stays
 * within hbm_budget_bytes (0: 90% of the device's free memory); all
 * resident runs advance together, each retiring at exactly its `steps`
 * (an advance lasts as long as the shortest remaining run, at most
 * max_chunk_steps), and runs that arrive meanwhile join at the next
 * advance. Each result holds the run's times and its global means at the
 * end. The engine must have no members; it has none again on return. A
 * run computes the same alone and in the mix. The reference serves one run
 * per page (runner/index.ts:61-78); this batched loop has no counterpart
 * there.
 */
typedef struct {
  exo_config config;
  int precision;
  int reserved;
  double arrival_s;
  int64_t steps; /* model steps, start-up steps included */
} exo_request;

typedef struct {
  double arrival_s, admitted_s, finished_s; /* seconds after the call starts */
  double tsurf, toa_net;                    /* global means at the end */
  int64_t steps;                            /* steps run */
  int status;                               /* EXO_OK or the failure code */
  int reserved;
} exo_result;

/* Progress report of the serving loop: called on the calling thread after
   each batched advance that ended at least one run, with the runs ended so
   far (finished or failed) out of `count`. */
typedef void (*exo_serve_progress)(void *user, int finished, int count, double elapsed_s);

typedef struct {
  size_t hbm_budget_bytes;
  int max_chunk_steps; /* 0: 32 */
  int prepare_threads; /* 0: 4 */
  exo_serve_progress progress; /* may be NULL */
  void *progress_user;
} exo_serve_options;

int exo_serve(exo_engine *e, const exo_request *requests, int count,
              const exo_serve_options *options, exo_result *results,
              double *wall_s);

/* Device-side timing of the last exo_step call (ms, HIP events). Replaces
   the timestamp queries of STEPPER.batch (src/model/climate/gcm/stepper/
   index.ts:175-276). */
int exo_last_step_ms(exo_engine *e, double *ms);

/* Kernel-duration probe: from this call on, every launch of `kernel` (a
   kernel name such as "k_leg_inv") is bracketed by a pair of HIP events
   on the engine's stream, for up to max_launches launches (1..65536;
   later launches are counted in `missed`). kernel = NULL stops it. The
   reference has no per-kernel timing; this serves the benchmark's
   roofline figures. */
int exo_probe_kernel(exo_engine *e, const char *kernel, int max_launches);
/* Summed duration (ms) and count of the probed launches so far; waits for
   the last one. */
int exo_probe_result(exo_engine *e, double *total_ms, int *launches,
                     int *missed);

/* ---- Averages and balance ------------------------------------------
   Monthly means of a member's published snapshots, kept on the device:
   GCM_AVERAGES.add / close / annual (src/model/climate/gcm/averages/
   index.ts:26-137) with the month history of HISTORY_MONTHS (:21-23) and
   the yearly global balances (yearly, :31-48), and ExoPlaSim's drift test
   GCM_BALANCE.assess (src/model/climate/gcm/balance/index.ts:6-68). The
   reference runs them on the host over read-back snapshots; here each
   published snapshot is added into an FP64 month sum on the device
   (exo_step), and only what is asked for is read back. */
#define EXO_MONTHS_PER_YEAR 12
#define EXO_HISTORY_MONTHS 132

typedef struct {
  int year, month; /* calendar month the mean covers */
  int samples;     /* snapshots averaged */
  int64_t step;    /* step of its last snapshot */
} exo_month_info;

typedef struct {
  int year;
  double toa_net_radiation;   /* W/m^2, global mean of the annual mean */
  double surface_net_heat;    /* W/m^2 */
  double surface_temperature; /* K */
} exo_yearly_balance;

#define EXO_BALANCE_SPINNING_UP 0
#define EXO_BALANCE_DRIFTING 1
#define EXO_BALANCE_BALANCED 2

typedef struct {
  int status;       /* EXO_BALANCE_* */
  int years_needed; /* spinning up: years still missing */
  double toa_drift_per_year, surface_drift_per_year; /* W/m^2 per year; NaN while spinning up */
} exo_balance_assessment;

/* Starts (enable = 1) or stops averaging for a member: starting allocates
   its month sum and the month history on the device (HISTORY_MONTHS x
   the snapshot size) and clears them; the member's next published
   snapshots are averaged. Needs a publishing interval (exo_set_publish_
   every). */
int exo_set_averaging(exo_engine *e, int member, int enable);
/* Closed monthly means held (<= EXO_HISTORY_MONTHS). */
int exo_month_count(exo_engine *e, int member, int *count);
/* Closed month `index` (0 = oldest held): its calendar month and, when
   values is not NULL, its mean snapshot (`count` floats: fields x NHOR,
   exo_read_snapshot's layout). Waits for queued steps. */
int exo_read_month(exo_engine *e, int member, int index, exo_month_info *info,
                   float *values, size_t count);
/* GCM_AVERAGES.annual (averages/index.ts:103-137): the mean of the last 12
   closed months, with the diurnal temperature range as the mean of the
   months' daily ranges; EXO_E_STATE before 12 months have closed. */
int exo_read_annual(exo_engine *e, int member, float *values, size_t count);
/* The member's yearly balances so far (one per complete calendar year of
   closed months), oldest first; *count = how many there are (up to max
   written). */
int exo_yearly_balances(exo_engine *e, int member, exo_yearly_balance *out,
                        int max, int *count);
/* The month being summed (GcmAverager.current, averages/index.ts:71-79):
   its calendar month, samples and last step (samples = 0 when none is
   open) and, when sum is not NULL, its FP64 sum (`count` doubles). With
   the closed months and the yearly balances it is the averager a
   checkpoint carries (GcmRunCheckpoint.averager, runner/types.ts:28-32). */
int exo_read_accumulator(exo_engine *e, int member, exo_month_info *info,
                         double *sum, size_t count);
/* Replaces a member's averages with a saved averager (GcmRunRequest.
   averager, runner/types.ts:58-59; gcm.worker.ts:197): n_months closed
   months (oldest first, <= EXO_HISTORY_MONTHS, `floats` values each, laid
   out one after another), n_years yearly balances, and the open month
   (current = NULL or current->samples = 0: none open; else current_sum
   holds `floats` doubles). Starts averaging when it is off. */
int exo_load_averager(exo_engine *e, int member, const exo_month_info *months,
                      const float *month_values, int n_months,
                      const exo_yearly_balance *years, int n_years,
                      const exo_month_info *current, const double *current_sum,
                      size_t floats);
#define EXO_CLIMATE_YEARS 10
/* The climatology of GCM_AVERAGES.view (averages/index.ts:177-222): for
   calendar month `month` (1..12), the mean over the last EXO_CLIMATE_YEARS
   complete calendar years held (twelve distinct months of one year) of
   that month's mean, summed in FP64 oldest year first and rounded to FP32.
   *years = the years averaged (0: no complete year, nothing written);
   info = that month's entry in the latest of them. */
int exo_read_climatology(exo_engine *e, int member, int month,
                         exo_month_info *info, float *values, size_t count,
                         int *years);
/* GCM_BALANCE.assess over yearly balances, oldest first (host only). */
int exo_balance_assess(const exo_yearly_balance *years, int count,
                       exo_balance_assessment *out);

/* ---- Boundary data ---------------------------------------------------
   The input side of the step: a terrain raster onto a Gaussian grid and
   ExoPlaSim's .sra surface files. */

/* GCM_BOUNDARY.fromTerrain (src/model/climate/gcm/boundary/index.ts:
   12-41; GCM_TERRAIN_REGRID.cell, terrain/regrid/index.ts:6-46): each cell
   of the nlat x 2 nlat Gaussian grid (nlat 32, 64 or 96; rows north
   first, longitudes from 0 east) gets the area-weighted land fraction
   (clamped to [0, 1]) and g x the area-weighted mean of elevation x land
   of the width x height raster (rows from 90N, columns from 180W) under
   it, in FP64 on the engine's device. The raster is validated as
   GCM_TERRAIN.validate does (terrain/index.ts:11-44), with its messages. */
int exo_boundary_from_terrain(exo_engine *e, const float *elevation_m,
                              const float *land, int width, int height,
                              int nlat, double gravity_ms2,
                              double *land_fraction, double *geopotential);
/* toSra (boundary/index.ts:43-57): the .sra text of one field (code 172
   land-sea mask, 129 orography) of an nlat x 2 nlat grid: a header of
   eight integers, then eight values a line as JavaScript's toFixed(3).
   *length = the text's bytes; out = NULL asks for the length only. */
int exo_boundary_sra(int code, int nlat, const double *values, char *out,
                     size_t capacity, size_t *length);

/* The values of one-field .sra text (the surface setup's read with one
   month, setup/surface/index.ts:76-107): whitespace-separated numbers, an
   eight-integer header, then the nlat x 2 nlat values into out. A text
   without a whole field fails with "N032_surf_0172.sra has 0 fields,
   expected 1" (the file name of `code`). */
int exo_boundary_read_sra(int code, int nlat, const char *text, size_t length,
                          double *out);

/* ---- Native PlaSim restart files ------------------------------------
   plasim_restart / plasim_status: Fortran sequential records (uint32
   byte count, bytes, the count again), each field a 16-character name
   record followed by its FP32 or FP64 values. Host only. */

/* NATIVE_RESTART.patch (src/model/climate/gcm/checkpoint/native/
   index.ts:3-46): copies the file's `bytes` into `out` (which may be
   `restart` itself) and overwrites every record of each named field with
   values[i] (counts[i] doubles, stored at the record's precision).
   EXO_E_FORMAT with the reference's message ("Truncated native restart
   record", "Invalid native restart record", "Invalid native restart field
   name", "Native restart shape mismatch for <name>", "Native restart
   fields missing: <names>"); `out` is then unspecified. */
int exo_native_patch(const void *restart, size_t bytes, int n_fields,
                     const char *const *names, const double *const *values,
                     const size_t *counts, void *out);
/* The first record of field `name`: byte offset of its values in the file
   and their byte count (4 or 8 per value). EXO_E_FORMAT on a malformed
   file, EXO_E_ARG when the field is absent. No counterpart in the
   reference, which only patches; checkers read fields through it. */
int exo_native_find(const void *restart, size_t bytes, const char *name,
                    size_t *offset, size_t *data_bytes);

#ifdef __cplusplus
}
#endif
#endif
