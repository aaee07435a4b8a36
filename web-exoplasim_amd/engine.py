"""Python host binding of libexoplasim_mi355x.so (ctypes over the C-ABI in
include/exoplasim_mi355x.h).

`Engine` holds one engine on one GPU with its members (planets); the
methods map one to one onto the C entry points. The library is loaded from
this package directory; a missing or unloadable library raises at once —
there is no CPU fallback for the step.
"""
import ctypes
import json
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexoplasim_mi355x.so")

EXO_PRECISION_FP32 = 0
EXO_PRECISION_SPLIT = 1

ERRORS = {-1: "argument", -2: "device", -3: "state", -4: "unstable", -5: "format"}


class ExoError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class ExoConfig(ctypes.Structure):
    """exo_config of include/exoplasim_mi355x.h."""

    _fields_ = [
        ("truncation", ctypes.c_int),
        ("levels", ctypes.c_int),
        ("timestep_minutes", ctypes.c_double),
        ("radius_earth", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("rotation_period_days", ctypes.c_double),
        ("pressure_bar", ctypes.c_double),
        ("gas_constant", ctypes.c_double),
        ("co2_ppm", ctypes.c_double),
        ("flux_wm2", ctypes.c_double),
        ("obliquity_deg", ctypes.c_double),
        ("eccentricity", ctypes.c_double),
        ("year_days", ctypes.c_double),
        ("ocean_depth_m", ctypes.c_double),
        ("ozone", ctypes.c_int),
        ("land_fraction", ctypes.c_double),
        ("terrain_max_m", ctypes.c_double),
        ("initial_temperature", ctypes.c_double),
        ("initial_humidity", ctypes.c_double),
        ("noise_kelvin", ctypes.c_double),
        ("sea_surface_k", ctypes.c_double),
        ("terrain_seed", ctypes.c_uint64),
        ("noise_seed", ctypes.c_uint64),
        ("storm_diagnostics", ctypes.c_int),
        ("vegetation", ctypes.c_int),
        ("initial_growth", ctypes.c_double),
        ("glaciers", ctypes.c_int),
        ("glacier_min_depth_m", ctypes.c_double),
        ("glacier_initial_height_m", ctypes.c_double),
        ("star_temperature_k", ctypes.c_double),
        ("synchronous", ctypes.c_int),
        ("substellar_longitude_deg", ctypes.c_double),
        ("fixed_orbit", ctypes.c_int),
        ("keplerian", ctypes.c_int),
        ("periapse_longitude_deg", ctypes.c_double),
        ("initial_mean_anomaly_deg", ctypes.c_double),
        ("soil_depth_scale", ctypes.c_double),
        ("soil_water_capacity_m", ctypes.c_double),
        ("soil_heat_capacity", ctypes.c_double),
        ("land_albedo", ctypes.c_double),
        ("ocean_albedo", ctypes.c_double),
        ("snow_ice_albedo", ctypes.c_double),
        ("wet_soil", ctypes.c_int),
        ("filter", ctypes.c_int),
        ("initial_state", ctypes.c_int),
        ("kick", ctypes.c_int),
        ("seed", ctypes.c_int32 * 8),
        ("clock_ms", ctypes.c_int64),
        ("nspinit", ctypes.c_int),
    ]


class ExoSurfaceField(ctypes.Structure):
    """exo_surface_field: one .sra field of a planet's surface."""

    _fields_ = [("code", ctypes.c_int), ("months", ctypes.c_int), ("values", ctypes.POINTER(ctypes.c_double))]


GASES = ("H2", "He", "N2", "O2", "Ar", "Ne", "Kr", "CO2", "H2O")


def atmosphere_resolve(partial_pressures=None, mode="partialPressures", pressure_bar=None, gas_constant=None):
    """GCM_ATMOSPHERE.resolve: (pressure bar, gas constant, CO2 ppm) of a
    gas mix given as {gas: bar or None}."""
    lib = library()
    pp = partial_pressures or {}
    arr = (ctypes.c_double * 9)(*[float("nan") if pp.get(g) is None else float(pp[g]) for g in GASES])
    out = [ctypes.c_double() for _ in range(3)]
    nan = float("nan")
    _check(lib.exo_atmosphere_resolve(arr, 1 if mode == "pressureAndGasConstant" else 0,
                                      nan if pressure_bar is None else float(pressure_bar),
                                      nan if gas_constant is None else float(gas_constant),
                                      *[ctypes.byref(o) for o in out]))
    return tuple(o.value for o in out)


def namelist_read(text, defaults):
    """GCM_NAMELIST.read: `defaults` ({key: number or list}) with the
    values one namelist file's text sets, arrays filled from the front."""
    lib = library()
    keys = list(defaults)
    lengths = [len(v) if isinstance(v, (list, tuple)) else 1 for v in defaults.values()]
    flat = []
    for v in defaults.values():
        flat += [float(x) for x in v] if isinstance(v, (list, tuple)) else [float(v)]
    karr = (ctypes.c_char_p * len(keys))(*[k.encode() for k in keys])
    larr = (ctypes.c_int * len(keys))(*lengths)
    varr = (ctypes.c_double * max(len(flat), 1))(*flat)
    _check(lib.exo_namelist_read(None if text is None else text.encode(), karr, larr, len(keys), varr))
    out, at = {}, 0
    for k, n, v in zip(keys, lengths, defaults.values()):
        out[k] = list(varr[at:at + n]) if isinstance(v, (list, tuple)) else varr[at]
        at += n
    return out


def config_from_namelists(truncation, levels, files, base=None):
    """exo_config_from_namelists: the ExoConfig a run's namelist files
    ({name: text}) carry; fields no namelist holds come from `base`
    (default_config() if None)."""
    lib = library()
    cfg = ExoConfig()
    if base is None:
        lib.exo_config_default(ctypes.byref(cfg))
    else:
        ctypes.memmove(ctypes.byref(cfg), ctypes.byref(base), ctypes.sizeof(ExoConfig))
    names = [n.encode() for n in files]
    texts = [None if t is None else t.encode() for t in files.values()]
    narr = (ctypes.c_char_p * max(len(names), 1))(*names)
    tarr = (ctypes.c_char_p * max(len(texts), 1))(*texts)
    _check(lib.exo_config_from_namelists(truncation, levels, narr, tarr, len(names), ctypes.byref(cfg)))
    return cfg


class ExoRequest(ctypes.Structure):
    """exo_request: one run of the serving loop."""

    _fields_ = [("config", ExoConfig), ("precision", ctypes.c_int), ("reserved", ctypes.c_int),
                ("arrival_s", ctypes.c_double), ("steps", ctypes.c_int64)]


class ExoResult(ctypes.Structure):
    """exo_result: a served run's times and its global means at the end."""

    _fields_ = [("arrival_s", ctypes.c_double), ("admitted_s", ctypes.c_double),
                ("finished_s", ctypes.c_double), ("tsurf", ctypes.c_double), ("toa_net", ctypes.c_double),
                ("steps", ctypes.c_int64), ("status", ctypes.c_int), ("reserved", ctypes.c_int)]


SERVE_PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double)


class ExoServeOptions(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_chunk_steps", ctypes.c_int),
                ("prepare_threads", ctypes.c_int), ("progress", SERVE_PROGRESS),
                ("progress_user", ctypes.c_void_p)]


def _month_info(info):
    return {"year": info.year, "month": info.month, "samples": info.samples, "step": info.step}


class ExoMonthInfo(ctypes.Structure):
    """exo_month_info: the calendar month a closed monthly mean covers."""

    _fields_ = [("year", ctypes.c_int), ("month", ctypes.c_int), ("samples", ctypes.c_int),
                ("step", ctypes.c_int64)]


class ExoYearlyBalance(ctypes.Structure):
    """exo_yearly_balance: global means of one calendar year's annual mean."""

    _fields_ = [("year", ctypes.c_int), ("toa_net_radiation", ctypes.c_double),
                ("surface_net_heat", ctypes.c_double), ("surface_temperature", ctypes.c_double)]


class ExoBalanceAssessment(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("years_needed", ctypes.c_int),
                ("toa_drift_per_year", ctypes.c_double), ("surface_drift_per_year", ctypes.c_double)]


BALANCE_STATUS = {0: "spinning-up", 1: "drifting", 2: "balanced"}

_lib = None


def library():
    """Loads the C-ABI library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ExoError(-2, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    S = ctypes.c_size_t
    D = ctypes.POINTER(ctypes.c_double)
    F = ctypes.POINTER(ctypes.c_float)
    sig = {
        "exo_config_default": (None, [ctypes.POINTER(ExoConfig)]),
        "exo_last_error": (ctypes.c_char_p, []),
        "exo_create": (I, [I, I, ctypes.POINTER(P)]),
        "exo_destroy": (None, [P]),
        "exo_add_member": (I, [P, ctypes.POINTER(ExoConfig), I, ctypes.POINTER(I)]),
        "exo_remove_member": (I, [P, I]),
        "exo_step": (I, [P, I]),
        "exo_step_noalloc": (I, [P, I]),
        "exo_sync": (I, [P]),
        "exo_member_step": (I, [P, I, ctypes.POINTER(ctypes.c_int64)]),
        "exo_member_date": (I, [P, I, ctypes.POINTER(I)]),
        "exo_member_date_at": (I, [P, I, ctypes.c_int64, ctypes.POINTER(I)]),
        "exo_math_eval": (I, [P, I, F, F, F, S]),
        "exo_arena_floats": (I, [P, I, ctypes.POINTER(S)]),
        "exo_slot": (I, [P, I, ctypes.c_char_p, ctypes.POINTER(S), ctypes.POINTER(S)]),
        "exo_slot_name": (ctypes.c_char_p, [I]),
        "exo_add_member_surface": (I, [P, ctypes.POINTER(ExoConfig), I, ctypes.POINTER(ExoSurfaceField), I,
                                       ctypes.POINTER(I)]),
        "exo_atmosphere_resolve": (I, [D, I, ctypes.c_double, ctypes.c_double, D, D, D]),
        "exo_namelist_read": (I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(I), I, D]),
        "exo_config_from_namelists": (I, [I, I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                                          I, ctypes.POINTER(ExoConfig)]),
        "exo_snapshot_field_name": (ctypes.c_char_p, [I, I]),
        "exo_read_field": (I, [P, I, ctypes.c_char_p, F, S]),
        "exo_write_field": (I, [P, I, ctypes.c_char_p, F, S]),
        "exo_read_table": (I, [P, I, ctypes.c_char_p, D, S]),
        "exo_table_count": (I, [P, I, ctypes.c_char_p, ctypes.POINTER(S)]),
        "exo_read_constants": (I, [P, I, I, D, S]),
        "exo_constant_name": (ctypes.c_char_p, [I]),
        "exo_phase_count": (I, [P, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "exo_dims": (I, [P, I, ctypes.POINTER(I)]),
        "exo_step_values": (I, [P, I, ctypes.c_int64, D]),
        "exo_last_step_ms": (I, [P, D]),
        "exo_probe_kernel": (I, [P, ctypes.c_char_p, I]),
        "exo_serve": (I, [P, ctypes.POINTER(ExoRequest), I, ctypes.POINTER(ExoServeOptions),
                          ctypes.POINTER(ExoResult), D]),
        "exo_probe_result": (I, [P, D, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "exo_set_publish_every": (I, [P, I]),
        "exo_set_step_graphs": (I, [P, I]),
        "exo_step_graphs": (I, [P, ctypes.POINTER(ctypes.c_int)]),
        "exo_set_step_lanes": (I, [P, I]),
        "exo_step_lanes": (I, [P, ctypes.POINTER(ctypes.c_int)]),
        "exo_set_transform_group": (I, [P, I, I, I]),
        "exo_transform_group": (I, [P, I, I, ctypes.POINTER(ctypes.c_int)]),
        "exo_autotune": (I, [P, I, ctypes.POINTER(ctypes.c_double)]),
        "exo_read_snapshot": (I, [P, I, F, S]),
        "exo_global_means": (I, [P, I, D, D]),
        "exo_checkpoint_size": (I, [P, I, ctypes.POINTER(S)]),
        "exo_checkpoint_save": (I, [P, I, ctypes.c_void_p, S]),
        "exo_checkpoint_restore": (I, [P, I, ctypes.c_void_p, S]),
        "exo_checkpoint_continue": (I, [P, I, ctypes.c_void_p, S, ctypes.POINTER(ExoConfig)]),
        "exo_set_averaging": (I, [P, I, I]),
        "exo_month_count": (I, [P, I, ctypes.POINTER(I)]),
        "exo_read_month": (I, [P, I, I, ctypes.POINTER(ExoMonthInfo), F, S]),
        "exo_read_annual": (I, [P, I, F, S]),
        "exo_yearly_balances": (I, [P, I, ctypes.POINTER(ExoYearlyBalance), I, ctypes.POINTER(I)]),
        "exo_balance_assess": (I, [ctypes.POINTER(ExoYearlyBalance), I, ctypes.POINTER(ExoBalanceAssessment)]),
        "exo_read_accumulator": (I, [P, I, ctypes.POINTER(ExoMonthInfo), D, S]),
        "exo_load_averager": (I, [P, I, ctypes.POINTER(ExoMonthInfo), F, I, ctypes.POINTER(ExoYearlyBalance), I,
                                  ctypes.POINTER(ExoMonthInfo), D, S]),
        "exo_read_climatology": (I, [P, I, I, ctypes.POINTER(ExoMonthInfo), F, S, ctypes.POINTER(I)]),
        "exo_native_patch": (I, [ctypes.c_char_p, S, I, ctypes.POINTER(ctypes.c_char_p),
                                 ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.POINTER(S),
                                 ctypes.c_void_p]),
        "exo_native_find": (I, [ctypes.c_char_p, S, ctypes.c_char_p, ctypes.POINTER(S), ctypes.POINTER(S)]),
        "exo_boundary_from_terrain": (I, [P, F, F, I, I, I, ctypes.c_double, D, D]),
        "exo_boundary_sra": (I, [I, I, D, ctypes.c_char_p, S, ctypes.POINTER(S)]),
        "exo_boundary_read_sra": (I, [I, I, ctypes.c_char_p, S, D]),
        "exo_add_member_boundary": (I, [P, ctypes.POINTER(ExoConfig), I, D, D, ctypes.POINTER(I)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def balance_assess(years):
    """GCM_BALANCE.assess over yearly balances (dicts or (toa, surface)
    pairs), oldest first: {"status", "years_needed", "toa_drift_per_year",
    "surface_drift_per_year"}. Host only."""
    lib = library()
    arr = (ExoYearlyBalance * max(1, len(years)))()
    for i, y in enumerate(years):
        if isinstance(y, dict):
            arr[i] = ExoYearlyBalance(y.get("year", i), y["toa_net_radiation"], y["surface_net_heat"],
                                      y.get("surface_temperature", 0.0))
        else:
            arr[i] = ExoYearlyBalance(i, y[0], y[1], 0.0)
    out = ExoBalanceAssessment()
    _check(lib.exo_balance_assess(arr, len(years), ctypes.byref(out)))
    return {"status": BALANCE_STATUS[out.status], "years_needed": out.years_needed,
            "toa_drift_per_year": out.toa_drift_per_year, "surface_drift_per_year": out.surface_drift_per_year}


def native_patch(restart, fields):
    """NATIVE_RESTART.patch: a copy of the native restart file `restart`
    (bytes) with each named field's records overwritten by its values
    ({name: sequence of floats}). Host only."""
    lib = library()
    data = bytes(restart)
    names = list(fields)
    arrays = [np.ascontiguousarray(np.asarray(fields[n], dtype=np.float64).reshape(-1)) for n in names]
    c_names = (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    c_values = (ctypes.POINTER(ctypes.c_double) * max(1, len(names)))(
        *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrays])
    c_counts = (ctypes.c_size_t * max(1, len(names)))(*[a.size for a in arrays])
    out = ctypes.create_string_buffer(len(data))
    _check(lib.exo_native_patch(data, len(data), len(names), c_names, c_values, c_counts, out))
    return out.raw


LAND_SEA_CODE, OROGRAPHY_CODE = 172, 129


def boundary_sra(code, nlat, values):
    """toSra: the .sra text of one field of an nlat x 2 nlat grid."""
    lib = library()
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
    if v.size != 2 * nlat * nlat:
        raise ExoError(-1, f"an N{nlat} field has {2 * nlat * nlat} values")
    n = ctypes.c_size_t()
    ptr = v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _check(lib.exo_boundary_sra(code, nlat, ptr, None, 0, ctypes.byref(n)))
    out = ctypes.create_string_buffer(n.value)
    _check(lib.exo_boundary_sra(code, nlat, ptr, out, n.value, ctypes.byref(n)))
    return out.raw[:n.value].decode()


def boundary_read_sra(code, nlat, text):
    """The nlat x 2 nlat values of a one-field .sra text."""
    lib = library()
    raw = text.encode() if isinstance(text, str) else bytes(text)
    out = np.empty(2 * nlat * nlat)
    _check(lib.exo_boundary_read_sra(code, nlat, raw, len(raw), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out.reshape(nlat, 2 * nlat)


def boundary_files(nlat, land_fraction, geopotential):
    """GCM_BOUNDARY.files: {file name: .sra text} of the land-sea mask and
    the orography."""
    prefix = f"N{nlat:03d}_surf_0"
    return {f"{prefix}{LAND_SEA_CODE}.sra": boundary_sra(LAND_SEA_CODE, nlat, land_fraction),
            f"{prefix}{OROGRAPHY_CODE}.sra": boundary_sra(OROGRAPHY_CODE, nlat, geopotential)}


def native_record(restart, name):
    """The bytes of field `name`'s values in a native restart file (FP32 or
    FP64 as the model that wrote it was built). Host only."""
    lib = library()
    data = bytes(restart)
    off, n = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib.exo_native_find(data, len(data), name.encode(), ctypes.byref(off), ctypes.byref(n)))
    return data[off.value:off.value + n.value]


def exported_symbols():
    """Names the C-ABI header declares (for the load/export check)."""
    return [n for n in _header_functions()]


def _header_functions():
    import re
    path = os.path.join(os.path.dirname(HERE), "include", "exoplasim_mi355x.h")
    text = open(path).read()
    return sorted(set(re.findall(r"\b(exo_[a-z0-9_]+)\s*\(", text)))


def default_config(**overrides):
    cfg = ExoConfig()
    library().exo_config_default(ctypes.byref(cfg))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


def _check(rc):
    if rc != 0:
        raise ExoError(rc, library().exo_last_error().decode())


def slot_names():
    lib = library()
    out = []
    i = 0
    while True:
        n = lib.exo_slot_name(i)
        if n is None:
            return out
        out.append(n.decode())
        i += 1


def snapshot_field_names(levels):
    """Ids of the snapshot planes of a member of `levels` levels."""
    lib = library()
    out = []
    while True:
        n = lib.exo_snapshot_field_name(levels, len(out))
        if n is None:
            return out
        out.append(n.decode())


def constant_names():
    lib = library()
    out = []
    i = 0
    while True:
        n = lib.exo_constant_name(i)
        if n is None:
            return out
        out.append(n.decode())
        i += 1


class Engine:
    """One engine (device < 0: host-only set-up, no stepping)."""

    def __init__(self, device=0, max_members=1):
        self.lib = library()
        h = ctypes.c_void_p()
        _check(self.lib.exo_create(device, max_members, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.exo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # members ------------------------------------------------------------
    def add_member(self, cfg=None, precision=EXO_PRECISION_FP32, land_fraction=None, geopotential=None):
        """Adds a planet; with land_fraction and geopotential (its grid's
        cells) its land and orography come from them instead of the seeded
        terrain."""
        cfg = cfg or default_config()
        mid = ctypes.c_int()
        if land_fraction is None and geopotential is None:
            _check(self.lib.exo_add_member(self.h, ctypes.byref(cfg), precision, ctypes.byref(mid)))
            return mid.value
        lf = np.ascontiguousarray(np.asarray(land_fraction, dtype=np.float64).reshape(-1))
        geo = np.ascontiguousarray(np.asarray(geopotential, dtype=np.float64).reshape(-1))
        nlat = {21: 32, 42: 64, 63: 96}.get(cfg.truncation, 0)
        if lf.size != 2 * nlat * nlat or geo.size != lf.size:
            raise ExoError(-1, f"T{cfg.truncation} boundary fields need {2 * nlat * nlat} values each")
        _check(self.lib.exo_add_member_boundary(self.h, ctypes.byref(cfg), precision,
                                                lf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                geo.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                ctypes.byref(mid)))
        return mid.value

    def add_member_surface(self, cfg=None, fields=None, precision=EXO_PRECISION_FP32):
        """Adds a planet whose surface comes from .sra fields: {code:
        array of months x the grid's cells}; codes absent take the
        reference's fallbacks."""
        cfg = cfg or default_config()
        arrays, table = [], (ExoSurfaceField * max(1, len(fields or {})))()
        nlat = {21: 32, 42: 64, 63: 96}.get(cfg.truncation, 0)
        nhor = 2 * nlat * nlat
        for i, (code, values) in enumerate(sorted((fields or {}).items())):
            a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
            if nhor == 0 or a.size % nhor:
                raise ExoError(-1, f"field {code} does not hold whole T{cfg.truncation} grids")
            arrays.append(a)
            table[i].code = int(code)
            table[i].months = a.size // nhor
            table[i].values = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        mid = ctypes.c_int()
        _check(self.lib.exo_add_member_surface(self.h, ctypes.byref(cfg), precision, table, len(arrays),
                                               ctypes.byref(mid)))
        return mid.value

    def remove_member(self, member):
        _check(self.lib.exo_remove_member(self.h, member))

    def dims(self, member):
        d = (ctypes.c_int * 4)()
        _check(self.lib.exo_dims(self.h, member, d))
        return tuple(d)

    def phases(self, member):
        p, s = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.exo_phase_count(self.h, member, ctypes.byref(p), ctypes.byref(s)))
        return p.value, s.value

    # stepping -----------------------------------------------------------
    def step(self, steps=1):
        _check(self.lib.exo_step(self.h, steps))

    def step_noalloc(self, steps=1):
        """exo_step without any allocation (calendar rows in the library's
        static device array; single launches)."""
        _check(self.lib.exo_step_noalloc(self.h, steps))

    def sync(self):
        _check(self.lib.exo_sync(self.h))

    def last_step_ms(self):
        v = ctypes.c_double()
        _check(self.lib.exo_last_step_ms(self.h, ctypes.byref(v)))
        return v.value

    def serve(self, requests, hbm_budget_bytes=0, max_chunk_steps=32, prepare_threads=4, progress=None):
        """Runs the serving loop over [(arrival_s, config, steps)] (FP32
        form); returns (results as dicts in request order, wall seconds).
        progress(finished, count, elapsed_s), when given, is called after
        each batched advance that ended runs."""
        n = len(requests)
        req = (ExoRequest * max(n, 1))()
        for i, (arrival, cfg, steps) in enumerate(requests):
            req[i].config = cfg
            req[i].precision = EXO_PRECISION_FP32
            req[i].arrival_s = float(arrival)
            req[i].steps = int(steps)
        res = (ExoResult * max(n, 1))()
        cb = SERVE_PROGRESS(lambda _u, done, total, t: progress(done, total, t)) if progress else SERVE_PROGRESS()
        opt = ExoServeOptions(hbm_budget_bytes, max_chunk_steps, prepare_threads, cb, None)
        wall = ctypes.c_double()
        _check(self.lib.exo_serve(self.h, req, n, ctypes.byref(opt), res, ctypes.byref(wall)))
        out = [{name: getattr(res[i], name) for name, _ in ExoResult._fields_ if name != "reserved"}
               for i in range(n)]
        return out, wall.value

    def probe_kernel(self, kernel, max_launches=4096):
        """Times each later launch of `kernel` with HIP events (None stops)."""
        _check(self.lib.exo_probe_kernel(self.h, kernel.encode() if kernel else None, max_launches))

    def probe_result(self):
        """(total ms, launches, missed launches) of the probed kernel."""
        t, n, miss = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        _check(self.lib.exo_probe_result(self.h, ctypes.byref(t), ctypes.byref(n), ctypes.byref(miss)))
        return t.value, n.value, miss.value

    def member_step(self, member):
        v = ctypes.c_int64()
        _check(self.lib.exo_member_step(self.h, member, ctypes.byref(v)))
        return v.value

    def member_date(self, member):
        """(year, month, day, hour, minute) of the member's step counter."""
        d = (ctypes.c_int * 5)()
        _check(self.lib.exo_member_date(self.h, member, d))
        return tuple(d)

    def member_date_at(self, member, step):
        """(year, month, day, hour, minute) of model step `step`."""
        d = (ctypes.c_int * 5)()
        _check(self.lib.exo_member_date_at(self.h, member, step, d))
        return tuple(d)

    def math_eval(self, function, a, b=None):
        """The device's transcendental routine `function` (exo_math_eval)
        over float32 arrays."""
        a = np.ascontiguousarray(a, dtype=np.float32)
        out = np.empty_like(a)
        bp = None
        if b is not None:
            b = np.ascontiguousarray(np.broadcast_to(b, a.shape), dtype=np.float32)
            bp = b.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        _check(self.lib.exo_math_eval(self.h, function, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), bp,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size))
        return out

    # layout and fields --------------------------------------------------
    def arena_floats(self, member):
        n = ctypes.c_size_t()
        _check(self.lib.exo_arena_floats(self.h, member, ctypes.byref(n)))
        return n.value

    def slot(self, member, name):
        off, n = ctypes.c_size_t(), ctypes.c_size_t()
        _check(self.lib.exo_slot(self.h, member, name.encode(), ctypes.byref(off), ctypes.byref(n)))
        return off.value, n.value

    def layout(self, member):
        return [(name,) + self.slot(member, name) for name in slot_names()]

    def read_field(self, member, name):
        _, n = self.slot(member, name)
        out = np.empty(n, dtype=np.float32)
        _check(self.lib.exo_read_field(self.h, member, name.encode(),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n))
        return out

    def write_field(self, member, name, values):
        values = np.ascontiguousarray(values, dtype=np.float32)
        _check(self.lib.exo_write_field(self.h, member, name.encode(),
                                        values.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), values.size))

    def read_table(self, member, name):
        n = ctypes.c_size_t()
        _check(self.lib.exo_table_count(self.h, member, name.encode(), ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.float64)
        _check(self.lib.exo_read_table(self.h, member, name.encode(),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.value))
        return out

    def constants(self, member, phase):
        names = constant_names()
        out = np.empty(len(names), dtype=np.float64)
        _check(self.lib.exo_read_constants(self.h, member, phase,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(names)))
        return dict(zip(names, out.tolist()))

    def step_values(self, member, step):
        out = np.empty(16, dtype=np.float64)
        _check(self.lib.exo_step_values(self.h, member, step,
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def set_transform_group(self, truncation, levels, members):
        """Members of one (truncation, levels) class per transform workgroup."""
        _check(self.lib.exo_set_transform_group(self.h, truncation, levels, members))

    def transform_group(self, truncation, levels):
        n = ctypes.c_int()
        _check(self.lib.exo_transform_group(self.h, truncation, levels, ctypes.byref(n)))
        return n.value

    def autotune(self, trial_steps=8):
        """Times the transform group sizes on the active batch and keeps
        the fastest per class; the state is left as it was. Returns the
        best trial's device ms per step."""
        ms = ctypes.c_double()
        _check(self.lib.exo_autotune(self.h, trial_steps, ctypes.byref(ms)))
        return ms.value

    def step_graphs(self):
        """Whether steps replay the captured step graph."""
        n = ctypes.c_int()
        _check(self.lib.exo_step_graphs(self.h, ctypes.byref(n)))
        return bool(n.value)

    def set_step_graphs(self, enable):
        """Replay steps from a captured HIP graph (True, the default) or
        launch their kernels one by one (False)."""
        _check(self.lib.exo_set_step_graphs(self.h, 1 if enable else 0))

    def step_lanes(self):
        """Streams a replayed step's members run on (1 to 4)."""
        n = ctypes.c_int()
        _check(self.lib.exo_step_lanes(self.h, ctypes.byref(n)))
        return n.value

    def set_step_lanes(self, lanes):
        """Run a replayed step's members on one stream (1) or split over 2
        to 4 concurrent streams (2 is the default); results are
        bit-identical."""
        _check(self.lib.exo_set_step_lanes(self.h, lanes))

    def set_publish_every(self, every):
        _check(self.lib.exo_set_publish_every(self.h, every))

    def read_snapshot(self, member):
        _, n = self.slot(member, "snapshot")
        out = np.empty(n, dtype=np.float32)
        _check(self.lib.exo_read_snapshot(self.h, member,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n))
        nhor = self.dims(member)[1] * self.dims(member)[2]
        return out.reshape(-1, nhor)

    # ---- averages and balance (GCM_AVERAGES / GCM_BALANCE on the device)

    def set_averaging(self, member, enable=True):
        _check(self.lib.exo_set_averaging(self.h, member, 1 if enable else 0))

    def month_count(self, member):
        n = ctypes.c_int()
        _check(self.lib.exo_month_count(self.h, member, ctypes.byref(n)))
        return n.value

    def read_month(self, member, index):
        """(info dict, mean snapshot (fields, NHOR)) of closed month `index`
        (0 = oldest held)."""
        _, n = self.slot(member, "snapshot")
        info = ExoMonthInfo()
        out = np.empty(n, dtype=np.float32)
        _check(self.lib.exo_read_month(self.h, member, index, ctypes.byref(info),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n))
        nhor = self.dims(member)[1] * self.dims(member)[2]
        return _month_info(info), out.reshape(-1, nhor)

    def read_annual(self, member):
        _, n = self.slot(member, "snapshot")
        out = np.empty(n, dtype=np.float32)
        _check(self.lib.exo_read_annual(self.h, member, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n))
        nhor = self.dims(member)[1] * self.dims(member)[2]
        return out.reshape(-1, nhor)

    def yearly_balances(self, member):
        n = ctypes.c_int()
        _check(self.lib.exo_yearly_balances(self.h, member, None, 0, ctypes.byref(n)))
        out = (ExoYearlyBalance * max(1, n.value))()
        _check(self.lib.exo_yearly_balances(self.h, member, out, n.value, ctypes.byref(n)))
        return [{"year": y.year, "toa_net_radiation": y.toa_net_radiation, "surface_net_heat": y.surface_net_heat,
                 "surface_temperature": y.surface_temperature} for y in out[:n.value]]

    def read_accumulator(self, member):
        """(info dict, FP64 sum (fields, NHOR)) of the month being summed, or
        (None, None) when none is open (GcmAverager.current)."""
        _, n = self.slot(member, "snapshot")
        info = ExoMonthInfo()
        _check(self.lib.exo_read_accumulator(self.h, member, ctypes.byref(info), None, 0))
        if info.samples == 0:
            return None, None
        out = np.empty(n, dtype=np.float64)
        _check(self.lib.exo_read_accumulator(self.h, member, ctypes.byref(info),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        nhor = self.dims(member)[1] * self.dims(member)[2]
        return _month_info(info), out.reshape(-1, nhor)

    def averager(self, member):
        """The member's whole averager, as a checkpoint carries it
        (GcmAverager: current, months, years)."""
        info, total = self.read_accumulator(member)
        months = [self.read_month(member, i) for i in range(self.month_count(member))]
        return {"current": None if info is None else dict(info, sum=total),
                "months": months, "years": self.yearly_balances(member)}

    def load_averager(self, member, averager):
        """Replaces the member's averages with `averager` (the dict
        averager() returns)."""
        _, n = self.slot(member, "snapshot")
        months = averager.get("months") or []
        years = averager.get("years") or []
        current = averager.get("current")
        infos = (ExoMonthInfo * max(1, len(months)))()
        values = np.zeros((max(1, len(months)), n), dtype=np.float32)
        for i, (info, mean) in enumerate(months):
            infos[i] = ExoMonthInfo(info["year"], info["month"], info["samples"], info["step"])
            values[i] = np.asarray(mean, dtype=np.float32).reshape(-1)
        balances = (ExoYearlyBalance * max(1, len(years)))()
        for i, y in enumerate(years):
            balances[i] = ExoYearlyBalance(y["year"], y["toa_net_radiation"], y["surface_net_heat"],
                                           y["surface_temperature"])
        cur = ExoMonthInfo()
        total = np.zeros(n, dtype=np.float64)
        if current is not None:
            cur = ExoMonthInfo(current["year"], current["month"], current["samples"], current["step"])
            total = np.ascontiguousarray(np.asarray(current["sum"], dtype=np.float64).reshape(-1))
        _check(self.lib.exo_load_averager(
            self.h, member, infos, values.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(months),
            balances, len(years), ctypes.byref(cur), total.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))

    def read_climatology(self, member, month):
        """(info dict, mean (fields, NHOR), years) of calendar month `month`
        over the last complete years held; (None, None, 0) with none."""
        _, n = self.slot(member, "snapshot")
        info = ExoMonthInfo()
        years = ctypes.c_int()
        out = np.empty(n, dtype=np.float32)
        _check(self.lib.exo_read_climatology(self.h, member, month, ctypes.byref(info),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n,
                                             ctypes.byref(years)))
        if years.value == 0:
            return None, None, 0
        nhor = self.dims(member)[1] * self.dims(member)[2]
        return _month_info(info), out.reshape(-1, nhor), years.value

    def boundary_from_terrain(self, elevation_m, land, nlat, gravity_ms2=9.80665):
        """GCM_BOUNDARY.fromTerrain on this engine's device: (land fraction,
        geopotential), each (nlat, 2 nlat), of a (height, width) raster."""
        elev = np.ascontiguousarray(np.asarray(elevation_m, dtype=np.float32))
        lnd = np.ascontiguousarray(np.asarray(land, dtype=np.float32))
        if elev.ndim != 2 or lnd.shape != elev.shape:
            raise ExoError(-1, "Terrain arrays do not match the image dimensions")
        height, width = elev.shape
        fraction = np.empty(2 * nlat * nlat)
        geo = np.empty(2 * nlat * nlat)
        _check(self.lib.exo_boundary_from_terrain(
            self.h, elev.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            lnd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), width, height, nlat, gravity_ms2,
            fraction.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            geo.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return fraction.reshape(nlat, 2 * nlat), geo.reshape(nlat, 2 * nlat)

    def global_means(self, member):
        t, f = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.exo_global_means(self.h, member, ctypes.byref(t), ctypes.byref(f)))
        return t.value, f.value

    def checkpoint(self, member):
        n = ctypes.c_size_t()
        _check(self.lib.exo_checkpoint_size(self.h, member, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(self.lib.exo_checkpoint_save(self.h, member, buf, n.value))
        return buf.raw

    def restore(self, member, data, previous=None):
        """Restores a checkpoint; with `previous` (the config it was saved
        under) the fields a changed setting invalidates take this member's
        initial values (GCM_CONTINUATION.fields)."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        if previous is None:
            _check(self.lib.exo_checkpoint_restore(self.h, member, buf, len(data)))
        else:
            _check(self.lib.exo_checkpoint_continue(self.h, member, buf, len(data), ctypes.byref(previous)))

    def read_arena(self, member):
        """Every slot, in arena order, as one float32 array."""
        parts = [self.read_field(member, name) for name in slot_names()]
        return np.concatenate(parts)
