"""Drives the FP64 CPU oracle (oracle/liboracle.so) over a member's FP64
set-up as the engine computed it, so the oracle and the GPU start from the
same inputs. Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIBS = {"fp64": os.path.join(ROOT, "oracle", "liboracle.so"),
               "fp32": os.path.join(ROOT, "oracle", "liboracle32.so"),
               "fp32fma": os.path.join(ROOT, "oracle", "liboracle32fma.so")}
CTYPES = {"fp64": (ctypes.c_double, np.float64), "fp32": (ctypes.c_float, np.float32),
          "fp32fma": (ctypes.c_float, np.float32)}


def _member_struct(ct):
    class _Member(ctypes.Structure):
        _fields_ = [
            ("ntru", ctypes.c_int), ("nlat", ctypes.c_int), ("nlon", ctypes.c_int),
            ("nlev", ctypes.c_int),
            ("arena", ctypes.POINTER(ct)),
            ("tables", ctypes.POINTER(ct)),
            ("constants", ctypes.POINTER(ct)),
            ("slot_offsets", ctypes.POINTER(ctypes.c_long)),
            ("table_offsets", ctypes.POINTER(ctypes.c_long)),
        ]
    return _Member


_MEMBER = {p: _member_struct(CTYPES[p][0]) for p in CTYPES}
_LIBS = {}


def oracle_library(precision="fp64"):
    if precision in _LIBS:
        return _LIBS[precision]
    path = ORACLE_LIBS[precision]
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    lib = ctypes.CDLL(path)
    ct = CTYPES[precision][0]
    Mem = _MEMBER[precision]
    for n in ("oracle_slot_names", "oracle_table_names", "oracle_constant_names"):
        getattr(lib, n).restype = ctypes.POINTER(ctypes.c_char_p)
    lib.oracle_step.argtypes = [ctypes.POINTER(Mem), ctypes.c_int]
    lib.oracle_physics.argtypes = [ctypes.POINTER(Mem)]
    lib.oracle_routine.argtypes = [ctypes.POINTER(Mem), ctypes.c_char_p]
    lib.oracle_routine.restype = ctypes.c_int
    P = ctypes.POINTER(ct)
    lib.oracle_transform_inverse.argtypes = [ctypes.POINTER(Mem), P, P, ctypes.c_int, ctypes.c_int]
    lib.oracle_transform_forward.argtypes = [ctypes.POINTER(Mem), P, P, ctypes.c_int]
    lib.oracle_radiation_fit.argtypes = [ctypes.POINTER(Mem), ctypes.c_int, P, ctypes.c_int, P]
    lib.oracle_radiation_fit.restype = ctypes.c_int
    lib.oracle_storm_potential.argtypes = [ctypes.POINTER(Mem), ctypes.c_int, P]
    lib.oracle_storm_potential.restype = ctypes.c_int
    lib.oracle_pack.argtypes = [ctypes.POINTER(Mem)]
    _LIBS[precision] = lib
    return lib


def _names(fn):
    out = []
    p = fn()
    i = 0
    while p[i] is not None:
        out.append(p[i].decode())
        i += 1
    return out


def _ptr(a):
    ct = ctypes.c_float if a.dtype == np.float32 else ctypes.c_double
    return a.ctypes.data_as(ctypes.POINTER(ct))


class OracleRun:
    """An FP64 copy of one engine member, stepped by the oracle."""

    def __init__(self, engine, member, table_overrides=None, precision="fp64"):
        self.precision = precision
        self.dtype = CTYPES[precision][1]
        self.lib = oracle_library(precision)
        self.engine = engine
        self.member = member
        self.dims = engine.dims(member)
        self.nphase, self.nstartup = engine.phases(member)
        self.layout = {name: (off, n) for name, off, n in engine.layout(member)}
        self.arena = engine.read_table(member, "initial").astype(self.dtype)
        self.table_names = _names(self.lib.oracle_table_names)
        self.const_names = _names(self.lib.oracle_constant_names)
        slot_names = _names(self.lib.oracle_slot_names)
        self.slot_off = np.array([self.layout[n][0] for n in slot_names], dtype=np.int64)
        overrides = table_overrides or {}
        parts, offs, at = [], [], 0
        self.table_at = {}
        for n in self.table_names:
            t = overrides.get(n)
            if t is None:
                t = engine.read_table(member, n)
            parts.append(np.asarray(t, dtype=self.dtype))
            offs.append(at)
            self.table_at[n] = (at, len(t))
            at += len(t)
        self.tables = np.concatenate(parts)
        self.table_off = np.array(offs, dtype=np.int64)
        self.bm1 = [engine.read_table(member, f"bm1_phase{p}").astype(self.dtype) for p in range(self.nphase)]
        self.consts = []
        for p in range(self.nphase):
            k = engine.constants(member, p)
            self.consts.append(np.array([k[n] for n in self.const_names], dtype=self.dtype))
        self.phase = 0
        self.step_no = None
        self.carbon_interval = int(round(engine.read_table(member, "calendar")[0]))

    def _struct(self, consts):
        m = _MEMBER[self.precision]()
        m.ntru, m.nlat, m.nlon, m.nlev = self.dims
        m.arena = _ptr(self.arena)
        m.tables = _ptr(self.tables)
        m.constants = _ptr(consts)
        m.slot_offsets = self.slot_off.ctypes.data_as(ctypes.POINTER(ctypes.c_long))
        m.table_offsets = self.table_off.ctypes.data_as(ctypes.POINTER(ctypes.c_long))
        return m

    def field(self, name):
        off, n = self.layout[name]
        return self.arena[off:off + n]

    def _load_step_values(self):
        """STEPPER.values into the arena's per-step slots, as the engine's
        params-to-state copy does (stepper/index.ts:242-255)."""
        if self.step_no is None:
            self.step_no = self.engine.member_step(self.member)
            self.carbon = 0
        vals = self.engine.step_values(self.member, self.step_no)
        interval = self.carbon_interval
        weathers = (self.carbon + 1) % interval == 0
        vals[12] = 1.0 if weathers else 0.0
        self.carbon = 0 if weathers else self.carbon + 1
        vals[15] = 0.0
        for name, lo, hi in (("radiation_step", 0, 5), ("surface_step", 5, 12),
                             ("carbon_step", 12, 13), ("snapshot_step", 13, 16)):
            off, n = self.layout[name]
            self.arena[off:off + n] = vals[lo:hi].astype(self.dtype)

    def step(self, steps=1, physics=False):
        for _ in range(steps):
            p = min(self.phase, self.nphase - 1)
            self._load_step_values()
            if self.phase >= self.nstartup:
                self.step_no += 1
            at, n = self.table_at["bm1"]
            self.tables[at:at + n] = self.bm1[p]
            consts = self.consts[p]
            m = self._struct(consts)
            self.lib.oracle_step(ctypes.byref(m), 1 if physics else 0)
            self.phase += 1

    def routine(self, name, phase=None):
        """Runs one parameterization of the oracle on the current arena
        with the constants of `phase` (default: the main loop's)."""
        consts = self.consts[-1 if phase is None else phase]
        m = self._struct(consts)
        if self.lib.oracle_routine(ctypes.byref(m), name.encode()) != 0:
            raise KeyError(name)

    FITS = {"lwr_h2o": 0, "lwr_co2": 1, "lwr_h2o_transmission": 2, "lwr_o3": 3, "swr_o3": 4, "swr_h2o": 5}

    def radiation_fit(self, name, paths):
        """One of the radiation's absorber fits at the given paths."""
        m = self._struct(self.consts[-1])
        u = np.ascontiguousarray(paths, dtype=self.dtype)
        out = np.zeros_like(u)
        if self.lib.oracle_radiation_fit(ctypes.byref(m), self.FITS[name], _ptr(u), len(u), _ptr(out)) != 0:
            raise KeyError(name)
        return out

    def storm_potential(self, column):
        """(outflow ratio, outflow temperature, outflow pressure, saturated
        CAPE, ambient CAPE, intensity) of the storm's potential-intensity
        iteration on one column, or None when it is skipped or fails."""
        m = self._struct(self.consts[-1])
        out = np.zeros(6, dtype=self.dtype)
        if self.lib.oracle_storm_potential(ctypes.byref(m), int(column), _ptr(out)) != 0:
            return None
        return out

    def pack(self):
        """The snapshot pack of one main-loop step on the current arena
        (tracking, and the planes when its snapshot_step says publish)."""
        m = self._struct(self.consts[-1])
        self.lib.oracle_pack(ctypes.byref(m))

    def set_constant(self, name, value, phase=None):
        self.consts[-1 if phase is None else phase][self.const_names.index(name)] = value

    def constant(self, name, phase=None):
        return float(self.consts[-1 if phase is None else phase][self.const_names.index(name)])

    def table(self, name):
        at, n = self.table_at[name]
        return self.tables[at:at + n]

    def inverse(self, sp, planes, dmu=False):
        consts = self.consts[-1]
        m = self._struct(consts)
        nlat, nlon = self.dims[1], self.dims[2]
        sp = np.ascontiguousarray(sp, dtype=self.dtype)
        out = np.zeros(planes * nlat * nlon, dtype=self.dtype)
        self.lib.oracle_transform_inverse(ctypes.byref(m), _ptr(sp), _ptr(out), planes, 1 if dmu else 0)
        return out

    def forward(self, grid, planes):
        consts = self.consts[-1]
        m = self._struct(consts)
        t = self.dims[0]
        grid = np.ascontiguousarray(grid, dtype=self.dtype)
        out = np.zeros(planes * (t + 1) * (t + 2), dtype=self.dtype)
        self.lib.oracle_transform_forward(ctypes.byref(m), _ptr(grid), _ptr(out), planes)
        return out
