"""MI355X-native ExoPlaSim time step: host package.

The compute path is libexoplasim_mi355x.so (hand-written gfx950 HIP kernels
behind the C-ABI of include/exoplasim_mi355x.h); this package holds the
Python host binding (engine.py) and the runner that mirrors the reference's
engine interface (runner.py).
"""
from .engine import (EXO_PRECISION_SPLIT, EXO_PRECISION_FP32, Engine,  # noqa: F401
                     ExoConfig, ExoError, constant_names, default_config,
                     balance_assess, boundary_files, boundary_read_sra, boundary_sra, library, native_patch,
                     native_record, slot_names, snapshot_field_names, atmosphere_resolve,
                     namelist_read, config_from_namelists)
