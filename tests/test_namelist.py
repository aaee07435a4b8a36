"""The namelist path (GCM_NAMELIST, namelist/index.ts:40-233, and the
set-up's reading of the files, setup/pumamod/index.ts:103-232,
setup/modules/index.ts:41-305): the Fortran-style read of one file
(exo_namelist_read), the files read back into an exo_config
(exo_config_from_namelists) with every key the engine cannot vary refused
by name, and the JavaScript host's edits/apply over the base namelists
(node/namelist.js) read back into the same configuration that the direct
GcmConfig mapping gives. CPU only.

tests/golden/base_namelists.json holds the reference's base namelist files
(public/gcm/t21/namelists.json; t42's and t63's are the same text).
"""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "web-exoplasim_amd", "node")
BASE = json.load(open(os.path.join(ROOT, "tests", "golden", "base_namelists.json")))["files"]

needs_node = pytest.mark.skipif(
    shutil.which("node") is None or not os.path.exists(os.path.join(NODE_DIR, "exoplasim_napi.node")),
    reason="node or the built addon is absent")


def _node(script, timeout=120):
    proc = subprocess.run(["node", "-e", script], capture_output=True, text=True, timeout=timeout, cwd=NODE_DIR)
    assert proc.returncode == 0, proc.stderr
    return json.loads(proc.stdout.strip().splitlines()[-1])


READ_TEXT = """ &plasim_nl
 KICK        =     2
 seed = 11,-3
 MPSTEP  =  2.5d1
 ndel = 4, 6 ,
 psurf = 1.0D+5
 EMPTY =
 JUNK = 12 abc
 HEX = 0x1F
 NEG = -.5e-1
 INF = -Infinity
 KICK = 3
\tTABBED\t=\t7\t
 NOEQUALS 5
 1BAD = 4
 /END
"""


def test_fortran_style_read_known_answers(exo):
    """Keys are case-blind and the file's last line for a key wins; a 'd'
    exponent reads as 'e'; a comma list is written over the defaults from
    the front (a trailing empty part reads 0); an empty value reads 0, a
    malformed one NaN; lines without '=' or with a key not starting with a
    letter are skipped, as are keys absent from the file."""
    got = exo.namelist_read(READ_TEXT, {
        "kick": 1, "SEED": [0] * 8, "mpstep": 0.0, "NDEL": [2.0] * 5, "PSURF": 101100.0, "EMPTY": 9.0,
        "JUNK": 1.0, "HEX": 0.0, "NEG": 0.0, "INF": 0.0, "TABBED": 0.0, "NOEQUALS": 1.0, "BAD": 2.0,
        "ABSENT": [1.5, 2.5]})
    assert got["kick"] == 3
    assert got["SEED"] == [11, -3, 0, 0, 0, 0, 0, 0]
    assert got["mpstep"] == 25.0
    assert got["NDEL"] == [4, 6, 0, 2, 2]
    assert got["PSURF"] == 1e5
    assert got["EMPTY"] == 0
    assert math.isnan(got["JUNK"])
    assert got["HEX"] == 31 and got["NEG"] == -0.05 and got["INF"] == -math.inf
    assert got["TABBED"] == 7 and got["NOEQUALS"] == 1 and got["BAD"] == 2
    assert got["ABSENT"] == [1.5, 2.5]
    # numbers beyond a key's length are dropped
    assert exo.namelist_read(" X = 1,2,3\n", {"X": [0.0, 0.0]})["X"] == [1, 2]
    assert exo.namelist_read(None, {"X": 4.0})["X"] == 4
    # CRLF line ends
    assert exo.namelist_read(" &g\r\n A = 1.5\r\n B=2\r\n /END\r\n", {"A": 0.0, "B": 0.0}) == {"A": 1.5, "B": 2}


@needs_node
def test_javascript_read_and_apply_agree(exo):
    """node/namelist.js's read gives the C read's numbers on the same text
    (NaN as null); apply replaces a key's line in place and adds a new key
    before /END."""
    defaults = {"KICK": 1, "SEED": [0] * 8, "MPSTEP": 0, "NDEL": [2] * 5, "PSURF": 101100, "EMPTY": 9,
                "HEX": 0, "NEG": 0, "TABBED": 0, "NOEQUALS": 1}
    out = _node(f"""
const N = require('./namelist.js')
const r = N.read({json.dumps(READ_TEXT)}, {json.dumps(defaults)})
const applied = N.apply(' &planet_nl\\n  ecCen =  0.0\\n OBLIQ = 1\\n /END\\n', {{ECCEN: '0.2', GA: '3.7'}})
console.log(JSON.stringify({{r, applied}}))
""")
    c = exo.namelist_read(READ_TEXT, {k: [float(x) for x in v] if isinstance(v, list) else float(v)
                                      for k, v in defaults.items()})
    assert out["r"] == c
    assert out["applied"] == " &planet_nl\n ECCEN = 0.2\n OBLIQ = 1\n GA = 3.7\n /END\n"


def _apply(text, values):
    """GCM_NAMELIST.apply for the tests' hand edits (keys the base writes
    one per line)."""
    lines = text.split("\n")
    for key, value in values.items():
        for i, line in enumerate(lines):
            if line.strip().upper().startswith(key) and line.split("=")[0].strip().upper() == key:
                lines[i] = f" {key} = {value}"
                break
        else:
            end = next(i for i, line in enumerate(lines) if line.strip().upper() == "/END")
            lines.insert(end, f" {key} = {value}")
    return "\n".join(lines)


def _planet_files(**edits):
    """The base files with a Mars-like planet's edits, written by hand from
    the edits' rules (namelist/index.ts:45-99)."""
    files = dict(BASE)
    files["planet_namelist"] = _apply(files["planet_namelist"], {
        "GA": "3.71", "PLARAD": "3389500", "GSOL0": "586", "ECCEN": "0.0934", "OBLIQ": "25.19", "NFIXORB": "1",
        "NGENKEPLERIAN": "0", "GASCON": "188.9", "SIDEREAL_YEAR": str(686.98 * 86400), "MVELP": "336.1",
        "MEANANOM0": "19.4", "ROTSPD": str(1 / 1.0275)})
    ndpy = max(math.floor(686.98 / 1.0275 / 12 + 0.5), 1) * 12
    files["plasim_namelist"] = _apply(files["plasim_namelist"], {
        "FIXEDLON": "0", "N_DAYS_PER_YEAR": str(ndpy), "PSURF": "610", "NVEG": "0", "NEQSIG": "0",
        "MPSTEP": "30", "NSTPW": "240", "FILTERKAPPA": "8", "NFILTEREXP": "8", "LANDHOSKN0": "15",
        "N_START_YEAR": "0", "NFILTER": "3", "KICK": "2", "SEED": "5,6,7,8,9,10,11,12"})
    files["radmod_namelist"] = _apply(files["radmod_namelist"], {
        "NFIXED": "0", "NSTARTEMP": "1", "STARBBTEMP": "5772", "NO3": "0", "CO2": "950000"})
    files["landmod_namelist"] = _apply(files["landmod_namelist"], {
        "NWETSOIL": "1", "DSOILZ": ",".join(str(d * 2) for d in (0.4, 0.8, 1.6, 3.2, 6.4)), "WSMAX": "0.1",
        "ALBLAND": "0.25", "DGROUNDALB": "0.25,0.25"})
    files["oceanmod_namelist"] = _apply(files["oceanmod_namelist"], {"MLDEPTH": "10"})
    files["vegmod_namelist"] = _apply(files["vegmod_namelist"], {"RINIDAGG": "0.1"})
    files["glacier_namelist"] = _apply(files["glacier_namelist"], {"NGLACIER": "1", "GLACELIM": "3", "ICESHEETH": "50"})
    files["hurricane_namelist"] = _apply(files["hurricane_namelist"], {"NSTORMDIAG": "1"})
    files["surfmod_namelist"] = _apply(files["surfmod_namelist"], {"NSPINIT": "1"})
    for name, values in edits.items():
        files[name] = _apply(files[name], values)
    return files


def test_config_from_namelists_fields(exo):
    """Each field comes from the key edits writes it into, inverted."""
    c = exo.config_from_namelists(42, 5, _planet_files())
    assert (c.truncation, c.levels) == (42, 5)
    assert c.gravity == 3.71 and c.radius_earth * 6371220 == 3389500 and c.flux_wm2 == 586
    assert (c.eccentricity, c.obliquity_deg, c.fixed_orbit, c.keplerian) == (0.0934, 25.19, 1, 0)
    assert c.gas_constant == 188.9 and c.year_days * 86400 == 686.98 * 86400
    assert 1 / c.rotation_period_days == 1 / 1.0275
    assert (c.periapse_longitude_deg, c.initial_mean_anomaly_deg) == (336.1, 19.4)
    assert c.pressure_bar * 1e5 == 610 and c.substellar_longitude_deg == 0
    assert (c.vegetation, c.timestep_minutes, c.filter, c.kick) == (0, 30, 3, 2)
    assert list(c.seed) == [5, 6, 7, 8, 9, 10, 11, 12]
    assert (c.synchronous, c.star_temperature_k, c.ozone, c.co2_ppm) == (0, 5772, 0, 950000)
    assert (c.wet_soil, c.soil_depth_scale, c.soil_water_capacity_m) == (1, 2, 0.1)
    assert math.isnan(c.soil_heat_capacity) and math.isnan(c.ocean_albedo) and math.isnan(c.snow_ice_albedo)
    assert c.land_albedo == 0.25
    assert (c.ocean_depth_m, c.initial_growth) == (10, 0.1)
    assert (c.glaciers, c.glacier_min_depth_m, c.glacier_initial_height_m) == (1, 3, 50)
    assert (c.storm_diagnostics, c.nspinit) == (1, 1)
    # fields no namelist carries keep the caller's values
    base = exo.default_config(land_fraction=0.17, initial_state=1)
    c2 = exo.config_from_namelists(42, 5, _planet_files(), base=base)
    assert (c2.land_fraction, c2.initial_state) == (0.17, 1)
    # a base periapse (102.7, p_earth's) is the null periapse
    c3 = exo.config_from_namelists(21, 10, _planet_files(planet_namelist={"MVELP": "102.7"}))
    assert math.isnan(c3.periapse_longitude_deg)


def test_module_defaults_for_absent_keys(exo):
    """A key no file sets takes its module's default: T21's 45-minute step
    for MPSTEP 0 (readnl), p_earth's gravity and radius, 360 ppm CO2, a
    50 m mixed layer, the star's spectrum off."""
    files = _planet_files()
    for key in ("GA", "PLARAD"):
        files["planet_namelist"] = "\n".join(l for l in files["planet_namelist"].split("\n")
                                             if l.split("=")[0].strip() != key)
    files["plasim_namelist"] = _apply(files["plasim_namelist"], {"MPSTEP": "0", "NSTPW": "160"})
    files["radmod_namelist"] = " &radmod_nl\n /END\n"
    files.pop("oceanmod_namelist")
    c = exo.config_from_namelists(21, 10, files)
    assert (c.gravity, c.radius_earth, c.timestep_minutes) == (9.80665, 1, 45)
    assert (c.co2_ppm, c.ozone, c.star_temperature_k, c.ocean_depth_m) == (360, 1, 0, 50)


@pytest.mark.parametrize("file,edit,named", [
    ("plasim_namelist", {"NKITS": "2"}, "NKITS"),
    ("plasim_namelist", {"NGPTFILTER": "0"}, "NGPTFILTER"),
    ("plasim_namelist", {"NLOWIO": "0"}, "NLOWIO"),
    ("plasim_namelist", {"NSTPW": "100"}, "NSTPW"),
    ("plasim_namelist", {"NDEL": "2,2,2,2,4"}, "NDEL"),
    ("plasim_namelist", {"NFILTER": "4"}, "NFILTER"),
    ("plasim_namelist", {"KICK": "5"}, "KICK"),
    ("plasim_namelist", {"N_DAYS_PER_YEAR": "360"}, "N_DAYS_PER_YEAR"),
    ("plasim_namelist", {"SEED": "1.5"}, "SEED"),
    ("radmod_namelist", {"TSWR1": "0.089"}, "TSWR1"),
    ("rainmod_namelist", {"NCLOUDS": "1", "GAMMA": "0.02"}, "GAMMA"),
    ("landmod_namelist", {"DSNOWALBMX": "0.7,0.7"}, "DSNOWALBMX"),
    ("landmod_namelist", {"DGROUNDALB": "0.25,0.3"}, "DGROUNDALB"),
    ("landmod_namelist", {"DSOILZ": "0.4,0.8,1.6,3.2,7"}, "DSOILZ"),
    ("seamod_namelist", {"ALBICE": "0.5"}, "ALBICE"),
    ("carbonmod_namelist", {"WMAX": "1"}, "WMAX"),
    ("surfmod_namelist", {"NOROMAX": "10"}, "NOROMAX"),
    ("hurricane_namelist", {"CKCD": "1.2"}, "CKCD"),
])
def test_unsupported_keys_are_refused_by_name(exo, file, edit, named):
    files = _planet_files()
    files[file] = _apply(files.get(file, f" &{file[:-9]}_nl\n /END\n"), edit)
    with pytest.raises(exo.ExoError, match=named):
        exo.config_from_namelists(42, 5, files)


def test_base_namelists_alone_are_not_a_run(exo):
    """The base files without a config's edits leave p_earth's sidereal
    year under a 360-day calendar, a year this engine does not run."""
    with pytest.raises(exo.ExoError, match="SIDEREAL_YEAR"):
        exo.config_from_namelists(21, 10, BASE)


RANDOM_CONFIGS = r"""
const N = require('./namelist.js')
const {engineConfig, namelistConfig} = require('./runner.js')
const base = JSON.parse(require('fs').readFileSync(process.argv[1], 'utf8')).files
let s = 20261204
const rnd = () => { s = (s * 1103515245 + 12345) % 2147483648; return s / 2147483648 }
const pick = (a) => a[Math.floor(rnd() * a.length)]
const maybe = (f) => (rnd() < 0.5 ? null : f())
const round = (x, d) => Number(x.toFixed(d))
const out = []
const texts = []
for (let i = 0; i < 300; i++) {
  const gases = {H2: null, He: null, N2: round(rnd() * 3, 4), O2: maybe(() => round(rnd(), 3)), Ar: null, Ne: null,
                 Kr: null, CO2: round(rnd() * 0.01, 6), H2O: null}
  const cfg = {
    targetTemperatureK: null, stormclim: rnd() < 0.5,
    atmosphere: rnd() < 0.7
      ? {mode: 'partialPressures', partialPressures: gases, pressureBar: null, gasConstant: null, ozone: rnd() < 0.5}
      : {mode: 'pressureAndGasConstant', partialPressures: gases, pressureBar: round(0.1 + rnd() * 5, 5),
         gasConstant: round(150 + rnd() * 3000, 3), ozone: rnd() < 0.5},
    layers: pick([5, 10]), periapseLongitudeDeg: maybe(() => round(rnd() * 360, 2)), keplerian: rnd() < 0.3,
    initialMeanAnomalyDeg: maybe(() => round(rnd() * 720 - 360, 2)),
    soilWaterCapacityM: maybe(() => round(0.05 + rnd(), 3)), soilHeatCapacity: maybe(() => round(1e6 + rnd() * 3e6, 0)),
    soilDepthScale: pick([1, 0.5, 1.5, round(0.1 + rnd() * 3, 3)]),
    landAlbedo: maybe(() => round(rnd(), 3)), oceanAlbedo: maybe(() => round(rnd(), 3)),
    snowIceAlbedo: maybe(() => round(rnd(), 3)), radiusEarth: round(0.3 + rnd() * 2, 4),
    gravityMs2: round(2 + rnd() * 20, 4), synchronous: rnd() < 0.3, substellarLongitudeDeg: round(rnd() * 360, 1),
    oceanDepthM: round(1 + rnd() * 200, 1), fluxWm2: round(100 + rnd() * 3000, 2),
    starTemperatureK: pick([0, 5780, round(2600 + rnd() * 5000, 0)]),
    yearDays: pick([360, 365, round(50 + rnd() * 700, 2)]), eccentricity: round(rnd() * 0.3, 4),
    rotationPeriodDays: pick([1, 1, round(0.3 + rnd() * 20, 3)]), obliquityDeg: round(rnd() * 90, 2),
    fixedOrbit: true, wetSoil: rnd() < 0.3, vegetation: pick([0, 1, 2]), initialGrowth: round(rnd(), 3),
    glaciers: {enabled: rnd() < 0.5, minDepthM: round(rnd() * 5, 2), initialHeightM: pick([-1, round(rnd() * 500, 1)])},
    physicsFilter: pick([['gp', 'exp', 'sp'], ['gp', 'sp'], ['none'], ['cesaro'], ['gp', 'lh', 'sp'], []]),
    timestepMinutes: pick([45, 30, 20, 15, 60]), truncation: pick([21, 42, 63]), earthSurfaceClimatology: false,
  }
  if (cfg.rotationPeriodDays === 1 && cfg.yearDays !== Math.round(cfg.yearDays)) cfg.yearDays = Math.round(cfg.yearDays)
  let direct, viaFiles
  try { direct = engineConfig(cfg, {}) } catch (x) { out.push({i, skipped: x.message}); continue }
  viaFiles = namelistConfig(cfg, {}, base)
  const nan = (o) => { for (const k of Object.keys(o)) if (typeof o[k] === 'number' && Number.isNaN(o[k])) o[k] = 'NaN'; return o }
  out.push({i, direct: nan(direct), viaFiles: nan(viaFiles)})
  if (i < 3) texts.push(N.files(base, cfg))
}
console.log(JSON.stringify({out, texts}))
"""


@needs_node
def test_edits_over_the_base_read_back_to_the_direct_mapping(exo, tmp_path):
    """300 seeded GcmConfigs (every field varied): node/namelist.js's edits
    applied over the base files and read back by exo_config_from_namelists
    give the exo_config of runner.js's direct mapping. The fields the
    set-up uses only through a product (PSURF = bar x 1e5, PLARAD, the
    sidereal year, 1/ROTSPD, the DSOILZ layers) agree in that product;
    every other field is equal."""
    base_file = tmp_path / "base.json"
    base_file.write_text(json.dumps({"files": BASE}))
    proc = subprocess.run(["node", "-e", RANDOM_CONFIGS, str(base_file)], capture_output=True, text=True,
                          timeout=300, cwd=NODE_DIR)
    assert proc.returncode == 0, proc.stderr
    result = json.loads(proc.stdout.strip().splitlines()[-1])
    compared = 0
    products = {"pressure_bar": 1e5, "radius_earth": 6371220, "year_days": 86400, "soil_depth_scale": 0.4}
    for case in result["out"]:
        if "skipped" in case:
            continue
        d, v = case["direct"], case["viaFiles"]
        for key in set(d) | set(v):
            if key not in d:  # fields engineConfig leaves at the defaults
                continue
            if key in products:
                assert d[key] * products[key] == v[key] * products[key], (case["i"], key, d[key], v[key])
            elif key == "rotation_period_days":
                assert 1 / d[key] == 1 / v[key], (case["i"], key)
            else:
                assert d[key] == v[key], (case["i"], key, d[key], v[key])
        compared += 1
    assert compared >= 250
    # the files a config writes are the base files, edited: the base's
    # other lines survive (KICK, SEED, NLOWIO) and each edit is one line
    plasim = result["texts"][0]["plasim_namelist"]
    for line in (" KICK        =     1", " SEED = 14963991,2,3,4,5,6,7,8", " NLOWIO      =     1"):
        assert line in plasim.split("\n")
    assert sum(l.split("=")[0].strip() == "MPSTEP" for l in plasim.split("\n")) == 1


@needs_node
def test_namelist_route_sets_up_the_same_planet(exo, tmp_path):
    """The set-up of a configuration reached through its namelist files
    equals the set-up of the direct mapping: every constant of every
    start-up phase and the initial state, bit for bit."""
    base_file = tmp_path / "base.json"
    base_file.write_text(json.dumps({"files": BASE}))
    result = json.loads(subprocess.run(["node", "-e", RANDOM_CONFIGS, str(base_file)], capture_output=True,
                                       text=True, timeout=300, cwd=NODE_DIR).stdout.strip().splitlines()[-1])
    cases = [c for c in result["out"] if "skipped" not in c][:4]
    names = [f[0] for f in exo.ExoConfig._fields_]

    def to_cfg(obj):
        cfg = exo.default_config()
        for k, val in obj.items():
            if k == "seed":
                for i, w in enumerate(val):
                    cfg.seed[i] = w
            elif k in names:
                setattr(cfg, k, float("nan") if val == "NaN" else val)
        return cfg

    with exo.Engine(-1, 2 * len(cases)) as e:
        for case in cases:
            a = e.add_member(to_cfg(case["direct"]))
            b = e.add_member(to_cfg(case["viaFiles"]))
            for phase in range(4):
                ka, kb = e.constants(a, phase), e.constants(b, phase)
                assert ka == kb or all(ka[k] == kb[k] or (np.isnan(ka[k]) and np.isnan(kb[k])) for k in ka), \
                    (case["i"], phase)
            for table in ("initial", "srp", "sop", "legendre_qc", "sid"):
                assert np.array_equal(e.read_table(a, table), e.read_table(b, table), equal_nan=True), \
                    (case["i"], table)
