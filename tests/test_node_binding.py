"""The Node-API binding (web-exoplasim_amd/node): it loads, exports the
engine calls, maps configs and errors as the C-ABI does, and GcmRun
drives a GPU run through snapshot, checkpoint and stop."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "web-exoplasim_amd", "node")
ADDON = os.path.join(NODE_DIR, "exoplasim_napi.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the built addon is absent")


def _node(script, timeout=120):
    proc = subprocess.run(["node", "-e", script], capture_output=True, text=True, timeout=timeout,
                          cwd=NODE_DIR)
    assert proc.returncode == 0, proc.stderr
    return json.loads(proc.stdout.strip().splitlines()[-1])


def test_addon_exports_and_host_engine():
    out = _node("""
const a = require('./exoplasim_napi.node')
const e = a.create(-1, 2)
const id = a.addMember(e, {truncation: 42, levels: 10, land_fraction: 0.3}, a.PRECISION_FP32)
const errors = []
for (const bad of [{truncation: 22}, {bogus: 1}, {levels: 'x'}]) {
  try { a.addMember(e, bad, 0) } catch (x) { errors.push(x.message) }
}
try { a.step(e, 1) } catch (x) { errors.push('step: ' + x.message) }
const r = {keys: Object.keys(a), dims: a.dims(e, id), date: a.memberDate(e, id),
           defaults: a.configDefault(), errors}
a.destroy(e)
console.log(JSON.stringify(r))
""")
    for name in ("create", "addMember", "step", "sync", "readSnapshot", "readField", "globalMeans",
                 "checkpoint", "restore", "memberDate"):
        assert name in out["keys"]
    assert out["dims"] == [42, 64, 128, 10]
    assert out["date"] == {"year": 1, "month": 1, "day": 1, "hour": 0, "minute": 0}
    assert out["defaults"]["truncation"] == 21
    assert out["errors"][0] == "T22 is not built"
    assert out["errors"][1] == "unknown config field bogus"
    assert out["errors"][2] == "levels must be a number"
    assert len(out["errors"]) == 4


# GCM.DEFAULT_CONFIG (src/model/climate/gcm/index.ts:4-44) with
# GCM_ATMOSPHERE.DEFAULT_CONFIG (atmosphere/index.ts:66-81), as JSON.
DEFAULT_CONFIG = {
    "targetTemperatureK": None, "stormclim": False,
    "atmosphere": {"mode": "partialPressures",
                   "partialPressures": {g: None for g in ("H2", "He", "N2", "O2", "Ar", "Ne", "Kr", "CO2", "H2O")},
                   "pressureBar": None, "gasConstant": None, "ozone": True},
    "layers": 10, "periapseLongitudeDeg": None, "keplerian": False, "initialMeanAnomalyDeg": None,
    "soilWaterCapacityM": None, "soilHeatCapacity": None, "soilDepthScale": 1, "landAlbedo": None,
    "oceanAlbedo": None, "snowIceAlbedo": None, "radiusEarth": 1, "gravityMs2": 9.80665, "synchronous": False,
    "substellarLongitudeDeg": 180, "oceanDepthM": 50, "fluxWm2": 1367, "starTemperatureK": 5780, "yearDays": 360,
    "eccentricity": 0.016715, "rotationPeriodDays": 1, "obliquityDeg": 23.5, "fixedOrbit": True, "wetSoil": False,
    "vegetation": 2, "initialGrowth": 0.5, "glaciers": {"enabled": True, "minDepthM": 2, "initialHeightM": -1},
    "physicsFilter": ["gp", "exp", "sp"], "timestepMinutes": 45, "truncation": 21,
    "earthSurfaceClimatology": False,
}


def test_engine_config_mapping():
    """The reference's default GcmConfig maps onto every exo_config field
    it implies, the atmosphere resolved as GCM_ATMOSPHERE.resolve does
    (1.011 bar of Earth air, R = 8314.46 / 28.970253, 300 ppm CO2)."""
    out = _node(f"""
const {{engineConfig}} = require('./runner.js')
const c = engineConfig({json.dumps(DEFAULT_CONFIG)})
for (const k of Object.keys(c)) if (Number.isNaN(c[k])) c[k] = 'NaN'
const edited = engineConfig(Object.assign({json.dumps(DEFAULT_CONFIG)}, {{
  truncation: 63, layers: 5, physicsFilter: ['gp', 'lh', 'sp'], landAlbedo: 0.3, vegetation: 0,
  atmosphere: {{mode: 'partialPressures', partialPressures: {{N2: 0.79, O2: 0.21, CO2: 0.0004}}, ozone: false}}}}),
  {{landFraction: 0.2, terrainSeed: 9}})
console.log(JSON.stringify({{c, edited}}))
""")
    c = out["c"]
    assert c["pressure_bar"] == 1.011 and c["co2_ppm"] == 300
    assert abs(c["gas_constant"] - 8314.46261815324 / 28.970253) < 1e-9
    assert (c["vegetation"], c["glaciers"], c["glacier_min_depth_m"], c["glacier_initial_height_m"]) == (2, 1, 2, -1)
    assert (c["star_temperature_k"], c["filter"], c["fixed_orbit"], c["synchronous"]) == (5780, 2, 1, 0)
    assert (c["eccentricity"], c["obliquity_deg"], c["substellar_longitude_deg"]) == (0.016715, 23.5, 180)
    for k in ("periapse_longitude_deg", "initial_mean_anomaly_deg", "soil_water_capacity_m", "land_albedo",
              "ocean_albedo", "snow_ice_albedo", "soil_heat_capacity"):
        assert c[k] == "NaN", k
    e = out["edited"]
    assert (e["truncation"], e["levels"], e["filter"], e["land_albedo"], e["vegetation"], e["ozone"]) == \
        (63, 5, 3, 0.3, 0, 0)
    assert abs(e["pressure_bar"] - 1.0004) < 1e-12 and abs(e["co2_ppm"] - 0.0004 / 1.0004 * 1e6) < 1e-6
    assert (e["land_fraction"], e["terrain_seed"]) == (0.2, 9)


def test_engine_config_refuses_what_it_cannot_honour():
    out = _node(f"""
const {{engineConfig, start}} = require('./runner.js')
const base = {json.dumps(DEFAULT_CONFIG)}
const errors = []
for (const edit of [{{targetTemperatureK: 288}}, {{fixedOrbit: false}}, {{bogus: 1}}, {{physicsFilter: ['gauss']}},
                    {{atmosphere: {{mode: 'x', partialPressures: {{}}}}}},
                    {{atmosphere: {{mode: 'partialPressures', partialPressures: {{N2: -1}}}}}}]) {{
  try {{ engineConfig(Object.assign({{}}, base, edit)); errors.push(null) }} catch (x) {{ errors.push(x.message) }}
}}
const resolution = {{layers: 10, truncation: 42, latitudes: 64, cpu: null, gpu: {{inputs: 't42/inputs.json'}}}}
for (const request of [{{config: base, resolution}}, {{config: Object.assign({{}}, base, {{truncation: 42}}), resolution, engine: 'cpu'}}]) {{
  try {{ start(request); errors.push(null) }} catch (x) {{ errors.push(x.message) }}
}}
console.log(JSON.stringify(errors))
""")
    assert out == [
        "The targeting runner (targetTemperatureK) is not built for the MI355X engine",
        "A free orbit (fixedOrbit false, Berger's series) is not built for the MI355X engine",
        "Unknown GcmConfig field bogus",
        "Unknown physics filter gauss",
        "Choose an atmosphere input mode.",
        "Nitrogen (N₂) partial pressure must be finite and nonnegative.",
        "Model assets do not match the configured resolution and layers",
        "T42 is not built for the CPU engine",
    ]


@pytest.mark.gpu
def test_gcm_run_from_the_reference_defaults_and_a_terrain():
    """GcmRun started as the page starts it: GCM.DEFAULT_CONFIG, a terrain
    raster (the land mask and orography as GCM_BOUNDARY.files makes them)
    and no Earth climatology. It runs the reference's atmosphere (1.011
    bar of air: the surface-pressure plane's mean over sea stays near
    101100 Pa) with vegetation and glaciers on, and its ocean starts at the
    freezing point under ice, as the reference's does without an SST
    file."""
    out = _node(f"""
const {{start}} = require('./runner.js')
const w = 360, h = 180
const elevationM = new Float32Array(w * h), land = new Float32Array(w * h)
for (let y = 0; y < h; y++) for (let x = 0; x < w; x++) {{
  const lat = 90 - (y + 0.5), lon = -180 + (x + 0.5)
  const inside = Math.abs(lat) < 40 && lon > -30 && lon < 60
  land[y * w + x] = inside ? 1 : 0
  elevationM[y * w + x] = inside ? 800 * Math.cos(lat * Math.PI / 90) : 0
}}
const snaps = []
const resolution = {{layers: 10, truncation: 21, latitudes: 32, cpu: null, gpu: {{inputs: 't21/inputs.json'}}}}
const run = start({{config: {json.dumps(DEFAULT_CONFIG)}, resolution, engine: 'gpu',
  terrain: {{name: 'box', width: w, height: h, elevationM, land}}, publishEvery: 8, restart: null,
  startPaused: false,
  onStatus: (u) => {{ if (u.status === 'failed') console.log(JSON.stringify({{failed: u.message}})) }},
  onDevice: () => {{}},
  onSnapshot: (s) => {{
    const nh = 2048, v = s.values
    const mask = v.subarray(15 * nh, 16 * nh), ps = v.subarray(6 * nh, 7 * nh), ice = v.subarray(24 * nh, 25 * nh)
    let sea = 0, psum = 0, iced = 0
    for (let i = 0; i < nh; i++) if (mask[i] < 0.5) {{ sea++; psum += ps[i]; iced += ice[i] }}
    snaps.push({{finite: v.every(Number.isFinite), sea, ps: psum / sea, ice: iced / sea,
                 land: nh - sea, ts: v[0]}})
    if (snaps.length === 3) run.stop()
  }},
  onCheckpoint: () => {{}}}})
setTimeout(() => console.log(JSON.stringify({{snaps}})), 60000).unref()
process.on('beforeExit', () => console.log(JSON.stringify({{snaps}})))
""", timeout=300)
    assert "failed" not in out, out
    snaps = out["snaps"]
    assert len(snaps) == 3 and all(s["finite"] for s in snaps)
    assert 0 < snaps[0]["land"] < 2048
    assert all(abs(s["ps"] - 101100.0) < 1500.0 for s in snaps), snaps
    # the whole ocean starts under sea ice about a metre thick (the
    # snapshot's seaIceThickness); its cover follows the ice model's
    # concentration, which the reference's set-up leaves at 0
    assert snaps[0]["ice"] > 0.5, snaps


@pytest.mark.gpu
def test_gcm_run_snapshots_checkpoint_stop():
    out = _node("""
const {start} = require('./runner.js')
const statuses = [], snaps = []
let restart = null
const run = start({config: {truncation: 21, layers: 10, timestepMinutes: 45}, publishEvery: 4,
  restart: null, startPaused: false,
  onStatus: (u) => { statuses.push(u.status); if (u.status === 'finished' || u.status === 'failed') done(u) },
  onDevice: () => {},
  onSnapshot: (s) => { snaps.push({step: s.step, n: s.values.length, finite: s.values.every(Number.isFinite),
                                   date: s.date}); if (snaps.length === 3) run.checkpoint() },
  onCheckpoint: (c) => { restart = c.restart.byteLength; run.stop() }})
function done(u) {
  console.log(JSON.stringify({statuses, snaps, restart, message: u.message}))
}
""", timeout=300)
    assert out["statuses"][:2] == ["loading", "running"], out
    assert out["statuses"][-1] == "finished", out
    assert "checkpointing" in out["statuses"]
    assert len(out["snaps"]) >= 3
    assert all(s["finite"] for s in out["snaps"])
    assert out["snaps"][0]["n"] == 101 * 2048
    assert out["restart"] > 2429594 * 4


@pytest.mark.gpu
def test_gcm_run_averages_and_averager_checkpoint():
    """onAverages reports newly closed months and the yearly balances
    (gcm.worker.ts:324-341); the checkpoint's averager restarts a run
    that carries on averaging from it."""
    out = _node("""
const {start} = require('./runner.js')
const config = {truncation: 21, layers: 5, timestepMinutes: 45}
const updates = []
let first = null
const summary = (a) => ({months: a.months.map((m) => [m.year, m.month, m.snapshot.step, m.snapshot.values.length]),
  current: a.current && {year: a.current.year, month: a.current.month, samples: a.current.samples,
                         sum: a.current.sum.length, f64: a.current.sum instanceof Float64Array,
                         step: a.current.last.step},
  years: a.years.length})
const run1 = start({config, publishEvery: 32, restart: null, startPaused: false,
  onStatus: (u) => { if (u.status === 'failed') { console.log(JSON.stringify({failed: u.message})) } },
  onDevice: () => {}, onSnapshot: () => {},
  onAverages: (u) => {
    updates.push({months: u.months.map((m) => [m.year, m.month, m.snapshot.step, m.snapshot.date.day]),
                  years: u.years.length})
    if (updates.length === 2) run1.checkpoint()
  },
  onCheckpoint: (c) => { first = c; run1.stop(); second(c) }})
function second(saved) {
  const run2 = start({config, publishEvery: 32, restart: saved.restart, averager: saved.averager,
    startPaused: false,
    onStatus: (u) => { if (u.status === 'failed') { console.log(JSON.stringify({failed: u.message})) } },
    onDevice: () => {}, onAverages: () => {},
    onSnapshot: () => run2.checkpoint(),
    onCheckpoint: (c) => {
      run2.stop()
      console.log(JSON.stringify({updates, first: summary(first.averager), second: summary(c.averager)}))
    }})
}
""", timeout=300)
    assert "failed" not in out, out
    updates = out["updates"]
    assert [len(u["months"]) for u in updates] == [1, 1]
    (y1, m1, s1, d1), (y2, m2, s2, d2) = updates[0]["months"][0], updates[1]["months"][0]
    assert (y1, m1 + 1) == (y2, m2) and s2 > s1 and d1 == d2 == 1
    first, second = out["first"], out["second"]
    assert [m[:3] for m in first["months"]] == [m[:3] for m in updates[0]["months"] + updates[1]["months"]]
    assert all(m[3] == 76 * 2048 for m in first["months"])
    cur = first["current"]
    assert cur["f64"] and cur["sum"] == 76 * 2048 and cur["samples"] >= 1
    # one more snapshot after the restart: added to the open month, or it closes that month
    if second["current"]["month"] == cur["month"]:
        assert second["months"] == first["months"]
        assert second["current"]["samples"] == cur["samples"] + 1
    else:
        assert second["months"][:-1] == first["months"][-len(second["months"]) + 1:]
        assert second["months"][-1][:2] == [cur["year"], cur["month"]]


def test_boundary_sra_through_node(exo):
    """The .sra values are formatted as Node's own Number.toFixed(3) does,
    over random values and exact binary ties (k / 2000 with k odd and a
    power-of-two denominator, e.g. 0.0625)."""
    out = _node("""
const {addon} = require('./runner.js')
const v = new Float64Array(2 * 32 * 32)
let seed = 7
const rand = () => (seed = (seed * 1103515245 + 12345) % 2147483648) / 2147483648
for (let i = 0; i < v.length; i++) v[i] = (rand() - 0.5) * Math.pow(10, Math.floor(rand() * 8) - 3)
const ties = [0.0625, -0.0625, 0.1875, 2.0625, -1024.0625, 1.0005, -0, 0]
ties.forEach((x, i) => { v[i] = x })
const text = addon.boundarySra(129, 32, v)
const mine = text.split('\\n').slice(1).join(' ').trim().split(/ +/)
const js = Array.from(v, (x) => x.toFixed(3))
const bad = js.map((s, i) => s === mine[i] ? -1 : i).filter((i) => i >= 0)
console.log(JSON.stringify({bad: bad.slice(0, 5), n: js.length, first: mine.slice(0, 8)}))
""")
    assert out["n"] == 2048 and out["bad"] == [], out
    assert out["first"] == ["0.063", "-0.063", "0.188", "2.063", "-1024.063", "1.000", "0.000", "0.000"]
