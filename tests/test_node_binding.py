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


def test_engine_config_mapping():
    out = _node("""
const {engineConfig} = require('./runner.js')
const c = engineConfig({truncation: 63, layers: 5, timestepMinutes: 30, radiusEarth: 1.5, gravityMs2: 9.0,
  rotationPeriodDays: 2, fluxWm2: 1000, obliquityDeg: 10, eccentricity: 0.1, yearDays: 200, oceanDepthM: 20,
  atmosphere: {ozone: false, pressureBar: 2, gasConstant: null}}, {landFraction: 0.2, terrainSeed: 9})
console.log(JSON.stringify(c))
""")
    assert out == {"truncation": 63, "levels": 5, "timestep_minutes": 30, "radius_earth": 1.5, "gravity": 9.0,
                   "rotation_period_days": 2, "flux_wm2": 1000, "obliquity_deg": 10, "eccentricity": 0.1,
                   "year_days": 200, "ocean_depth_m": 20, "ozone": 0, "pressure_bar": 2, "land_fraction": 0.2,
                   "terrain_seed": 9}


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
