"use strict"
// GcmRun over the MI355X engine: the shape of GCM_RUNNER.start /
// GCM_WEBGPU_RUNNER.start (src/model/climate/gcm/runner/index.ts:61-78,
// runner/types.ts:47-74). A request carries the config, publishEvery, an
// optional restart and the callbacks; the returned handle has pause,
// resume, checkpoint and stop. Status updates use the reference's
// statuses (loading, running, paused, checkpointing, finished, failed)
// and its error wording (gcm.worker.ts:47-51, :317-323).
//
// The engine runs on the calling thread: each slice enqueues publishEvery
// steps, waits for them, publishes one snapshot, then yields to the event
// loop (setImmediate), so pause/stop/checkpoint land between slices.
//
// Averages follow gcm.worker.ts:197 and :324-364: every published snapshot
// goes into the member's monthly means (on the device); a slice that closes
// a month reports the newly closed months and the yearly balances through
// onAverages; a checkpoint carries the averager (GcmAverager: the open
// month, the closed months, the yearly balances), and request.averager
// restores it.

const fs = require("fs")
const path = require("path")
const addon = require(path.join(__dirname, "exoplasim_napi.node"))
const NAMELIST = require(path.join(__dirname, "namelist.js"))

// GcmConfig's fields (src/model/climate/gcm/types.ts:28-62).
const CONFIG_KEYS = [
	"targetTemperatureK", "stormclim", "atmosphere", "layers", "periapseLongitudeDeg", "keplerian",
	"initialMeanAnomalyDeg", "soilWaterCapacityM", "soilHeatCapacity", "soilDepthScale", "landAlbedo",
	"oceanAlbedo", "snowIceAlbedo", "radiusEarth", "gravityMs2", "synchronous", "substellarLongitudeDeg",
	"oceanDepthM", "fluxWm2", "starTemperatureK", "yearDays", "eccentricity", "rotationPeriodDays",
	"obliquityDeg", "fixedOrbit", "wetSoil", "vegetation", "initialGrowth", "glaciers", "physicsFilter",
	"timestepMinutes", "truncation", "earthSurfaceClimatology",
]
// GCM_ATMOSPHERE.GASES order (atmosphere/index.ts:14-64).
const GASES = ["H2", "He", "N2", "O2", "Ar", "Ne", "Kr", "CO2", "H2O"]
// physicsFilter's NFILTER switches (namelist/index.ts:27-34); "gp" and
// "sp" set NGPTFILTER / NSPVFILTER, which the base namelist already sets.
const NFILTER = { none: 0, cesaro: 1, exp: 2, lh: 3 }

function nullable(value) {
	return value === null || value === undefined ? NaN : value
}

function flag(value) {
	return value ? 1 : 0
}

// GcmConfig (types.ts:28-62) onto exo_config, as GCM_NAMELIST.edits
// writes it into the namelists (namelist/index.ts:40-182), the
// atmosphere through GCM_ATMOSPHERE.resolve. Every field is honoured or
// refused: an unknown field, a targeting run and a free (Berger) orbit
// throw. `planet` (optional) carries the synthetic-planet fields of the
// seeded terrain (landFraction, terrainMaxM, seeds, seaSurfaceK, ...),
// which the reference takes from terrain rasters instead.
function engineConfig(config, planet) {
	for (const key of Object.keys(config))
		if (!CONFIG_KEYS.includes(key)) throw new Error(`Unknown GcmConfig field ${key}`)
	if (config.targetTemperatureK !== null && config.targetTemperatureK !== undefined)
		throw new Error("The targeting runner (targetTemperatureK) is not built for the MI355X engine")
	if (config.fixedOrbit === false)
		throw new Error("A free orbit (fixedOrbit false, Berger's series) is not built for the MI355X engine")
	const out = {
		truncation: config.truncation,
		levels: config.layers,
		timestep_minutes: config.timestepMinutes,
		radius_earth: config.radiusEarth,
		gravity: config.gravityMs2,
		rotation_period_days: config.rotationPeriodDays,
		flux_wm2: config.fluxWm2,
		obliquity_deg: config.obliquityDeg,
		eccentricity: config.eccentricity,
		year_days: config.yearDays,
		ocean_depth_m: config.oceanDepthM,
		storm_diagnostics: config.stormclim === undefined ? undefined : flag(config.stormclim),
		vegetation: config.vegetation,
		initial_growth: config.initialGrowth,
		star_temperature_k: config.starTemperatureK,
		synchronous: config.synchronous === undefined ? undefined : flag(config.synchronous),
		substellar_longitude_deg: config.substellarLongitudeDeg,
		fixed_orbit: config.fixedOrbit === undefined ? undefined : flag(config.fixedOrbit),
		keplerian: config.keplerian === undefined ? undefined : flag(config.keplerian),
		soil_depth_scale: config.soilDepthScale,
		wet_soil: config.wetSoil === undefined ? undefined : flag(config.wetSoil),
	}
	for (const [key, name] of [
		["periapseLongitudeDeg", "periapse_longitude_deg"],
		["initialMeanAnomalyDeg", "initial_mean_anomaly_deg"],
		["soilWaterCapacityM", "soil_water_capacity_m"],
		["soilHeatCapacity", "soil_heat_capacity"],
		["landAlbedo", "land_albedo"],
		["oceanAlbedo", "ocean_albedo"],
		["snowIceAlbedo", "snow_ice_albedo"],
	])
		if (key in config) out[name] = nullable(config[key])
	if (config.glaciers) {
		out.glaciers = flag(config.glaciers.enabled)
		out.glacier_min_depth_m = config.glaciers.minDepthM
		out.glacier_initial_height_m = config.glaciers.initialHeightM
	}
	if (config.physicsFilter) {
		let nfilter = 0
		for (const filter of config.physicsFilter) {
			if (filter === "gp" || filter === "sp") continue
			if (!(filter in NFILTER)) throw new Error(`Unknown physics filter ${filter}`)
			nfilter = NFILTER[filter]
		}
		out.filter = nfilter
	}
	const atmosphere = config.atmosphere
	if (atmosphere) {
		const mode = atmosphere.mode
		if (mode !== "partialPressures" && mode !== "pressureAndGasConstant")
			throw new Error("Choose an atmosphere input mode.")
		const partial = GASES.map((gas) => nullable((atmosphere.partialPressures || {})[gas]))
		const resolved = addon.atmosphereResolve(
			partial,
			mode === "pressureAndGasConstant",
			nullable(atmosphere.pressureBar),
			nullable(atmosphere.gasConstant),
		)
		out.pressure_bar = resolved.pressureBar
		out.gas_constant = resolved.gasConstant
		out.co2_ppm = resolved.co2Ppm
		if (atmosphere.ozone !== undefined) out.ozone = flag(atmosphere.ozone)
	}
	const p = planet || {}
	const extra = {
		land_fraction: p.landFraction,
		terrain_max_m: p.terrainMaxM,
		initial_temperature: p.initialTemperatureK,
		initial_humidity: p.initialHumidity,
		noise_kelvin: p.noiseKelvin,
		sea_surface_k: p.seaSurfaceK === null ? NaN : p.seaSurfaceK,
		terrain_seed: p.terrainSeed,
		noise_seed: p.noiseSeed,
		// A fresh start is the reference's prolog: initfd with the base
		// namelist's KICK 1 and SEED (setup/index.ts:434-449); the seeded
		// synthetic state (isothermal, humid, ln(ps) from the orography)
		// only when planet.syntheticStart asks for it.
		initial_state: p.syntheticStart ? 0 : 1,
	}
	if (p.co2Ppm !== undefined) extra.co2_ppm = p.co2Ppm
	const merged = Object.assign(out, extra)
	for (const key of Object.keys(merged)) if (merged[key] === undefined) delete merged[key]
	return merged
}

// The base namelist files of a request (runner/webgpu/index.ts:44-58):
// request.namelists, the namelist files of request.inputs, or public/gcm's
// t<T>/namelists.json under request.publicDir; null when there are none.
function baseNamelists(request) {
	if (request.namelists) return request.namelists
	const inputs = request.inputs && request.inputs.files
	if (inputs && Object.keys(inputs).some((name) => name.endsWith("_namelist"))) return inputs
	if (request.publicDir) {
		const earth = request.config.earthSurfaceClimatology && request.resolution && request.resolution.gpu
		const file = path.join(request.publicDir, earth ? request.resolution.gpu.inputs : `t${request.config.truncation}/namelists.json`)
		if (fs.existsSync(file)) {
			const files = JSON.parse(fs.readFileSync(file, "utf8")).files
			return Object.fromEntries(Object.entries(files).filter(([name]) => name.endsWith("_namelist")))
		}
	}
	return null
}

// A full GcmConfig through its namelist files, as the reference's set-up
// receives it: GCM_NAMELIST.edits applied over the base files, read back
// by exo_config_from_namelists. The fields no namelist carries (the
// synthetic planet's, the kind of fresh start) come from engineConfig.
function namelistConfig(config, planet, base) {
	const direct = engineConfig(config, planet)
	return addon.configFromNamelists(config.truncation, config.layers, NAMELIST.files(base, config), direct)
}

// Every resolution this engine runs (the shape of GCM_RUNNER.resolutions,
// runner/index.ts:12-59): T21, T42 and T63 at 5 and 10 levels, on the GPU
// engine; `inputs` is the resolution's inputs.json under public/gcm.
function resolutions() {
	const out = []
	for (const [truncation, latitudes] of [[21, 32], [42, 64], [63, 96]])
		for (const layers of [5, 10])
			out.push({ layers, truncation, latitudes, cpu: null, gpu: { inputs: `t${truncation}/inputs.json`, downloadBytes: 0 } })
	return out
}

// GCM_INPUTS.keepSurfaceFile (runner/inputs/index.ts:5-27) for this config.
function keepSurfaceFile(name, config) {
	if (!config.earthSurfaceClimatology) return false
	if (config.landAlbedo !== null && config.landAlbedo !== undefined && name.endsWith("0174.sra")) return false
	if (config.soilWaterCapacityM !== null && config.soilWaterCapacityM !== undefined && name.endsWith("0229.sra"))
		return false
	const sheet = config.glaciers ? Number(config.glaciers.initialHeightM) : NaN
	return !(sheet > 0 && ["0174.sra", "1740.sra", "0210.sra", "0232.sra"].some((s) => name.endsWith(s)))
}

// The records of one .sra file (get_surf_array, setup/surface/index.ts:
// 76-107): an eight-integer header, then the grid's values, per month.
function sraField(name, text, latitudes) {
	const code = Number(name.slice(-8, -4))
	const tokens = text.trim().split(/\s+/).map(Number)
	const points = 2 * latitudes * latitudes
	const record = 8 + points
	const months = Math.floor(tokens.length / record)
	const values = new Float64Array(months * points)
	for (let m = 0; m < months; m++) values.set(tokens.slice(m * record + 8, (m + 1) * record), m * points)
	return { code, months, values }
}

// The surface fields of a run (runner/webgpu/index.ts:39-80's files): the
// Earth climatologies kept for the config, then the terrain's land mask
// and orography (GCM_BOUNDARY.files).
function surfaceFields(engine, request) {
	const config = request.config
	const latitudes = { 21: 32, 42: 64, 63: 96 }[config.truncation]
	const files = {}
	if (config.earthSurfaceClimatology) {
		let inputs = request.inputs
		if (!inputs && request.publicDir && request.resolution && request.resolution.gpu)
			inputs = JSON.parse(fs.readFileSync(path.join(request.publicDir, request.resolution.gpu.inputs), "utf8"))
		if (!inputs) throw new Error(`Missing t${config.truncation}/inputs.json in public/gcm`)
		for (const [name, text] of Object.entries(inputs.files))
			if (name.endsWith(".sra") && keepSurfaceFile(name, config)) files[name] = text
	}
	if (request.terrain)
		Object.assign(files, boundaryFiles(engine, request.terrain, latitudes, config.gravityMs2 === undefined ? 9.80665 : config.gravityMs2))
	return Object.entries(files).map(([name, text]) => sraField(name, text, latitudes))
}

// A closed month as GcmMonthlyMean (averages/index.ts:50-69: the month's
// first day as its date, its last snapshot's step).
function monthlyMean(layers, m) {
	return {
		year: m.year,
		month: m.month,
		snapshot: { layers, step: m.step, date: { year: m.year, month: m.month, day: 1, hour: 0, minute: 0 }, values: m.values },
	}
}

// GcmAverager of the member: current (the open month, with its last
// snapshot), months, years.
function readAverager(engine, id, layers, latest) {
	const months = []
	const count = addon.monthCount(engine, id)
	for (let i = 0; i < count; i++) months.push(monthlyMean(layers, addon.readMonth(engine, id, i)))
	const open = addon.readAccumulator(engine, id)
	const current = open && {
		year: open.year,
		month: open.month,
		sum: open.sum,
		samples: open.samples,
		last: latest || { layers, step: open.step, date: addon.memberDate(engine, id), values: addon.readSnapshot(engine, id) },
	}
	return { current, months, years: addon.yearlyBalances(engine, id) }
}

// request.averager (GcmAverager) onto the device.
function loadAverager(engine, id, averager) {
	const months = (averager.months || []).map((m) => ({
		year: m.year,
		month: m.month,
		samples: 0,
		step: m.snapshot.step,
		values: m.snapshot.values,
	}))
	const c = averager.current
	const current = c ? { year: c.year, month: c.month, samples: c.samples, step: c.last.step, sum: c.sum } : null
	addon.loadAverager(engine, id, months, averager.years || [], current)
}

// GCM_RUNNER.start's checks (runner/index.ts:61-78), then the run.
function start(request) {
	if (request.resolution) {
		if (request.config.layers !== request.resolution.layers || request.config.truncation !== request.resolution.truncation)
			throw new Error("Model assets do not match the configured resolution and layers")
		const engine = request.engine || "gpu"
		if (!request.resolution[engine])
			throw new Error(`T${request.resolution.truncation} is not built for the ${engine.toUpperCase()} engine`)
	}
	const state = {
		paused: !!request.startPaused,
		stopped: false,
		checkpoint: false,
		engine: null,
		id: -1,
		lastClosedStep: -1, // step of the newest closed month already reported
		latest: null, // the last published snapshot
	}
	const status = (s, message = "") => request.onStatus && request.onStatus({ status: s, message })
	const fail = (error) => {
		status("failed", error && error.message ? error.message : String(error))
		close()
	}
	const close = () => {
		if (state.engine) addon.destroy(state.engine)
		state.engine = null
	}

	// Months closed since the last report, oldest first, and every yearly
	// balance (gcm.worker.ts:324-341); nothing when no month closed.
	const reportAverages = () => {
		const count = addon.monthCount(state.engine, state.id)
		const fresh = []
		for (let i = count - 1; i >= 0; i--) {
			const info = addon.readMonth(state.engine, state.id, i, false)
			if (info.step <= state.lastClosedStep) break
			fresh.unshift(i)
		}
		if (!fresh.length) return
		const months = fresh.map((i) => monthlyMean(request.config.layers, addon.readMonth(state.engine, state.id, i)))
		state.lastClosedStep = months[months.length - 1].snapshot.step
		request.onAverages && request.onAverages({ months, years: addon.yearlyBalances(state.engine, state.id) })
	}

	const slice = () => {
		if (!state.engine) return
		try {
			if (state.stopped) {
				close()
				status("finished")
				return
			}
			if (state.checkpoint) {
				state.checkpoint = false
				status("checkpointing")
				const restart = addon.checkpoint(state.engine, state.id)
				const averager = readAverager(state.engine, state.id, request.config.layers, state.latest)
				request.onCheckpoint && request.onCheckpoint({ targeting: null, restart, averager })
				status(state.paused ? "paused" : "running")
			}
			if (state.paused) return
			const every = Math.max(1, request.publishEvery | 0)
			addon.step(state.engine, every)
			addon.sync(state.engine)
			const values = addon.readSnapshot(state.engine, state.id)
			state.latest = {
				layers: request.config.layers,
				step: addon.memberStep(state.engine, state.id),
				date: addon.memberDate(state.engine, state.id),
				values,
			}
			request.onSnapshot && request.onSnapshot(state.latest)
			reportAverages()
			setImmediate(slice)
		} catch (error) {
			fail(error)
		}
	}

	setImmediate(() => {
		try {
			status("loading")
			state.engine = addon.create(request.device === undefined ? 0 : request.device, 1)
			const precision = request.precision === "split" ? addon.PRECISION_SPLIT : addon.PRECISION_FP32
			const base = baseNamelists(request)
			const toEngine = (c) => (base ? namelistConfig(c, request.planet, base) : engineConfig(c, request.planet))
			const config = toEngine(request.config)
			const fields = surfaceFields(state.engine, request)
			// without a surface-temperature climatology or a synthetic SST the
			// ocean starts as the reference's does (setup/surface/index.ts:
			// 483-527)
			if ((request.terrain || request.config.earthSurfaceClimatology) && !(request.planet && "seaSurfaceK" in request.planet))
				config.sea_surface_k = NaN
			state.id = fields.length
				? addon.addMemberSurface(state.engine, config, precision, fields)
				: addon.addMember(state.engine, config, precision)
			addon.setPublishEvery(state.engine, Math.max(1, request.publishEvery | 0))
			// a continuation under an edited config resets what the edit
			// invalidates (gcm.worker.ts:110-124, continuation/index.ts:60-85)
			if (request.restart)
				addon.restore(
					state.engine,
					state.id,
					request.restart,
					request.previousConfig ? toEngine(request.previousConfig) : null,
				)
			addon.setAveraging(state.engine, state.id, true)
			if (request.averager) {
				loadAverager(state.engine, state.id, request.averager)
				const held = request.averager.months || []
				if (held.length) state.lastClosedStep = held[held.length - 1].snapshot.step
			}
			request.onDevice && request.onDevice("AMD Instinct MI355X (HIP, gfx950)")
			status(state.paused ? "paused" : "running")
			setImmediate(slice)
		} catch (error) {
			fail(error)
		}
	})

	return {
		pause() {
			state.paused = true
			status("paused")
		},
		resume() {
			if (!state.paused) return
			state.paused = false
			status("running")
			setImmediate(slice)
		},
		checkpoint() {
			state.checkpoint = true
			if (state.paused) setImmediate(slice)
		},
		stop() {
			state.stopped = true
			setImmediate(slice)
		},
	}
}

// GCM_BOUNDARY.files(GCM_BOUNDARY.fromTerrain(...)) (boundary/index.ts:
// 12-79, preparation.worker.ts:48): the land-sea mask and orography .sra
// files of a terrain raster on the Gaussian grid of `latitudes`.
function boundaryFiles(engine, terrain, latitudes, gravityMs2) {
	const { landFraction, geopotential } = addon.boundaryFromTerrain(engine, terrain, latitudes, gravityMs2)
	const prefix = `N${String(latitudes).padStart(3, "0")}_surf_0`
	return {
		[`${prefix}172.sra`]: addon.boundarySra(172, latitudes, landFraction),
		[`${prefix}129.sra`]: addon.boundarySra(129, latitudes, geopotential),
	}
}

module.exports = {
	start,
	resolutions,
	engineConfig,
	namelistConfig,
	baseNamelists,
	addon,
	readAverager,
	loadAverager,
	boundaryFiles,
	surfaceFields,
	keepSurfaceFile,
}
