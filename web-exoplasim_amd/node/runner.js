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

const path = require("path")
const addon = require(path.join(__dirname, "exoplasim_napi.node"))

// GcmConfig (src/model/climate/gcm/types.ts:29-66) onto exo_config. The
// synthetic-planet fields (landFraction, terrainMaxM, seeds, ...) come from
// request.planet; the reference reads them from terrain rasters instead.
function engineConfig(config, planet) {
	const atmosphere = config.atmosphere || {}
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
		ozone: atmosphere.ozone === undefined ? undefined : atmosphere.ozone ? 1 : 0,
		pressure_bar: atmosphere.pressureBar === null ? undefined : atmosphere.pressureBar,
		gas_constant: atmosphere.gasConstant === null ? undefined : atmosphere.gasConstant,
		storm_diagnostics: config.stormclim === undefined ? undefined : config.stormclim ? 1 : 0,
	}
	const p = planet || {}
	const extra = {
		land_fraction: p.landFraction,
		terrain_max_m: p.terrainMaxM,
		initial_temperature: p.initialTemperatureK,
		initial_humidity: p.initialHumidity,
		noise_kelvin: p.noiseKelvin,
		sea_surface_k: p.seaSurfaceK,
		co2_ppm: p.co2Ppm,
		terrain_seed: p.terrainSeed,
		noise_seed: p.noiseSeed,
	}
	const merged = Object.assign(out, extra)
	for (const key of Object.keys(merged)) if (merged[key] === undefined) delete merged[key]
	return merged
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

function start(request) {
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
			state.id = addon.addMember(state.engine, engineConfig(request.config, request.planet), precision)
			addon.setPublishEvery(state.engine, Math.max(1, request.publishEvery | 0))
			// a continuation under an edited config resets what the edit
			// invalidates (gcm.worker.ts:110-124, continuation/index.ts:60-85)
			if (request.restart)
				addon.restore(
					state.engine,
					state.id,
					request.restart,
					request.previousConfig ? engineConfig(request.previousConfig, request.planet) : null,
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

module.exports = { start, engineConfig, addon, readAverager, loadAverager, boundaryFiles }
