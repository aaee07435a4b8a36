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

function start(request) {
	const state = { paused: !!request.startPaused, stopped: false, checkpoint: false, engine: null, id: -1 }
	const status = (s, message = "") => request.onStatus && request.onStatus({ status: s, message })
	const fail = (error) => {
		status("failed", error && error.message ? error.message : String(error))
		close()
	}
	const close = () => {
		if (state.engine) addon.destroy(state.engine)
		state.engine = null
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
				request.onCheckpoint &&
					request.onCheckpoint({ targeting: null, restart, averager: request.averager || null })
				status(state.paused ? "paused" : "running")
			}
			if (state.paused) return
			const every = Math.max(1, request.publishEvery | 0)
			addon.step(state.engine, every)
			addon.sync(state.engine)
			const values = addon.readSnapshot(state.engine, state.id)
			request.onSnapshot &&
				request.onSnapshot({
					layers: request.config.layers,
					step: addon.memberStep(state.engine, state.id),
					date: addon.memberDate(state.engine, state.id),
					values,
				})
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
			if (request.restart) addon.restore(state.engine, state.id, request.restart)
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

module.exports = { start, engineConfig, addon }
