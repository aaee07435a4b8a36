"use strict"
// The namelist files of a run, host side: GCM_NAMELIST (src/model/climate/
// gcm/namelist/index.ts:40-243). edits() turns a GcmConfig into the values
// each namelist file gets, apply() writes them into a file's text, parse()
// and read() read a file back as the set-up does. The engine reads the
// files through exoplasim_napi's configFromNamelists
// (exo_config_from_namelists).

const path = require("path")
const addon = require(path.join(__dirname, "exoplasim_napi.node"))

// ExoPlaSim's Python configure() constants the edits use
const EARTH_RADIUS_M = 6371220
const NSTPW_WINDOW_MINUTES = 7200
const SECONDS_PER_DAY = 86400
const SOIL_LAYER_DEPTHS_M = [0.4, 0.8, 1.6, 3.2, 6.4]
// physicsFilter: the switch each filter kind sets
const FILTER_SWITCH = {
	gp: ["NGPTFILTER", "1"],
	sp: ["NSPVFILTER", "1"],
	none: ["NFILTER", "0"],
	cesaro: ["NFILTER", "1"],
	exp: ["NFILTER", "2"],
	lh: ["NFILTER", "3"],
}
const GASES = ["H2", "He", "N2", "O2", "Ar", "Ne", "Kr", "CO2", "H2O"]

const bit = (value) => (value ? "1" : "0")
const nan = (value) => (value === null || value === undefined ? NaN : value)

// GCM_ATMOSPHERE.resolve through the engine's exo_atmosphere_resolve.
function resolveAtmosphere(atmosphere) {
	const mode = atmosphere.mode
	if (mode !== "partialPressures" && mode !== "pressureAndGasConstant")
		throw new Error("Choose an atmosphere input mode.")
	return addon.atmosphereResolve(
		GASES.map((gas) => nan((atmosphere.partialPressures || {})[gas])),
		mode === "pressureAndGasConstant",
		nan(atmosphere.pressureBar),
		nan(atmosphere.gasConstant),
	)
}

function checkRange(values, ok, message) {
	for (const v of values) if (v !== null && !ok(v)) throw new Error(message)
}

// edits (namelist/index.ts:40-182): {file: {KEY: text}} for a full
// GcmConfig, the values written as JavaScript's String() writes them.
function edits(config) {
	const { pressureBar, gasConstant, co2Ppm } = resolveAtmosphere(config.atmosphere)
	const planet = {
		GA: String(config.gravityMs2),
		PLARAD: String(config.radiusEarth * EARTH_RADIUS_M),
		GSOL0: String(config.fluxWm2),
		ECCEN: String(config.eccentricity),
		OBLIQ: String(config.obliquityDeg),
		NFIXORB: bit(config.fixedOrbit),
		NGENKEPLERIAN: bit(config.keplerian),
		GASCON: String(gasConstant),
		SIDEREAL_YEAR: String(config.yearDays * SECONDS_PER_DAY),
	}
	if (config.layers !== 5 && config.layers !== 10) throw new Error("Choose 5 or 10 atmospheric layers")
	checkRange(
		[config.periapseLongitudeDeg, config.initialMeanAnomalyDeg],
		(v) => Number.isFinite(v) && v >= -360 && v <= 360,
		"Orbital angles must be finite and between -360 and 360 degrees",
	)
	if (config.periapseLongitudeDeg !== null) {
		if (config.periapseLongitudeDeg < 0) throw new Error("Periapse longitude must be between 0 and 360 degrees")
		planet.MVELP = String(config.periapseLongitudeDeg)
	}
	if (config.initialMeanAnomalyDeg !== null) planet.MEANANOM0 = String(config.initialMeanAnomalyDeg)

	const plasim = {
		FIXEDLON: String(config.substellarLongitudeDeg),
		N_DAYS_PER_YEAR: String(config.yearDays),
		PSURF: String(pressureBar * 1e5),
		NVEG: String(config.vegetation),
		NEQSIG: "0",
		MPSTEP: String(config.timestepMinutes),
		NSTPW: String(Math.floor(NSTPW_WINDOW_MINUTES / config.timestepMinutes)),
		FILTERKAPPA: "8",
		NFILTEREXP: "8",
		LANDHOSKN0: "15",
	}
	if (config.layers === 5) plasim.NSELA = "0"
	if (config.rotationPeriodDays !== 1) {
		planet.ROTSPD = String(1 / config.rotationPeriodDays)
		const months = Math.max(Math.floor(config.yearDays / config.rotationPeriodDays / 12 + 0.5), 1)
		plasim.N_DAYS_PER_YEAR = String(months * 12)
	}
	plasim.N_START_YEAR = plasim.N_DAYS_PER_YEAR === "365" ? "1" : "0"
	for (const kind of config.physicsFilter) {
		const [key, value] = FILTER_SWITCH[kind]
		plasim[key] = value
	}

	const land = { NWETSOIL: bit(config.wetSoil) }
	checkRange(
		[config.soilDepthScale, config.soilWaterCapacityM, config.soilHeatCapacity],
		(v) => Number.isFinite(v) && v > 0,
		"Soil depth, water capacity and heat capacity must be positive",
	)
	land.DSOILZ = SOIL_LAYER_DEPTHS_M.map((depth) => depth * config.soilDepthScale).join(",")
	if (config.soilWaterCapacityM !== null) land.WSMAX = String(config.soilWaterCapacityM)
	if (config.soilHeatCapacity !== null) land.SOILCAP = String(config.soilHeatCapacity)
	const sea = {}
	checkRange(
		[config.landAlbedo, config.oceanAlbedo, config.snowIceAlbedo],
		(v) => Number.isFinite(v) && v >= 0 && v <= 1,
		"Albedo overrides must be between 0 and 1",
	)
	const twoBands = (v) => `${v},${v}`
	if (config.landAlbedo !== null) {
		land.ALBLAND = String(config.landAlbedo)
		land.DGROUNDALB = twoBands(config.landAlbedo)
	}
	if (config.oceanAlbedo !== null) {
		sea.ALBSEA = String(config.oceanAlbedo)
		sea.DOCEANALB = twoBands(config.oceanAlbedo)
	}
	if (config.snowIceAlbedo !== null) {
		const bands = twoBands(String(config.snowIceAlbedo))
		Object.assign(sea, { ALBICE: String(config.snowIceAlbedo), DICEALBMN: bands, DICEALBMX: bands })
		Object.assign(land, { DSNOWALBMN: bands, DGLACALBMN: bands, DSNOWALB: bands })
	}

	const radmod = { NFIXED: bit(config.synchronous) }
	if (config.starTemperatureK > 0) {
		radmod.NSTARTEMP = "1"
		radmod.STARBBTEMP = String(config.starTemperatureK)
	}
	Object.assign(radmod, {
		NO3: bit(config.atmosphere.ozone),
		NRADINT: "1",
		CO2: String(co2Ppm),
		NPBROADEN: "1",
		NECHAM: "1",
		NECHAM6: "0",
	})
	const out = {
		planet_namelist: planet,
		plasim_namelist: plasim,
		radmod_namelist: radmod,
		landmod_namelist: land,
	}
	if (Object.keys(sea).length > 0) out.seamod_namelist = sea
	out.oceanmod_namelist = { MLDEPTH: String(config.oceanDepthM) }
	out.vegmod_namelist = {
		NCVEG: "1",
		FORGROW: "1.0",
		RINIDSC: "1.0",
		RINIDMR: "2.0",
		RINISOIL: "0.0",
		RINIVEG: "0.0",
		RINIDAGG: String(config.initialGrowth),
	}
	out.hurricane_namelist = { NSTORMDIAG: bit(config.stormclim), HC_CAPTURE: "0" }
	out.glacier_namelist = {
		NGLACIER: bit(config.glaciers.enabled),
		GLACELIM: String(config.glaciers.minDepthM),
		ICESHEETH: String(config.glaciers.initialHeightM),
	}
	return out
}

// apply (namelist/index.ts:184-194): each KEY's first line (any case,
// leading spaces) becomes " KEY = value"; a new key goes in before the
// group's /END.
function apply(text, values) {
	let result = text
	for (const [key, value] of Object.entries(values)) {
		const line = ` ${key} = ${value}`
		const existing = new RegExp(`^\\s*${key}\\s*=.*$`, "im")
		result = existing.test(result) ? result.replace(existing, line) : result.replace(/^\s*\/END/im, `${line}\n /END`)
	}
	return result
}

// parse (namelist/index.ts:196-206): a file's KEY = value lines, keys
// upper-cased, values as written.
function parse(text) {
	const values = {}
	for (const line of text.split(/\r?\n/)) {
		const m = /^\s*([A-Za-z_][A-Za-z0-9_]*)\s*=\s*(.*?)\s*$/.exec(line)
		if (m) values[m[1].toUpperCase()] = m[2]
	}
	return values
}

// read (namelist/index.ts:212-233): the defaults with the file's values,
// arrays filled from the front.
function read(text, defaults) {
	const values = parse(text || "")
	const out = Object.assign({}, defaults)
	for (const [key, fallback] of Object.entries(defaults)) {
		const value = values[key.toUpperCase()]
		if (value === undefined) continue
		const parts = value.split(",").map((part) => Number(part.trim().replace(/d/i, "e")))
		if (Array.isArray(fallback)) {
			const array = fallback.slice()
			parts.forEach((part, i) => {
				array[i] = part
			})
			out[key] = array
		} else out[key] = parts[0]
	}
	return out
}

// The namelist files of a run: the base files with the config's edits
// applied (runner/webgpu/index.ts:61-79). A file the edits name that the
// base lacks starts as an empty group.
function files(base, config) {
	const out = {}
	for (const [name, text] of Object.entries(base)) if (name.endsWith("_namelist")) out[name] = text
	for (const [name, values] of Object.entries(edits(config))) {
		const group = name.replace(/_namelist$/, "_nl")
		out[name] = apply(out[name] === undefined ? ` &${group}\n /END\n` : out[name], values)
	}
	return out
}

module.exports = {
	edits,
	apply,
	parse,
	read,
	files,
	wetSoilOverridden: (config) => config.wetSoil && config.starTemperatureK > 0,
}
