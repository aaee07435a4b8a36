// Types of runner.js, the GcmRun over the MI355X engine. They restate the
// reference's GcmConfig (src/model/climate/gcm/types.ts:28-62), its
// atmosphere (atmosphere/types.ts:21-36), GcmRunRequest / GcmRun
// (runner/types.ts:47-74) and GcmResolution (runner/types.ts:86-101), so a
// caller of GCM_RUNNER.start can call start() with the same request.

export type GcmGas = "H2" | "He" | "N2" | "O2" | "Ar" | "Ne" | "Kr" | "CO2" | "H2O"
export type GcmAtmosphereMode = "partialPressures" | "pressureAndGasConstant"

export interface GcmAtmosphereConfig {
	mode: GcmAtmosphereMode
	partialPressures: Record<GcmGas, number | null>
	pressureBar: number | null
	gasConstant: number | null
	ozone: boolean
}

export type GcmPhysicsFilter = "gp" | "sp" | "none" | "cesaro" | "exp" | "lh"

export interface GcmConfig {
	targetTemperatureK: number | null // must be null: targeting is refused
	stormclim: boolean
	atmosphere: GcmAtmosphereConfig
	layers: 5 | 10
	periapseLongitudeDeg: number | null
	keplerian: boolean
	initialMeanAnomalyDeg: number | null
	soilWaterCapacityM: number | null
	soilHeatCapacity: number | null
	soilDepthScale: number
	landAlbedo: number | null
	oceanAlbedo: number | null
	snowIceAlbedo: number | null
	radiusEarth: number
	gravityMs2: number
	synchronous: boolean
	substellarLongitudeDeg: number
	oceanDepthM: number
	fluxWm2: number
	starTemperatureK: number
	yearDays: number
	eccentricity: number
	rotationPeriodDays: number
	obliquityDeg: number
	fixedOrbit: boolean // false is refused
	wetSoil: boolean
	vegetation: 0 | 1 | 2
	initialGrowth: number
	glaciers: { enabled: boolean; minDepthM: number; initialHeightM: number }
	physicsFilter: ReadonlyArray<GcmPhysicsFilter>
	timestepMinutes: number
	truncation: 21 | 42 | 63
	earthSurfaceClimatology: boolean
}

export interface GcmDate {
	year: number
	month: number
	day: number
	hour: number
	minute: number
}

export interface GcmSnapshot {
	layers: 5 | 10
	step: number
	date: GcmDate
	values: Float32Array
}

export interface GcmTerrain {
	name: string
	width: number
	height: number
	elevationM: Float32Array
	land: Float32Array
}

export interface GcmGpuInputs {
	inputs: string
	downloadBytes: number
}

export interface GcmResolution {
	layers: 5 | 10
	truncation: 21 | 42 | 63
	latitudes: number
	cpu: null
	gpu: GcmGpuInputs | null
}

export type GcmRunStatus = "loading" | "running" | "paused" | "checkpointing" | "finished" | "failed"

export interface GcmMonthlyMean {
	year: number
	month: number
	snapshot: GcmSnapshot
}

export interface GcmYearlyBalance {
	year: number
	toaNetRadiation: number
	surfaceNetHeat: number
	surfaceTemperature: number
}

export interface GcmAverager {
	current: { year: number; month: number; sum: Float64Array; samples: number; last: GcmSnapshot } | null
	months: GcmMonthlyMean[]
	years: GcmYearlyBalance[]
}

export interface GcmRunCheckpoint {
	targeting: null
	restart: ArrayBuffer
	averager: GcmAverager
}

// The synthetic planet of the seeded terrain (no counterpart in the
// reference, which reads terrain rasters): used when `terrain` is absent.
export interface SyntheticPlanet {
	landFraction?: number
	terrainMaxM?: number
	initialTemperatureK?: number
	initialHumidity?: number
	noiseKelvin?: number
	seaSurfaceK?: number | null
	co2Ppm?: number
	terrainSeed?: number | bigint
	noiseSeed?: number | bigint
	// true: start from the seeded synthetic state (isothermal at
	// initialTemperatureK, initialHumidity, ln(ps) from the orography)
	// instead of the reference's initfd
	syntheticStart?: boolean
}

export interface GcmRunRequest {
	config: GcmConfig
	previousConfig?: GcmConfig | null
	engine?: "gpu" | "cpu"
	resolution?: GcmResolution
	terrain?: GcmTerrain
	publishEvery: number
	restart: ArrayBuffer | null
	averager?: GcmAverager | null
	startPaused: boolean
	onSnapshot(snapshot: GcmSnapshot): void
	onAverages?(update: { months: GcmMonthlyMean[]; years: GcmYearlyBalance[] }): void
	onStatus(update: { status: GcmRunStatus; message: string }): void
	onDevice(description: string): void
	onCheckpoint(saved: GcmRunCheckpoint): void
	// beyond the reference: the Earth climatologies' inputs.json (or the
	// public/gcm directory to read resolution.gpu.inputs from), the HIP
	// device, the Legendre operand form and the synthetic planet
	inputs?: { files: Record<string, string> }
	publicDir?: string
	device?: number
	precision?: "plain" | "split"
	planet?: SyntheticPlanet
}

export interface GcmRun {
	pause(): void
	resume(): void
	checkpoint(): void
	stop(): void
}

export function start(request: GcmRunRequest): GcmRun
export function resolutions(): GcmResolution[]
export function engineConfig(config: Partial<GcmConfig>, planet?: SyntheticPlanet): Record<string, number>
