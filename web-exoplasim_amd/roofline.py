"""Algorithmic work of the step's kernels, for roofline figures.

Counts are per model step and member, from the transform plane counts of
one step (the job lists the step enqueues):

* inverse Legendre (k_leg_inv): 7L+3 scalar planes (sp2fc/sp2fcdmu) and
  5L dv2uv levels; forward Legendre (k_leg_fwd): 4L+1 scalar planes
  (fc2sp), L uv2dv levels with kinetic energy, L without, 2L qtend
  levels. (The dissipation's mean-mode round trip is not a transform
  here: k_mean_dekin computes its one surviving mode directly.) FLOPs use the north/south-symmetric count with an FMA as 2:
  2 NLAT NCSP per scalar plane, 8 per dv2uv level, 10 per uv2dv+KE level,
  8 per uv2dv level, 6 per qtend level. Bytes: every input plane read
  and every output plane written once (a Fourier plane is NLAT NFC
  floats, a spectral plane 2 NCSP): inverse 17L+3 spectral planes in and
  17L+3 Fourier planes out; forward 15L+1 Fourier planes in (the forward
  DFT's rows) and 10L+1 spectral planes out. Their intensity, about
  6 FLOP/B on the mix, is below the FP32 MFMA ridge (157.3 TFLOP/s over
  8 TB/s, 19.7 FLOP/B) and far below FP16's, so the bound they are
  measured against is HBM.
* zonal DFTs (k_dft_inv, k_dft_fwd): 17L+4 inverse and 15L+1 forward rows
  of every latitude; each row reads or writes NFC Fourier and NLON grid
  floats once: 4 (NLON + NFC) bytes. Bound: HBM.
* column physics (HBM): the distinct arena floats per column a kernel's
  routines read from outside the kernel plus those they write, counted
  from the kernel source, 4 bytes each:
    k_surface_fluxes (fixer apply, surflx, vdiff): reads 10L + 7, writes
      6L + 11 -> 16L + 18 floats;
    k_shortwave (solang, mko3, swr): reads 5L + 7, writes 4L + 7 ->
      9L + 14;
    k_longwave (lwr, radstep): reads 8L + 4, writes 8L + 4 -> 16L + 8;
    k_convection (kuo, mkshallow, mkdca's column pass): reads 5L + 3,
      writes 8L + 2 -> 13L + 5.
  and, counting every plane of each grid slot the routines touch (a slot
  read and written counts twice):
    k_rain (mkdca apply, mklsp, mkrain): 14L + 10;
    k_clouds (mkclouds): 8L + 10;
    k_physics_c (land surface and albedo, sea ice to getiflx): 7L + 101;
    k_physics_d (ice finish, slab ocean, sea surface, carbon, frame out):
      17L + 79;
  and for the dynamics' grid-point kernels:
    k_calcgp: reads gp, gpj, gpmt and 6 level fields, writes gp, 12 level
      fields, gvpp and dp0 -> 18L + 6;
    k_physics_in: reads gp and 4 level fields, zeroes 4 half-level and 4
      level tendencies, writes 4 level fields and dp -> 16L + 7.
"""
import os

FP32_PEAK_TFLOPS = 157.3
FP16_PEAK_TFLOPS = 2500.0  # dense FP16 MFMA: the split form's bound (its 3 products per term not counted)
HBM_PEAK_GBS = 8000.0

# The kernel bench.py times by default: the one with the largest share of
# device time in that workload's rocprofv3 kernel trace
# (profiles/ensemble64_fp32_kernel_trace.md).
DOMINANT = {"batched": "k_leg_inv", "plain": "k_leg_inv"}
# further kernels the bench line reports beside the dominant one
SECONDARY = {"batched": ("k_leg_fwd", "k_longwave", "k_dft_inv", "k_surface_fluxes", "k_dft_fwd"), "plain": ()}


def _dims(cfg):
    trunc, levels = cfg.truncation, cfg.levels
    nlat = {21: 32, 42: 64, 63: 96}[trunc]
    nlon = 2 * nlat
    ntp1 = trunc + 1
    ncsp = ntp1 * (ntp1 + 1) // 2
    nfc = 2 * ntp1
    return nlat, nlon, levels, ncsp, nfc


def work_per_step(kernel, cfg):
    """(units, algorithmic work) of one model step of one member."""
    nlat, nlon, L, ncsp, nfc = _dims(cfg)
    kernel = kernel.replace("_split", "")
    if kernel == "k_leg_inv":
        return "flop", nlat * ncsp * (2 * (7 * L + 3) + 8 * 5 * L)
    if kernel == "k_leg_fwd":
        return "flop", nlat * ncsp * (2 * (4 * L + 1) + 10 * L + 8 * L + 6 * 2 * L)
    if kernel == "k_dft_inv":
        return "byte", 4 * (nlon + nfc) * nlat * (17 * L + 4)
    if kernel == "k_dft_fwd":
        return "byte", 4 * (nlon + nfc) * nlat * (15 * L + 1)
    column = {"k_surface_fluxes": 16 * L + 18, "k_shortwave": 9 * L + 14, "k_longwave": 16 * L + 8,
              "k_convection": 13 * L + 5, "k_rain": 14 * L + 10, "k_clouds": 8 * L + 10,
              "k_physics_c": 7 * L + 101, "k_physics_d": 17 * L + 79, "k_calcgp": 18 * L + 6,
              "k_physics_in": 16 * L + 7}
    if kernel in column:
        return "byte", 4 * column[kernel] * nlat * nlon
    raise KeyError(f"no algorithmic work model for {kernel}")


def bytes_per_step(kernel, cfg):
    """Algorithmic HBM bytes of one model step of one member for the
    kernels whose work_per_step counts FLOPs (None for the others)."""
    nlat, nlon, L, ncsp, nfc = _dims(cfg)
    kernel = kernel.replace("_split", "")
    if kernel == "k_leg_inv":
        return 4 * (17 * L + 3) * (2 * ncsp + nlat * nfc)
    if kernel == "k_leg_fwd":
        return 4 * ((15 * L + 1) * nlat * nfc + (10 * L + 1) * 2 * ncsp)
    return None


LAUNCHES_PER_STEP = {"k_leg_inv": 3, "k_dft_inv": 3, "k_leg_fwd": 3, "k_dft_fwd": 3, "k_surface_fluxes": 1,
                     "k_shortwave": 1, "k_longwave": 1, "k_convection": 1, "k_rain": 1, "k_clouds": 1,
                     "k_physics_c": 1, "k_physics_d": 1, "k_calcgp": 1, "k_physics_in": 1}


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                               "ensemble64_fp32_hbm_traffic.md")


def measured_traffic(kernel, path=TRAFFIC_PROFILE):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the ensemble (scaled by the copy
    calibration it states), or None when it has no line for the kernel."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return float(parts[2])
    return None


def figure(exo, kernel, cfgs, total_ms, launches, missed, steps):
    """The bench line's roofline object for `kernel` over `steps` steps of
    `cfgs`, from the probed launch durations."""
    units, per_step, bytes_step = None, 0, 0
    for c in cfgs:
        units, w = work_per_step(kernel, c)
        per_step += w
        b = bytes_per_step(kernel, c)
        bytes_step = None if b is None or bytes_step is None else bytes_step + b
    expected = LAUNCHES_PER_STEP[kernel.replace("_split", "")] * steps
    if launches + missed != expected:
        raise RuntimeError(f"{kernel}: {launches + missed} launches seen, {expected} expected")
    per_launch = per_step / LAUNCHES_PER_STEP[kernel.replace("_split", "")]
    avg_s = total_ms / launches / 1e3
    extra = {}
    if units == "flop":
        flop_peak = FP16_PEAK_TFLOPS if kernel.endswith("_split") else FP32_PEAK_TFLOPS
        ridge = flop_peak * 1e12 / (HBM_PEAK_GBS * 1e9)
        intensity = per_step / bytes_step if bytes_step else None
        extra = {"flop_per_launch": per_launch, "achieved_tflops": per_launch / avg_s / 1e12,
                 "mfma_peak_tflops": flop_peak, "intensity_flop_per_byte": intensity,
                 "ridge_flop_per_byte": ridge}
        if intensity is not None and intensity < ridge:
            # below the ridge: HBM bounds it, measured in algorithmic bytes
            per_launch = bytes_step / LAUNCHES_PER_STEP[kernel.replace("_split", "")]
            units = "byte"
    if units == "flop":
        achieved = per_launch / avg_s / 1e12
        peak = extra["mfma_peak_tflops"]
        unit, bound = "TFLOP/s", "mfma"
    else:
        achieved = per_launch / avg_s / 1e9
        peak, unit, bound = HBM_PEAK_GBS, "GB/s", "hbm"
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
            "traffic": measured_traffic(kernel),
            "traffic_source": "profiles/" + os.path.basename(TRAFFIC_PROFILE) + " (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "copy-calibrated, averaged over every dispatch of a 4-step run)", "kernel": kernel, "avg_launch_us": avg_s * 1e6, "launches": launches,
            "algorithmic_per_launch": per_launch, "algorithmic_unit": units, **extra}
