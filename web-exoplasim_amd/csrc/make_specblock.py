"""Writes specblock.inc: ExoPlaSim's spectral albedo table (965 wavelengths
in micrometres and the spectral albedos, in per cent, of snow, sea ice,
glacier, bare ground and ocean at each; specblock.f90 at ExoPlaSim
5f8dff4), read from the numbers of the reference's copy of it
(src/model/climate/gcm/physics/radiation/spectrum/specblock/index.ts).
The star-weighted band albedos of the radiation set-up integrate over it.

    python web-exoplasim_amd/csrc/make_specblock.py [REFERENCE_ROOT]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
KEYS = ("wavelengths", "iceblend", "iceblendmin", "iceblendmax", "seaicemin", "seaicemax", "glacalbmin",
        "groundblend", "oceanblend")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    path = os.path.join(ref, "src/model/climate/gcm/physics/radiation/spectrum/specblock/index.ts")
    text = re.sub(r"//[^\n]*", "", open(path).read())
    arrays = {}
    for key in KEYS:
        m = re.search(r"\b" + key + r"\s*:\s*\[([^\]]*)\]", text)
        arrays[key] = [v.strip() for v in m.group(1).split(",") if v.strip()]
        float(arrays[key][0])
    n = len(arrays["wavelengths"])
    assert all(len(v) == n for v in arrays.values()), {k: len(v) for k, v in arrays.items()}
    out = ["// Generated by make_specblock.py: ExoPlaSim's specblock.f90 spectral albedo",
           f"// table, {n} wavelengths (micrometres) and albedos (per cent).",
           f"constexpr int kSpecPoints = {n};"]
    for key in KEYS:
        body = ",\n".join("    " + ", ".join(arrays[key][i:i + 10]) for i in range(0, n, 10))
        out.append(f"constexpr double kSpec_{key}[kSpecPoints] = {{\n{body}}};")
    with open(os.path.join(HERE, "specblock.inc"), "w") as f:
        f.write("\n".join(out) + "\n")
    print(f"wrote specblock.inc: {n} points x {len(KEYS)} arrays")


if __name__ == "__main__":
    main()
