"""Writes tests/golden/radiation_fit_constants.json: the two radiation fit
coefficients that no physical property fixes, read out of the reference's
WGSL source text (nothing of the reference is imported or run):

* the slope per decade of path of the CO2 absorptivity's logarithmic branch,
  `aco2 = C * log10(sumco2) + ZACO2C` (physics/radiation/wgsl/lwr.wgsl:122);
* the exponent of the ozone absorption's ultraviolet term,
  `1.082 * o3 / precise_pow(1.0 + 138.6 * o3, E)`
  (physics/radiation/wgsl/swr.wgsl:36).

    python tests/golden/make_fit_constants.py [REFERENCE_ROOT]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
RAD = "src/model/climate/gcm/physics/radiation/wgsl"
NUM = r"([0-9]+\.[0-9]*(?:[eE][-+]?[0-9]+)?)"


def find(path, pattern):
    """The one match of `pattern` in the file: (value, 1-based line)."""
    with open(path) as f:
        lines = f.read().split("\n")
    hits = [(float(m.group(1)), i + 1) for i, line in enumerate(lines) for m in [re.search(pattern, line)] if m]
    if len(hits) != 1:
        raise SystemExit(f"{path}: expected one match of {pattern!r}, found {len(hits)}")
    return hits[0]


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    lwr = os.path.join(root, RAD, "lwr.wgsl")
    swr = os.path.join(root, RAD, "swr.wgsl")
    slope, slope_line = find(lwr, r"aco2\s*=\s*" + NUM + r"\s*\*\s*log10\(\s*sumco2\s*\)")
    expo, expo_line = find(swr, r"precise_pow\(\s*1\.0\s*\+\s*[0-9.]+\s*\*\s*o3\s*,\s*" + NUM + r"\s*\)")
    out = {
        "lwr_co2_log_slope": {"value": slope, "source": f"{RAD}/lwr.wgsl:{slope_line}"},
        "swr_o3_uv_exponent": {"value": expo, "source": f"{RAD}/swr.wgsl:{expo_line}"},
    }
    path = os.path.join(HERE, "radiation_fit_constants.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", path, out)


if __name__ == "__main__":
    main()
