"""Native PlaSim restart files: exo_native_patch (host C++) against the
oracle's restatement of NATIVE_RESTART.patch (oracle/native_restart.py,
native/index.ts:3-46), both pinned on hand-built Fortran sequential files
whose byte offsets are known: patched values read back at their offsets
(FP64 exact, FP32 rounded), every other byte unchanged, and each of the
reference's errors on the file that provokes it."""
import struct

import numpy as np
import pytest

from oracle import native_restart as oracle_native


def _record(payload):
    return struct.pack("<I", len(payload)) + payload + struct.pack("<I", len(payload))


def _field(name, payload):
    return _record(name.ljust(16).encode()) + _record(payload)


def _file():
    """nstep (int32), sz (6 x FP32), dwmax (3 x FP64), dalbcl (2 x FP64)."""
    parts = [("nstep", struct.pack("<i", 1440)),
             ("sz", np.arange(6, dtype="<f4").tobytes()),
             ("dwmax", np.array([0.5, 0.25, 0.125], dtype="<f8").tobytes()),
             ("dalbcl", np.array([0.1, 0.2], dtype="<f8").tobytes())]
    blob = b"".join(_field(n, p) for n, p in parts)
    offsets, at = {}, 0
    for n, p in parts:
        at += 4 + 16 + 4
        offsets[n] = at + 4
        at += 4 + len(p) + 4
    return blob, offsets


def _both(exo, blob, fields):
    """(engine result or error text, oracle result or error text)."""
    try:
        ours = exo.native_patch(blob, fields)
    except exo.ExoError as e:
        ours = str(e)
    try:
        ref = oracle_native.patch(blob, fields)
    except oracle_native.NativeRestartError as e:
        ref = str(e)
    return ours, ref


def test_patch_writes_at_the_record_precision(exo):
    blob, off = _file()
    fields = {"dwmax": [1.0 / 3.0, 2.0, -7.5], "sz": [np.pi, 1e-8, 2.0, 3.0, 4.0, 5.0]}
    ours, ref = _both(exo, blob, fields)
    assert isinstance(ours, bytes) and ours == ref
    assert len(ours) == len(blob)
    got = np.frombuffer(ours, dtype="<f8", count=3, offset=off["dwmax"])
    np.testing.assert_array_equal(got, fields["dwmax"])
    got = np.frombuffer(ours, dtype="<f4", count=6, offset=off["sz"])
    np.testing.assert_array_equal(got, np.float32(fields["sz"]))
    # everything outside the two patched value records is unchanged
    mask = np.ones(len(blob), dtype=bool)
    mask[off["dwmax"]:off["dwmax"] + 24] = False
    mask[off["sz"]:off["sz"] + 24] = False
    assert bytes(np.frombuffer(ours, np.uint8)[mask]) == bytes(np.frombuffer(blob, np.uint8)[mask])
    assert exo.native_record(ours, "dalbcl") == np.array([0.1, 0.2], dtype="<f8").tobytes()
    assert exo.native_patch(blob, {}) == blob


@pytest.mark.parametrize("case, want", [
    ("truncated", "Truncated native restart record"),
    ("bad_marker", "Invalid native restart record"),
    ("long_name", "Invalid native restart field name"),
    ("shape", "Native restart shape mismatch for sz"),
    ("missing", "Native restart fields missing: ghost, phantom"),
])
def test_patch_errors(exo, case, want):
    blob, off = _file()
    fields = {"dwmax": [0.0, 0.0, 0.0]}
    if case == "truncated":
        blob = blob + b"\x01\x00\x00"
    elif case == "bad_marker":
        blob = blob[:-1] + b"\x7f"
    elif case == "long_name":
        blob = _record(b"x" * 17) + _record(b"")
    elif case == "shape":
        fields = {"sz": [1.0] * 5}
    elif case == "missing":
        fields = {"ghost": [1.0], "dwmax": [1.0, 2.0, 3.0], "phantom": [2.0]}
    ours, ref = _both(exo, blob, fields)
    assert ours == ref == want


def test_find_reports_absent_fields(exo):
    blob, off = _file()
    with pytest.raises(exo.ExoError, match="Native restart fields missing: nope"):
        exo.native_record(blob, "nope")
    assert exo.native_record(blob, "nstep") == struct.pack("<i", 1440)
