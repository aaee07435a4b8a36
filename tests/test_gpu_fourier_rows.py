"""Fourier slots at the boundary (GPU).

Inside the arena a Fourier slot is coefficient-major, so that one wavenumber
of every row is contiguous for the transforms. Field reads, writes and
checkpoints show the reference's arrangement instead (LAYOUT.read,
layout/index.ts:66-108): row r's 2(T+1) coefficients together, rows north
to south per plane.

* After one step the physics frame's Fourier rows (fgt, fgq, fgp, from
  the spectral state, and fgtdt, fgqdt, the tendencies on their way back)
  agree with the FP64 oracle's rows. Any other arrangement of the same
  numbers is off by O(1), so 1e-3 relative RMS pins the arrangement; the
  precision of the values is pinned by the spectral step-1 gate, 1e-5.
* A written slot reads back unchanged, and the checkpoint payload holds
  the slot exactly as read; a restore puts it back where the transforms
  find it: one step after a restore equals one step after the save, bit
  for bit.
"""
import json

import numpy as np
import pytest

from oracle_harness import OracleRun

pytestmark = pytest.mark.gpu

ROWS = ("fgt", "fgq", "fgp", "fgtdt", "fgqdt")


def _rel(a, b):
    return np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300)


def test_fourier_rows_match_the_oracle(exo):
    cfg = exo.default_config(truncation=21, levels=10, timestep_minutes=45, land_fraction=0.4)
    gpu = exo.Engine(0, 1)
    host = exo.Engine(-1, 1)
    try:
        g = gpu.add_member(cfg)
        h = host.add_member(cfg)
        o64 = OracleRun(host, h)
        gpu.step(1)
        gpu.sync()
        o64.step(1, physics=True)
        nfc = 2 * (21 + 1)
        for name in ROWS:
            a = gpu.read_field(g, name).astype(np.float64)
            b = np.asarray(o64.field(name), dtype=np.float64)
            rel = _rel(a, b)
            # the coefficient-major image of the same numbers, for scale
            wrong = _rel(a.reshape(-1, nfc).T.reshape(-1), b)
            print(f"{name}: rel {rel:.2e} (coefficient-major image {wrong:.2e})")
            assert np.sqrt(np.mean(b ** 2)) > 0, name
            assert rel < 1e-3 < wrong, name
    finally:
        gpu.close()
        host.close()


def _payload(blob):
    """The arena part of a checkpoint: after the 17-byte magic, a 4-byte
    header length, the header JSON and padding to 4 bytes."""
    assert blob.startswith(b"GCM-WEBGPU-STATE\n")
    n = int.from_bytes(blob[17:21], "little")
    header = json.loads(blob[21:21 + n])
    assert [f[0] for f in header["fields"]][:3] == ["sd", "st", "sz"]
    return np.frombuffer(blob[(21 + n + 3) // 4 * 4:], dtype=np.float32)


def test_fourier_rows_round_trip_and_checkpoint(exo):
    cfg = exo.default_config(truncation=21, levels=5, timestep_minutes=45)
    eng = exo.Engine(0, 1)
    try:
        m = eng.add_member(cfg)
        eng.step(8)
        eng.sync()
        off, n = eng.slot(m, "fgq")
        pattern = np.arange(n, dtype=np.float32) * np.float32(0.5)
        eng.write_field(m, "fgq", pattern)
        np.testing.assert_array_equal(eng.read_field(m, "fgq"), pattern)
        blob = eng.checkpoint(m)
        payload = _payload(blob)
        np.testing.assert_array_equal(payload[off:off + n], pattern)
        fgt_off, fgt_n = eng.slot(m, "fgt")
        np.testing.assert_array_equal(payload[fgt_off:fgt_off + fgt_n], eng.read_field(m, "fgt"))
        eng.step(1)
        eng.sync()
        after = eng.read_arena(m)
        eng.restore(m, blob)
        np.testing.assert_array_equal(eng.read_field(m, "fgq"), pattern)
        eng.step(1)
        eng.sync()
        np.testing.assert_array_equal(eng.read_arena(m), after)
    finally:
        eng.close()
