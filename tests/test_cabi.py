"""The C-ABI library loads and exports every symbol its header declares
(no GPU needed), and its host-only paths behave as documented."""
import ctypes
import re
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "exoplasim_mi355x.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(exo):
    lib = exo.library()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_slot_names_cover_the_registry(exo):
    names = exo.slot_names()
    assert names[0] == "sd" and names[-1] == "snapshot_step"
    assert len(names) == len(set(names))
    assert exo.library().exo_slot_name(len(names)) is None


def test_host_only_engine_rejects_device_calls(exo):
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config())
        with pytest.raises(exo.ExoError) as err:
            e.step(1)
        assert err.value.code == -3
        with pytest.raises(exo.ExoError):
            e.read_field(m, "sd")


def test_invalid_configuration_is_rejected(exo):
    with exo.Engine(-1, 1) as e:
        with pytest.raises(exo.ExoError) as err:
            e.add_member(exo.default_config(levels=7))
        assert "5 or 10" in str(err.value)
        with pytest.raises(exo.ExoError) as err:
            e.add_member(exo.default_config(truncation=31))
        assert "T31 is not built" in str(err.value)


def test_restore_rejects_foreign_bytes(exo):
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config())
        with pytest.raises(exo.ExoError) as err:
            e.restore(m, b"EXOPLASIM-CHECKPOINT\n" + b"\0" * 64)
        assert "was not saved by the WebGPU model" in str(err.value)
        bad = b"GCM-WEBGPU-STATE\n" + (20).to_bytes(4, "little") + b'{"fields":[["st",1]]}'
        with pytest.raises(exo.ExoError) as err:
            e.restore(m, bad + b"\0" * 64)
        assert "does not match this version of the model" in str(err.value)
