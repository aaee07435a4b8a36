"""Continuing a run under an edited configuration (GPU): GCM_CONTINUATION.
fields (continuation/index.ts:60-85) as the worker applies it on restart
(gcm.worker.ts:110-124). A planet saved after 40 steps is restored into a
planet whose CO2 and/or ocean depth differ: the fields the edit
invalidates (co2, dqco2 for CO2; ymld, xmld, dmld for the mixed layer)
take the new planet's initial values, every other field is the saved one,
bit for bit; without an edit the restore is plain."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CO2 = ("co2", "dqco2")
MIXED = ("ymld", "xmld", "dmld")


def _cfg(exo, co2, depth):
    return exo.default_config(truncation=21, levels=5, co2_ppm=co2, ocean_depth_m=depth)


@pytest.mark.parametrize("co2, depth, reset", [
    (560.0, 50.0, CO2),
    (360.0, 20.0, MIXED),
    (560.0, 20.0, CO2 + MIXED),
    (360.0, 50.0, ()),
])
def test_continuation_resets_what_the_edit_invalidates(exo, co2, depth, reset):
    saved_cfg = _cfg(exo, 360.0, 50.0)
    eng = exo.Engine(0, 2)
    try:
        a = eng.add_member(saved_cfg)
        eng.step(40)
        eng.sync()
        blob = eng.checkpoint(a)
        saved = {name: eng.read_field(a, name) for name in exo.slot_names()}
        b = eng.add_member(_cfg(exo, co2, depth))
        initial = {name: eng.read_field(b, name) for name in CO2 + MIXED}
        eng.restore(b, blob, previous=saved_cfg)
        assert eng.member_step(b) == eng.member_step(a)
        for name in exo.slot_names():
            got = eng.read_field(b, name)
            want = initial[name] if name in reset else saved[name]
            np.testing.assert_array_equal(got, want, err_msg=name)
        if reset:
            assert any(not np.array_equal(initial[n], saved[n]) for n in reset)
        # the continued run steps on
        eng.step(4)
        eng.sync()
        assert np.all(np.isfinite(eng.read_snapshot(b)))
    finally:
        eng.close()
