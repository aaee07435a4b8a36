"""Continuing a run under an edited configuration (GPU): GCM_CONTINUATION.
fields (continuation/index.ts:60-85) as the worker applies it on restart
(gcm.worker.ts:110-124). A planet saved after 40 steps is restored into a
planet whose settings differ: the fields the edit invalidates (co2, dqco2
for CO2; ymld, xmld, dmld for the mixed layer; dalbcl, dalbcl1, dalbcl2
for the land albedo or the star's temperature; dwmax for the soil water
capacity) take the new planet's initial values, every other field is the
saved one, bit for bit; without an edit the restore is plain."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CO2 = ("co2", "dqco2")
MIXED = ("ymld", "xmld", "dmld")
ALBEDO = ("dalbcl", "dalbcl1", "dalbcl2")
SOIL = ("dwmax",)
NULL = float("nan")  # a null GcmConfig value
BASE = dict(co2_ppm=360.0, ocean_depth_m=50.0, land_albedo=NULL, star_temperature_k=5780.0,
            soil_water_capacity_m=NULL)


def _cfg(exo, **edit):
    return exo.default_config(truncation=21, levels=5, land_fraction=0.4, **{**BASE, **edit})


@pytest.mark.parametrize("edit, reset", [
    (dict(co2_ppm=560.0), CO2),
    (dict(ocean_depth_m=20.0), MIXED),
    (dict(co2_ppm=560.0, ocean_depth_m=20.0), CO2 + MIXED),
    (dict(land_albedo=0.3), ALBEDO),
    (dict(star_temperature_k=4500.0), ALBEDO),
    (dict(soil_water_capacity_m=0.3), SOIL),
    ({}, ()),
], ids=["co2", "depth", "co2+depth", "land-albedo", "star", "soil", "none"])
def test_continuation_resets_what_the_edit_invalidates(exo, edit, reset):
    saved_cfg = _cfg(exo)
    eng = exo.Engine(0, 2)
    try:
        a = eng.add_member(saved_cfg)
        eng.step(40)
        eng.sync()
        blob = eng.checkpoint(a)
        saved = {name: eng.read_field(a, name) for name in exo.slot_names()}
        b = eng.add_member(_cfg(exo, **edit))
        initial = {name: eng.read_field(b, name) for name in CO2 + MIXED + ALBEDO + SOIL}
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
