"""The bit-exact contracts against the reference's own declarations (CPU):
the state arena's slots, in order, with their shapes and sizes, and the
snapshot's field order, compared with tests/golden/registry.json, which
tests/golden/make_registry.py reads out of the reference's TypeScript
source text (stepper/index.ts:38-50 and each domain's fields();
fields/index.ts:81-128 ORDER and fieldCount). The checkpoint payload is
the arena in this slot order (checkpoint/restart/index.ts:14-20), so the
same comparison fixes the checkpoint's field order and sizes."""
import json
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "registry.json")


@pytest.fixture(scope="module")
def registry():
    with open(GOLDEN) as f:
        return json.load(f)


def _plane(shape, t, nlat, nlon):
    # planeFloats (layout/index.ts:14-20)
    return {"spectral": (t + 1) * (t + 2), "fourier": nlat * 2 * (t + 1), "grid": nlat * nlon,
            "table": 1}[shape]


@pytest.mark.parametrize("truncation,levels", [(21, 10), (21, 5), (42, 10), (63, 5)])
def test_arena_slots_match_the_reference_registry(exo, registry, truncation, levels):
    ref = registry["slots"][str(levels)]
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config(truncation=truncation, levels=levels))
        t, nlat, nlon, _ = e.dims(m)
        layout = e.layout(m)
        total = e.arena_floats(m)
    assert [n for n, _, _ in layout] == [n for n, _, _ in ref]
    at = 0
    for (name, off, n), (rname, shape, count) in zip(layout, ref):
        assert off == at, name
        assert n == _plane(shape, t, nlat, nlon) * count, (name, shape, count, n)
        at += n
    assert at == total


@pytest.mark.parametrize("levels", [5, 10])
def test_snapshot_order_matches_the_reference(exo, registry, levels):
    names = exo.snapshot_field_names(levels)
    assert names == registry["snapshot"][str(levels)]
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config(levels=levels))
        _, nlat, nlon, _ = e.dims(m)
        _, n = e.slot(m, "snapshot")
    assert n == len(names) * nlat * nlon
