"""exo_step_noalloc, the variant of exo_step that allocates nothing,
through the C-ABI: on two calls in a row (the second smaller: fewer
members and fewer steps) its members equal, bit for bit, the same members
stepped by exo_step on another engine; a request beyond its static
calendar buffer fails before any work; exo_step continues after it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXO_NOALLOC_ROWS = 16384


def test_noalloc_steps_match_exo_step(exo):
    cfgs = [exo.default_config(truncation=21, levels=5),
            exo.default_config(truncation=42, levels=10, land_fraction=0.3),
            exo.default_config(truncation=21, levels=10, terrain_seed=7)]
    ref, fixed = exo.Engine(0, 4), exo.Engine(0, 4)
    try:
        for e in (ref, fixed):
            e.set_publish_every(2)
        ids = [[e.add_member(c) for c in cfgs] for e in (ref, fixed)]
        assert ids[0] == ids[1]
        members = ids[0]

        def compare(label):
            for e in (ref, fixed):
                e.sync()
            for m in members:
                a, b = ref.read_arena(m), fixed.read_arena(m)
                assert np.array_equal(a, b), f"{label}: member {m} differs"
                assert ref.member_step(m) == fixed.member_step(m)

        ref.step(8)  # start-up steps and main-loop steps, publishing
        fixed.step_noalloc(8)
        compare("first call")
        # the second call is smaller: one member fewer, fewer steps
        for e in (ref, fixed):
            e.remove_member(members[1])
        members = [members[0], members[2]]
        ref.step(3)
        fixed.step_noalloc(3)
        compare("second call")
        # beyond the static buffer: rejected, nothing stepped
        before = fixed.member_step(members[0])
        with pytest.raises(exo.ExoError) as err:
            fixed.step_noalloc(EXO_NOALLOC_ROWS // len(members) + 1)
        assert err.value.code == -1 and "exo_step_noalloc" in str(err.value)
        assert fixed.member_step(members[0]) == before
        # the ordinary entry point continues from the same state
        ref.step(4)
        fixed.step(4)
        compare("exo_step after exo_step_noalloc")
        ts, _ = fixed.global_means(members[0])
        assert 150.0 < ts < 400.0
    finally:
        ref.close()
        fixed.close()
