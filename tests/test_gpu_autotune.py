"""GPU tests of the transform-group autotuner: every group size gives
bit-identical members, and tuning leaves the state, step counters and
published snapshots exactly as they were."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _planets(exo):
    workloads = importlib.import_module(exo.__name__ + ".workloads")
    return workloads.ensemble(exo, seed=23, mix=((21, 5, 5), (42, 10, 3)))


def _run(exo, cfgs, steps, tune_at=None, groups=None):
    with exo.Engine(0, len(cfgs)) as e:
        e.set_publish_every(3)
        if groups:
            for (t, l), n in groups.items():
                e.set_transform_group(t, l, n)
        ids = [e.add_member(c) for c in cfgs]
        ms = None
        if tune_at is not None:
            e.step(tune_at)
            ms = e.autotune(trial_steps=4)
            picked = {k: e.transform_group(*k) for k in ((21, 5), (42, 10))}
            assert all(1 <= n <= 8 for n in picked.values())
            assert e.step_graphs() in (True, False)
            steps -= tune_at
        e.step(steps)
        e.sync()
        arenas = [e.read_arena(i) for i in ids]
        snaps = [e.read_snapshot(i) for i in ids]
        counters = [e.member_step(i) for i in ids]
    return arenas, snaps, counters, ms


@pytest.mark.parametrize("groups", [{(21, 5): 1, (42, 10): 1}, {(21, 5): 3, (42, 10): 2},
                                    {(21, 5): 4, (42, 10): 4}])
def test_group_sizes_are_bit_identical(exo, groups):
    cfgs = _planets(exo)
    ref = _run(exo, cfgs, 9)
    got = _run(exo, cfgs, 9, groups=groups)
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("tune_at", [1, 6])
def test_autotune_leaves_the_run_unchanged(exo, tune_at):
    """Tuned during the start-up phases (step 1) or in the main loop."""
    cfgs = _planets(exo)
    ref = _run(exo, cfgs, 12)
    got = _run(exo, cfgs, 12, tune_at=tune_at)
    assert got[3] > 0
    assert ref[2] == got[2]
    for a, b in zip(ref[0], got[0]):
        assert np.array_equal(a, b)
    for a, b in zip(ref[1], got[1]):
        assert np.array_equal(a, b)
