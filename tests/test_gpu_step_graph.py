"""GPU test of the captured step graph: a batch stepped by graph replays
(on one stream, or split over two concurrent streams) holds the same
state, bit for bit, as the same batch stepped by single launches, across
start-up phases, snapshot publishing, averaging, eager work between calls
and a change of members (which re-captures)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _planets(exo):
    import importlib
    workloads = importlib.import_module(exo.__name__ + ".workloads")
    return workloads.ensemble(exo, seed=11, mix=((21, 5, 2), (42, 10, 1)))


@pytest.mark.parametrize("lanes", [1, 2, 3])
def test_graph_replays_match_single_launches(exo, lanes):
    """With lanes = 2 the replayed step runs the T42 member on one stream
    and the two T21 members on another, concurrently; with 3, each member
    on a stream of its own. The single launches run on one stream, or in
    two lanes against the one-stream replay."""
    cfgs = _planets(exo)
    extra = exo.default_config()
    engines = [exo.Engine(0, 4), exo.Engine(0, 4)]
    try:
        engines[0].set_step_graphs(False)
        engines[0].set_step_lanes(2 if lanes == 1 else 1)
        engines[1].set_step_graphs(True)
        engines[1].set_step_lanes(lanes)
        assert engines[1].step_lanes() == lanes
        ids = []
        for e in engines:
            e.set_publish_every(2)
            ids.append([e.add_member(c) for c in cfgs])
        assert ids[0] == ids[1]
        members = ids[0]
        for e in engines:
            e.set_averaging(members[2], True)

        def compare(label):
            for e in engines:
                e.sync()
            for m in members:
                a, b = (e.read_arena(m) for e in engines)
                assert np.array_equal(a, b), f"{label}: member {m} arena differs"
                sa, sb = (e.read_snapshot(m) for e in engines)
                assert np.array_equal(sa, sb), f"{label}: member {m} snapshot differs"

        # start-up phases (one step each, single launches) then main steps
        for chunk in (1, 3, 5, 8, 8):
            for e in engines:
                e.step(chunk)
            compare(f"after {chunk} steps")
            # eager work on the stream between replays
            for e in engines:
                e.global_means(members[0])
        # a member leaves and another joins: the graph is captured again
        for e in engines:
            e.remove_member(members[1])
        members = [members[0], members[2]]
        for e in engines:
            members_new = e.add_member(extra)
        members.append(members_new)
        for chunk in (4, 6):
            for e in engines:
                e.step(chunk)
            compare(f"after re-admission and {chunk} steps")
        ts = [e.global_means(m)[0] for e in engines for m in members]
        assert all(150.0 < t < 400.0 for t in ts)
    finally:
        for e in engines:
            e.close()


def test_recapture_while_replays_are_queued(exo):
    """A member joins while earlier replays may still be queued (no sync in
    between): the executable they use is replaced only after they finish,
    and the state equals single launches bit for bit."""
    cfgs = _planets(exo)
    engines = [exo.Engine(0, 4), exo.Engine(0, 4)]
    try:
        engines[0].set_step_graphs(False)
        engines[1].set_step_graphs(True)
        ids = [[e.add_member(c) for c in cfgs[:2]] for e in engines]
        for e in engines:
            e.step(12)  # start-up, then replays of one captured step
        joined = [e.add_member(cfgs[2]) for e in engines]
        for e in engines:
            e.step(6)  # the batch changed: captured again
        for e in engines:
            e.sync()
        assert ids[0] == ids[1] and joined[0] == joined[1]
        for m in ids[0] + [joined[0]]:
            a, b = (e.read_arena(m) for e in engines)
            assert np.array_equal(a, b), f"member {m} arena differs"
    finally:
        for e in engines:
            e.close()
