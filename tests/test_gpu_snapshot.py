"""GPU snapshot pack against the oracle's restatement of pack.wgsl, plane
by plane (all 101 planes at 10 levels, all 76 at 5).

The pack reads the step's final state and the extrema tracking of the
step before. So the test reads a member's arena just before a publishing
step (its tracking and snapshot) and just after it (the state the pack
read, and the GPU's snapshot and tracking); the FP32 oracle packs the
after-state with the before-tracking, and every plane and tracking row
must equal the GPU's bit for bit. Two planes are held to 1e-6 relative
instead: the surface net heat (a five-term sum the GPU may contract into
fused multiply-adds) and the lowest level's relative humidity (an
exponential, device libm against host libm). Publishing steps 4 (before
any output window closed), 160 (a window closes: extrema roll over) and
164 (after the close) are checked, for a T21 L10 and a T42 L5 member in
one batch."""
import numpy as np
import pytest

from oracle_harness import OracleRun

pytestmark = pytest.mark.gpu

CASES = [(21, 10), (42, 5)]
PUBLISH_EVERY = 4


def _slot(layout, name):
    for n, off, cnt in layout:
        if n == name:
            return off, cnt
    raise KeyError(name)


def test_snapshot_planes_match_the_oracle_pack(exo):
    cfgs = [exo.default_config(truncation=t, levels=l) for t, l in CASES]
    host = exo.Engine(-1, len(cfgs))
    hosts = [host.add_member(c) for c in cfgs]
    with exo.Engine(0, len(cfgs)) as gpu:
        gpu.set_publish_every(PUBLISH_EVERY)
        ids = [gpu.add_member(c) for c in cfgs]
        layout = gpu.layout(ids[0]), gpu.layout(ids[1])
        checked = 0
        for target in (4, 160, 164):
            # step until the next step to run is `target`
            while gpu.member_step(ids[0]) < target:
                gpu.step(1)
            gpu.sync()
            assert gpu.member_step(ids[0]) == target
            before = [gpu.read_arena(m) for m in ids]
            gpu.step(1)
            gpu.sync()
            after = [gpu.read_arena(m) for m in ids]
            for k, (cfg, m) in enumerate(zip(cfgs, ids)):
                lay = layout[k]
                nh = gpu.dims(m)[1] * gpu.dims(m)[2]
                L = cfg.levels
                run = OracleRun(host, hosts[k], precision="fp32")
                x = after[k].copy()
                for name in ("snapshot", "snapshot_tracking"):
                    off, n = _slot(lay, name)
                    x[off:off + n] = before[k][off:off + n]
                off, n = _slot(lay, "snapshot_step")
                assert x[off + 2] == 1.0, "the step publishes"
                run.arena[:] = x
                run.pack()
                names = exo.snapshot_field_names(L)
                off, n = _slot(lay, "snapshot")
                got = after[k][off:off + n].reshape(-1, nh)
                want = run.arena[off:off + n].reshape(-1, nh)
                assert got.shape[0] == len(names)
                loose = {17, 51 + 5 * (L - 1) + 3}
                for p, name in enumerate(names):
                    if p in loose:
                        np.testing.assert_allclose(got[p], want[p], rtol=1e-6, atol=1e-30, err_msg=name)
                    else:
                        assert np.array_equal(got[p], want[p]), f"step {target} T{cfg.truncation} L{L} {name}"
                off, n = _slot(lay, "snapshot_tracking")
                assert np.array_equal(after[k][off:off + n], run.arena[off:off + n]), f"tracking at {target}"
                checked += 1
        # the window closed at step 160: the published extrema are that window's
        off, n = _slot(layout[0], "snapshot_tracking")
        nh = gpu.dims(ids[0])[1] * gpu.dims(ids[0])[2]
        assert np.all(after[0][off + 8 * nh:off + 9 * nh] == 1.0)
    host.close()
    assert checked == 6
