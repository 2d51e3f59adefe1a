"""Host-side logic that needs no device: schedule table lookup, request
stream generation, latency percentiles, scene generators."""
import json

import numpy as np


def test_schedule_lookup_buckets_and_defaults(tmp_path):
    from vxsim import tune
    # one-scene buckets: 64^3 (262,144 REVs) and 192x192x96 (3.5 M) get their own
    assert [tune.bucket_of(x) for x in (1, 1.9e4, 2e4, 1e5, 262144, 589824, 3.5e6, 1e9)] == [0, 0, 1, 1, 2, 3, 4, 4]
    assert [tune.bucket_of(x, "batch") for x in (1, 1.9e4, 2e4, 1e5, 3e6, 1e9)] == [0, 0, 1, 1, 3, 3]
    assert tune.lookup(5e4, "single", table={}) == tune.DEFAULT["single"]
    t = {"single": {"1": {"iters": 32, "zcols": 0, "deal": 1, "update_threads": 256}},
         "batch": {"0": {"chunk": 8, "iters": 8}}}
    assert tune.lookup(5e4, "single", table=t) == {"iters": 32, "zcols": 0, "deal": 1, "update_threads": 256}
    assert tune.lookup(5e3, "batch", table=t) == {"chunk": 8, "iters": 8}
    assert tune.lookup(5e5, "batch", table=t) == tune.DEFAULT["batch"]
    # a table written before the shape knobs keeps the default shape
    assert tune.lookup(5e4, "single", table={"single": {"1": {"iters": 8}}}) == dict(tune.DEFAULT["single"], iters=8)
    path = tmp_path / "t.json"
    path.write_text(json.dumps(t))
    assert tune.load_table(str(path))["single"]["1"]["iters"] == 32


def test_shipped_schedule_table_is_well_formed():
    """The measured table next to vxsim.tune: every one-scene bucket has a
    schedule and a launch shape the library accepts."""
    from vxsim import tune
    t = tune.load_table(tune.TABLE)
    for b in range(len(tune.BUCKETS["single"])):
        e = tune.lookup([1e4, 1e5, 3e5, 6e5, 4e6][b], "single", table=t)
        assert e["iters"] % 2 == 0 and 2 <= e["iters"] <= 256
        assert e["zcols"] in (0, 1, 2, 4, 8, 16, 32, 64)
        assert e["deal"] in (0, 1, 2, 3)
        assert e["update_threads"] in (256, 512)
    for b in range(3):
        e = tune.lookup([1e4, 1e5, 1e6][b], "batch", table=t)
        assert e["chunk"] >= 1 and e["iters"] % 2 == 0


def test_request_stream_is_seeded_and_in_range():
    from vxsim import serve
    a = serve.request_stream(rate=50.0, seconds=0.5, rev_lo=200, rev_hi=5000, steps=2, seed=3)
    b = serve.request_stream(rate=50.0, seconds=0.5, rev_lo=200, rev_hi=5000, steps=2, seed=3)
    assert len(a) == len(b) > 5
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(np.array_equal(x.scene.positions, y.scene.positions) for x, y in zip(a, b))
    assert all(0 < r.arrival < 0.5 for r in a) and sorted(r.arrival for r in a) == [r.arrival for r in a]
    assert all(100 <= r.scene.n <= 10000 and r.scene.kinematic.any() for r in a)
    p = serve.percentiles([1.0, 2.0, 3.0, 4.0, None])
    assert p[50] == 2.5 and 3.9 < p[99] <= 4.0


def test_large_block_scene():
    from vxsim import scenes
    s = scenes.wood_block_blade(dims=(24, 20, 10), blade=(6, 2, 4))
    targeted = ~np.isnan(s.targets[:, 2])
    assert targeted.sum() == 6 * 2 and np.isnan(s.targets[targeted, :2]).all()
    assert (s.keys[targeted, 2] == 9).all() and (s.kinematic[s.keys[:, 2] == 0] == 1).all()
    assert s.bond_count() == 23 * 20 * 10 + 24 * 19 * 10 + 24 * 20 * 9


class _FakeBatch:
    """Records vx_batch_hold calls instead of making them."""
    def __init__(self):
        self.calls = []

    def hold(self, slot, hold=True):
        self.calls.append((slot, bool(hold)))


def _server(active_revs, aging_s=10.0):
    from vxsim import serve
    S = serve.Server.__new__(serve.Server)
    S.batch, S.active_revs, S.aging_s, S.policy = _FakeBatch(), active_revs, aging_s, "sjf"
    return S


def _req(rid, n, arrival=0.0, steps=10):
    from types import SimpleNamespace
    from vxsim import serve
    return serve.Request(rid=rid, arrival=arrival, scene=SimpleNamespace(n=n), steps=steps, params=None)


def test_sjf_runs_the_small_scenes_first_and_the_large_together():
    """Small scenes (below half the active budget) run smallest first up to the
    budget; large ones are held while any small one runs, then all run
    together; a large one that has waited the aging time runs with them."""
    S = _server(active_revs=1000)
    running = {0: _req(0, 900), 1: _req(1, 100), 2: _req(2, 480), 3: _req(3, 490), 4: _req(4, 2000)}
    held = {}
    S._choose(running, held, now=1.0)
    # small: 100, 480, 490 (< 500); 100 + 480 fits, + 490 would pass 1000
    runs = lambda: {s for s in running if not held.get(s, False)}  # noqa: E731
    assert runs() == {1, 2}
    # holds are sent only when a scene's state changes
    n = len(S.batch.calls)
    S._choose(running, held, now=1.0)
    assert len(S.batch.calls) == n
    # the small ones finish: the large ones run together
    for s in (1, 2, 3):
        running.pop(s)
    S._choose(running, held, now=2.0)
    assert runs() == {0, 4}
    # a new small request holds them again, until one of them has waited the aging time
    running[5] = _req(5, 50, arrival=2.0)
    S._choose(running, held, now=3.0)
    assert runs() == {5}
    S._choose(running, held, now=10.5)
    assert runs() == {0, 4, 5}


def test_sjf_orders_by_size_and_steps_with_aging():
    S = _server(active_revs=1 << 20, aging_s=10.0)
    small_late, big_early = _req(0, 1000, arrival=5.0), _req(1, 3000, arrival=0.0)
    assert S._key(small_late, 5.0) < S._key(big_early, 5.0)
    # after 40 s of waiting the large request's key has dropped below the new one's
    assert S._key(big_early, 40.0) < S._key(_req(2, 1000, arrival=40.0), 40.0)
