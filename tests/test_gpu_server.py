"""Continuous-batching server: results must equal the same request stepped alone (bit-exact),
admission must respect the HBM budget, cancellation must not disturb other requests."""
import numpy as np
import pytest

from helpers import flat_state

pytestmark = pytest.mark.gpu


def _reqs(k=8, seed=77):
    from pe_amd import scenes
    rng = np.random.default_rng(seed)
    counts = scenes.log_uniform_counts(rng, k, 16, 300)
    hor = rng.integers(5, 25, k)
    return [(scenes.grid_world(int(c), scenes.world_seed(seed, i)), int(h)) for i, (c, h) in enumerate(zip(counts, hor))]


def _alone(sc, horizon):
    import pe_amd
    w = pe_amd.World(sc)
    w.step(horizon)
    w.sync()
    return flat_state(w.state())


def test_server_mix_equals_alone():
    from pe_amd.server import Server
    reqs = _reqs()
    print(f"[server] sizes {[r[0].n for r in reqs]} horizons {[r[1] for r in reqs]}", flush=True)
    srv = Server(hbm_budget_bytes=2 << 30, max_worlds=64)
    ids = {}
    done = {}
    # staggered arrivals: two requests per tick, while earlier ones run
    pending = list(enumerate(reqs))
    for _ in range(200):
        for _ in range(2):
            if pending:
                i, (sc, h) = pending.pop(0)
                ids[srv.submit(sc, h)] = i
        srv.tick()
        for rid in srv.poll():
            done[ids[rid]] = flat_state(srv.result(rid))
        if len(done) == len(reqs):
            break
    assert len(done) == len(reqs)
    for i, (sc, h) in enumerate(reqs):
        assert np.array_equal(done[i], _alone(sc, h)), f"request {i}"
    st = srv.stats()
    assert st["ticks"] > 0 and st["p99_ms"] >= st["p50_ms"] > 0


def test_server_budget_limits_admission():
    from pe_amd.server import Server
    from pe_amd import scenes
    sc = scenes.grid_world(200, 5)
    fp = Server.footprint(sc)
    srv = Server(hbm_budget_bytes=int(2.5 * fp), max_worlds=16, max_world_bodies=256)
    rids = [srv.submit(sc, 6) for _ in range(5)]
    seen_live = []
    finished = []
    for _ in range(100):
        live, _ = srv.tick()
        seen_live.append(live)
        finished += srv.poll()
        if len(finished) == 5:
            break
    assert sorted(finished) == sorted(rids)
    assert max(seen_live) <= 2, seen_live


def test_server_cancel_leaves_others_unchanged():
    from pe_amd.server import Server
    reqs = _reqs(4, seed=91)
    srv = Server(hbm_budget_bytes=1 << 30, max_worlds=16)
    ids = [srv.submit(sc, h) for sc, h in reqs]
    srv.tick()
    srv.cancel(ids[1])
    res = {}
    for _ in range(100):
        srv.tick()
        for rid in srv.poll():
            res[rid] = flat_state(srv.result(rid))
        if len(res) == 3:
            break
    for k in (0, 2, 3):
        assert np.array_equal(res[ids[k]], _alone(*reqs[k])), f"request {k}"


def test_server_shape_mix_equals_alone():
    """Requests holding cylinders, convex hulls and a heightfield (their meshes placed in the
    server's mesh arena, the ranges of a cancelled and of finished requests reused by later
    ones) end bit-identical to the same scene stepped alone."""
    from pe_amd import scenes
    from pe_amd.server import Server
    sizes = [40, 90, 25, 120, 60, 35, 75]
    reqs = [(scenes.shape_mix_scene(n, seed=300 + i, volume=(8, 6, 8), grid=8, ground_size=(10, 10), hull_pool=4),
             int(h)) for i, (n, h) in enumerate(zip(sizes, [12, 30, 8, 20, 16, 10, 14]))]
    print(f"[server shapes] sizes {[r[0].n for r in reqs]} meshes {[len(r[0].meshes) for r in reqs]}", flush=True)
    srv = Server(hbm_budget_bytes=1 << 30, max_worlds=8)
    ids, done = {}, {}
    pending = list(enumerate(reqs))
    victim = srv.submit(reqs[1][0], 200)   # cancelled while live: its ranges are reused
    for t in range(300):
        if pending and t % 3 == 0:
            i, (sc, h) = pending.pop(0)
            ids[srv.submit(sc, h)] = i
        if t == 4:
            srv.cancel(victim)
        srv.tick()
        for rid in srv.poll():
            if rid in ids:
                done[ids[rid]] = flat_state(srv.result(rid))
        if len(done) == len(reqs):
            break
    assert len(done) == len(reqs)
    for i, (sc, h) in enumerate(reqs):
        assert np.array_equal(done[i], _alone(sc, h)), f"request {i}"
