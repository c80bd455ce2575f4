"""Capacity errors are reported, never swallowed: a world whose broadphase finds more pairs than its
capacity has its extra pairs dropped (the reference never truncates,
broad_phase_sweep_and_prune.cpp:19-34), so the step's result must come back as PE_ERR_CAPACITY --
from pe_world_sync on the plain path (also when only an early step of a multi-step call
overflowed), from pe_batch_sync for a batch, and from pe_server_result for the request that
overflowed -- while the worlds beside it are unaffected."""
import numpy as np
import pytest

from helpers import flat_state

pytestmark = pytest.mark.gpu

PE_ERR_CAPACITY = -3


def cluster(n=40, r=0.25, spread=0.3, speed=12.0, seed=5):
    """n spheres packed in a ball of radius `spread` 5 m above a kinematic ground, each moving
    radially outwards at `speed`: many pairs at the first step, few a handful of steps later."""
    from pe_amd.scenes import SHAPE_BOX, SHAPE_SPHERE, SceneBuilder
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    b.add(SHAPE_BOX, (40.0, 1.0, 40.0), 1.0, (0.0, -0.5, 0.0), kin=1)
    for _ in range(n):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p = np.array([0.0, 5.0, 0.0]) + d * spread * rng.random() ** (1.0 / 3.0)
        b.add(SHAPE_SPHERE, (r,), 100.0, p, v=tuple(d * speed), mu=0.5, e=0.0)
    return b.scene(name="cluster")


def _pair_counts(sc, steps):
    from pyoracle import OracleWorld
    o = OracleWorld(sc, 32)
    out = []
    for _ in range(steps):
        o.step(1)
        out.append(len(o.pairs()))
    return out


def test_plain_overflow_in_an_early_step_is_reported_at_sync():
    import pe_amd
    from pe_amd._lib import PEError
    sc = cluster()
    counts = _pair_counts(sc, 5)
    cap = 200
    print(f"[overflow plain] pairs per step {counts}, max_pairs {cap}", flush=True)
    assert counts[0] > cap and counts[-1] <= cap     # only the first steps overflow
    w = pe_amd.World(sc, max_pairs=cap)
    w.step(5)
    with pytest.raises(PEError) as e:
        w.sync()
    assert e.value.rc == PE_ERR_CAPACITY and "max_pairs" in str(e.value)
    w.sync()           # reported once, then cleared
    w.step(1)
    w.sync()           # the later steps fit
    # the broadphase stage entry point reports it too
    w2 = pe_amd.World(sc, max_pairs=cap)
    w2.integrate()
    with pytest.raises(PEError) as e:
        w2.broadphase()
    assert e.value.rc == PE_ERR_CAPACITY
    w.close()
    w2.close()


def test_batch_overflow_is_reported_and_others_unaffected():
    import pe_amd
    from pe_amd import scenes
    from pe_amd._lib import PEError
    dense = cluster(r=0.5, speed=0.0)               # 780 pairs among 41 bodies: beyond 16 n + 64 = 720
    assert _pair_counts(dense, 1)[0] > 16 * dense.n + 64
    ok = [scenes.grid_world(60, 31), scenes.grid_world(150, 32)]
    b = pe_amd.Batch([ok[0], dense, ok[1]])
    b.step(3)
    with pytest.raises(PEError) as e:
        b.sync()
    assert e.value.rc == PE_ERR_CAPACITY and "worlds 1" in str(e.value)
    assert b.counts()["overflow_worlds"] == 1
    b.sync()           # reported once
    for k, i in ((0, 0), (1, 2)):
        w = pe_amd.World(ok[k])
        w.step(3)
        w.sync()
        assert np.array_equal(flat_state(b.world_state(i)), flat_state(w.state())), f"world {i}"
        w.close()
    b.close()


def test_server_overflowing_request_gets_a_capacity_error():
    import pe_amd
    from pe_amd import scenes
    from pe_amd._lib import PEError
    from pe_amd.server import Server
    dense = cluster(r=0.5, speed=0.0)
    ok = scenes.grid_world(80, 33)
    srv = Server(hbm_budget_bytes=1 << 30, max_worlds=8)
    bad_id = srv.submit(dense, 4)
    ok_id = srv.submit(ok, 6)
    got = {}
    for _ in range(40):
        srv.tick()
        for rid in srv.poll():
            try:
                got[rid] = flat_state(srv.result(rid))
            except PEError as e:
                got[rid] = e.rc
        if len(got) == 2:
            break
    assert got[bad_id] == PE_ERR_CAPACITY
    w = pe_amd.World(ok)
    w.step(6)
    w.sync()
    assert np.array_equal(got[ok_id], flat_state(w.state()))
    assert srv.request_stats()["capacity_errors"] == 1
    w.close()
    srv.close()


def test_server_retires_every_finished_world_in_one_tick():
    """More finished worlds in one tick than the result buffer holds at once (64 of the largest
    admissible size): every one is retired at its horizon, in rounds, and equals its scene stepped
    alone for exactly that many steps.  Uploads stage 8 of the largest requests per tick, so request
    k is admitted at tick k // 8 + 1; horizons 20 - k // 8 make all 70 finish at tick 20."""
    import pe_amd
    from pe_amd import scenes
    from pe_amd.server import Server
    reqs = [scenes.grid_world(16, 600 + k) for k in range(70)]
    hor = [20 - k // 8 for k in range(70)]
    srv = Server(hbm_budget_bytes=1 << 30, max_worlds=128, max_world_bodies=17)
    ids = {srv.submit(sc, h): k for k, (sc, h) in enumerate(zip(reqs, hor))}
    got = {}
    per_tick = []
    for _ in range(30):
        srv.tick()
        fin = srv.poll()
        per_tick.append(len(fin))
        for rid in fin:
            got[ids[rid]] = flat_state(srv.result(rid))
        if len(got) == len(reqs):
            break
    print(f"[retire] finished per tick {per_tick}", flush=True)
    assert len(got) == len(reqs)
    assert max(per_tick) > 64
    for k in (0, 33, 63, 64, 69):
        w = pe_amd.World(reqs[k])
        w.step(hor[k])
        w.sync()
        assert np.array_equal(got[k], flat_state(w.state())), f"request {k}"
        w.close()
    srv.close()
