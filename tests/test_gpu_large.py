"""Parity at the sizes the benchmark runs: batched worlds of 161-512 bodies (the two longest solver
launches of the 4096-world batch), a served request of thousands of bodies, and worlds whose
solver does not fit LDS (more than 2 pairs per body, or too many bodies), which take the
global-memory solve path.  Each world must be bit-identical to the same world stepped alone on
the single-world path, and the single-world path must track the CPU oracle in lock-step
(tolerances as in test_gpu_world.py: state / pairs / colour classes bit-exact; contact points,
velocities and impulses 1e-5 relative)."""
import math

import numpy as np
import pytest

from helpers import flat_state, impulse_error

pytestmark = pytest.mark.gpu

REL = 1e-5
ABS = 1e-6
# the schedule the 4096-world benchmark runs (class upper bounds, block sizes, solver group bodies)
from pe_amd.batch import load_schedule  # noqa: E402

_SCHED = load_schedule("batched_4096_worlds")
BENCH_UPPER = _SCHED["class_upper_bodies"]
BENCH_BLOCK = _SCHED["block"]
BENCH_GROUP = _SCHED["solver_group_bodies"]


def _close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


def dense_block(k=7, seed=21, gap=1.01):
    """k^3 unit boxes on a lattice of spacing gap, each turned by up to 3 degrees about a random
    axis, resting on a kinematic ground: about 4 AABB pairs per body (5 to 8 in the first steps),
    beyond the 2 per body the solver's LDS layout holds, so the solve runs on the global-memory
    path."""
    from pe_amd.scenes import SHAPE_BOX, SceneBuilder
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    b.add(SHAPE_BOX, (40.0, 1.0, 40.0), 1.0, (0.0, -0.5, 0.0), kin=1)
    for i in range(k):
        for j in range(k):
            for m in range(k):
                ax = rng.normal(size=3)
                ax /= np.linalg.norm(ax)
                ang = math.radians(rng.uniform(-3.0, 3.0))
                K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
                R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
                pos = ((i - k / 2) * gap, 0.56 + j * gap, (m - k / 2) * gap)
                b.add(SHAPE_BOX, (1.0, 1.0, 1.0), 1000.0, pos, rot=R, mu=0.5, e=0.0)
    return b.scene(name=f"dense_block{k}")


def _lockstep(sc, steps, label=""):
    """The single-world path and the oracle in lock-step for `steps` steps, every stage compared
    on every step (the oracle restarts from the device state after each step and keeps its own
    contact cache, as the device does)."""
    import pe_amd
    from pyoracle import OracleWorld
    g, o = pe_amd.World(sc), OracleWorld(sc, 32)
    worst = 0.0
    for step in range(steps):
        g.integrate(); o.integrate()
        g.sync()
        assert np.array_equal(flat_state(g.state()), flat_state(o.state()).astype(np.float32)), f"{label} state {step}"
        g.broadphase(); o.broadphase()
        gp = g.pairs()
        assert np.array_equal(gp, o.pairs()), f"{label} pairs {step}"
        g.narrowphase(); o.narrowphase()
        gn, gpts = g.contacts()
        on, _, opts = o.contacts()
        assert np.array_equal(gn, on), f"{label} point counts {step}"
        vm = np.arange(8)[None, :] < on[:, None]
        assert _close(gpts[vm], opts[vm]), f"{label} contact points {step}"
        g.solve(); o.solve()
        g.sync()
        order, cs = g.solve_order()
        gcol = np.full(len(gp), -1)
        for c in range(len(cs) - 1):
            gcol[order[cs[c]:cs[c + 1]]] = c
        assert np.array_equal(gcol, o.color_of()), f"{label} colour classes {step}"
        _, _, gv, gw = g.state()
        _, _, ov, ow = o.state()
        assert _close(gv, ov) and _close(gw, ow), f"{label} velocities {step}"
        err, _ = impulse_error(g.impulses(), o.impulses(), on)
        worst = max(worst, err)
        assert err <= REL, f"{label} impulses {step}: {err:.3e}"
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    c = g.counts()
    g.close()
    print(f"[lockstep] {label}: n={sc.n} pairs {c['pairs']} points {c['points']} colours {c['colors']}, "
          f"largest relative impulse error {worst:.3e}", flush=True)
    return c


def test_batch_161_512_classes_equal_single_world_and_oracle():
    """Worlds of 161-512 bodies from the 4096-world benchmark's distribution (boxes and spheres on
    a jittered lattice), stepped 160 steps in one batch at the default configuration and at the
    benchmark's tuned one (graph-replayed): every world bit-identical to the single-world path;
    the largest world in lock-step with the oracle over all 160 steps."""
    import pe_amd
    from pe_amd import scenes
    sizes = [165, 230, 300, 384, 385, 450, 512]
    ws = [scenes.grid_world(n, scenes.world_seed(1003, 900 + k)) for k, n in enumerate(sizes)]
    print(f"[batch 161-512] sizes {[w.n for w in ws]}", flush=True)
    a = pe_amd.Batch(ws)
    a.step(160)
    a.sync()
    b = pe_amd.Batch(ws)
    b.set_config(BENCH_UPPER, BENCH_BLOCK)
    b.set_solver_groups(BENCH_GROUP)
    b.capture_graph()
    b.step(150)
    b.sync()
    b.step(10)
    b.sync()
    assert a.counts()["overflow_worlds"] == 0 and b.counts()["overflow_worlds"] == 0
    for k, sc in enumerate(ws):
        w = pe_amd.World(sc)
        w.step(160)
        w.sync()
        single = flat_state(w.state())
        w.close()
        assert np.array_equal(flat_state(a.world_state(k)), single), f"default config, world {k} (n={sc.n})"
        assert np.array_equal(flat_state(b.world_state(k)), single), f"benchmark config, world {k} (n={sc.n})"
    c = _lockstep(ws[-1], 160, label="512-body world")
    assert c["points"] > 0


def test_global_solve_path_batch_and_oracle():
    """A dense block (about 4 pairs per body) takes the global-memory solve path inside the LDS
    kernel build; a 2,500-body world is too large for the solver's LDS at all and launches the
    unconstrained build.  Both, batched with small worlds, are bit-identical to the single-world
    path; the dense block tracks the oracle in lock-step."""
    import pe_amd
    from pe_amd import scenes
    blk = dense_block(7)
    big = scenes.grid_world(2500, 4242)
    small = [scenes.grid_world(n, 77 + n) for n in (20, 90, 200)]
    ws = [blk, big] + small
    print(f"[global path] sizes {[w.n for w in ws]}", flush=True)
    for cfg in (None, "bench"):
        b = pe_amd.Batch(ws)
        if cfg:
            b.set_config(BENCH_UPPER, BENCH_BLOCK)
            b.set_solver_groups(BENCH_GROUP)
            b.capture_graph()
        b.step(40)
        b.sync()
        cnt = b.counts()
        assert cnt["overflow_worlds"] == 0
        for k, sc in enumerate(ws):
            w = pe_amd.World(sc)
            w.step(40)
            w.sync()
            if k == 0:
                wc = w.counts()
                assert wc["pairs"] > 2 * sc.n, wc   # really beyond the LDS layout's 2 pairs per body
            assert np.array_equal(flat_state(b.world_state(k)), flat_state(w.state())), f"config {cfg}, world {k}"
            w.close()
        b.close()
    _lockstep(blk, 40, label="dense block")


def test_server_large_request_equals_alone():
    """A 3,000-body mixed-shape request (over its heightfield; solved on the global path) and a
    dense block served beside smaller requests end bit-identical to each scene stepped alone."""
    import pe_amd
    from pe_amd import scenes
    from pe_amd.server import Server
    reqs = [(scenes.served_shape_world(3000, 5151), 25), (dense_block(5, seed=3), 30),
            (scenes.served_shape_world(60, 5152), 12), (scenes.grid_world(300, 5153), 20)]
    print(f"[server large] sizes {[r[0].n for r in reqs]}", flush=True)
    srv = Server(hbm_budget_bytes=8 << 30, max_worlds=32, max_world_bodies=4096)
    ids, done = {}, {}
    pending = list(enumerate(reqs))
    for t in range(200):
        if pending and t % 2 == 0:
            i, (sc, h) = pending.pop(0)
            ids[srv.submit(sc, h)] = i
        srv.tick()
        for rid in srv.poll():
            done[ids[rid]] = flat_state(srv.result(rid))
        if len(done) == len(reqs):
            break
    assert len(done) == len(reqs)
    rs = srv.request_stats()
    print(f"[server large] request stats {rs}", flush=True)
    assert rs["finished"] == len(reqs) and rs["capacity_errors"] == 0
    for i, (sc, h) in enumerate(reqs):
        w = pe_amd.World(sc)
        w.step(h)
        w.sync()
        assert np.array_equal(done[i], flat_state(w.state())), f"request {i} (n={sc.n})"
        w.close()
    srv.close()
