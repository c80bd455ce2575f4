"""Batched stepping: every world in a batch must equal the same world stepped alone by the
single-world path (bit-exact), whatever else is in the batch and whatever the launch
configuration or graph replay, and must track the CPU oracle like the single-world path."""
import numpy as np
import pytest

from helpers import flat_state

pytestmark = pytest.mark.gpu


def _worlds(k=12, seed=1003):
    from pe_amd import scenes
    rng = np.random.default_rng(seed)
    counts = scenes.log_uniform_counts(rng, k, 8, 160)
    return [scenes.mixed_world(int(c), scenes.world_seed(seed, w)) for w, c in enumerate(counts)]


@pytest.fixture(scope="module")
def worlds():
    return _worlds()


def test_batch_equals_single_world(worlds):
    import pe_amd
    from pe_amd.batch import Batch
    print(f"[batch-vs-single] {len(worlds)} worlds, sizes {[w.n for w in worlds]}", flush=True)
    b = Batch(worlds)
    b.step(30)
    b.sync()
    assert b.counts()["overflow_worlds"] == 0
    for i, sc in enumerate(worlds):
        w = pe_amd.World(sc)
        w.step(30)
        w.sync()
        assert np.array_equal(flat_state(b.world_state(i)), flat_state(w.state())), f"world {i}"


def test_batch_graph_and_config_do_not_change_results(worlds):
    from pe_amd.batch import Batch
    a = Batch(worlds)
    a.step(20)
    a.sync()
    b = Batch(worlds)
    b.set_config([16, 48, 100, 4096], [64, 128, 512, 256])
    b.capture_graph()
    b.step(5)
    b.sync()          # eager work between replays
    _ = b.counts()
    b.step(15)
    b.sync()
    for i in range(len(worlds)):
        assert np.array_equal(flat_state(a.world_state(i)), flat_state(b.world_state(i))), f"world {i}"
    # the solver grouping (one world per workgroup, or several solved as one system) does not
    # change a result either
    for rows in ([0, 0, 0, 0], [2048, 2048, 2048, 2048], [64, 300, 0, 1000]):
        c = Batch(worlds)
        c.set_config([16, 48, 100, 4096], [64, 128, 512, 256])
        c.set_solver_groups(rows)
        # class-by-class timed steps, then steps on the streams their costs dealt them to
        t = c.class_times(3)
        assert len(t) == 4 and all(k1 >= 0 and k3 >= 0 for k1, k3 in t)
        c.step(17)
        c.sync()
        for i in range(len(worlds)):
            assert np.array_equal(flat_state(a.world_state(i)), flat_state(c.world_state(i))), f"world {i} rows {rows}"


def test_batch_snapshot_restore_and_autotune(worlds):
    """A restored snapshot steps exactly like the state it saved, and the tuner (which times
    every candidate from a snapshot) leaves the worlds where it found them."""
    from pe_amd.batch import Batch
    a = Batch(worlds)
    a.step(20)
    a.sync()
    b = Batch(worlds)
    b.step(12)
    b.snapshot()
    b.step(8)
    b.sync()
    for i in range(len(worlds)):
        assert np.array_equal(flat_state(a.world_state(i)), flat_state(b.world_state(i))), f"world {i}"
    b.step(5)
    b.restore()
    assert b.steps_done == 12
    b.step(8)
    b.sync()
    for i in range(len(worlds)):
        assert np.array_equal(flat_state(a.world_state(i)), flat_state(b.world_state(i))), f"restored world {i}"
    c = Batch(worlds)
    c.step(12)
    tuned = c.autotune(candidates=(64, 256), steps=1, group_candidates=(0, 256))
    print(f"[snapshot] tuned {tuned}", flush=True)
    assert c.steps_done == 12
    c.step(8)
    c.sync()
    for i in range(len(worlds)):
        assert np.array_equal(flat_state(a.world_state(i)), flat_state(c.world_state(i))), f"tuned world {i}"


def test_batch_subset_alone_vs_in_mix(worlds):
    """A world's result does not depend on its neighbours in the batch."""
    from pe_amd.batch import Batch
    full = Batch(worlds)
    full.step(25)
    full.sync()
    sub_idx = [1, 4, 7]
    sub = Batch([worlds[i] for i in sub_idx])
    sub.step(25)
    sub.sync()
    for k, i in enumerate(sub_idx):
        assert np.array_equal(flat_state(full.world_state(i)), flat_state(sub.world_state(k)))


def test_batch_tracks_oracle(worlds):
    from pe_amd.batch import Batch
    from pyoracle import OracleWorld
    b = Batch(worlds[:4])
    b.step(20)
    b.sync()
    for i in range(4):
        o = OracleWorld(worlds[i], 32)
        o.step(20)
        pg = b.world_state(i)[0]
        po = o.state()[0]
        d = np.linalg.norm(pg - po, axis=1)
        assert np.sqrt(np.mean(d ** 2)) <= 1e-3 and d.max() <= 1e-2, f"world {i}"
