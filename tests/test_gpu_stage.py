"""The stage entry points (pe_stage_broadphase / narrowphase / solve, on caller device buffers)
against the single-world pipeline: fed the records a World's integrate stage wrote, each stage
returns exactly (bit for bit) what the World's own stage computes -- the same kernels behind a
different boundary -- step after step, with the caller carrying the contact slots from one
narrowphase to the next as the reference carries its ContactResult objects."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mesh_table(sc):
    """The world's mesh ids: one per (scene mesh, concave) in order of first use by a body."""
    from pe_amd.scenes import SHAPE_CONCAVE, SHAPE_CONVEX
    ids, meshes, concave = {}, [], []
    for i in range(sc.n):
        t = int(sc.shape_type[i])
        if t in (SHAPE_CONVEX, SHAPE_CONCAVE):
            key = (int(sc.mesh_index[i]), t == SHAPE_CONCAVE)
            if key not in ids:
                ids[key] = len(meshes)
                meshes.append(sc.meshes[key[0]])
                concave.append(1 if key[1] else 0)
    return meshes, concave


def _scenes():
    from helpers import mixed_small, small_rain
    from pe_amd import scenes
    return {"rain300": small_rain(300), "mixed200": mixed_small(200),
            "shapes": scenes.shape_mix_scene(200, seed=9, volume=(12, 8, 12), grid=16, ground_size=(16, 16))}


@pytest.mark.parametrize("name", ["rain300", "mixed200", "shapes"])
def test_stages_equal_world_stages(name):
    import pe_amd
    from pe_amd.hip import DeviceArray, synchronize
    from pe_amd.stage import Stage
    sc = _scenes()[name]
    print(f"[stage] scene={name} n={sc.n}", flush=True)
    w = pe_amd.World(sc)
    n = sc.n
    max_pairs = 16 * n + 1024
    st = Stage(n, max_pairs)
    meshes, concave = _mesh_table(sc)
    if meshes:
        st.set_meshes(meshes, concave)
    points = DeviceArray.from_numpy(np.zeros((max_pairs, 8, 13), np.float32))
    prev = 0
    checked = 0
    for step in range(70):
        axis = w.axis()
        w.integrate()
        w.sync()
        aabb, body, solver = (DeviceArray.from_numpy(x) for x in w.records())
        # stage 1
        pa, pb, res = st.broadphase(aabb, axis)
        npg = w.broadphase()
        synchronize()
        r = res.numpy()
        gp = w.pairs()
        assert r[0] == npg and r[2] == 0, (step, r, npg)
        assert r[1] == w.axis(), step
        sp = np.stack([pa.numpy()[:npg], pb.numpy()[:npg]], 1)
        assert np.array_equal(sp, gp), f"pairs step {step}"
        # stage 2 (the slots carry the previous step's distances in)
        npts, swap, res2 = st.narrowphase(body, pa, pb, npg, prev, points)
        w.narrowphase()
        synchronize()
        gn, gpts = w.contacts()
        assert int(res2.numpy()[0]) == 0
        assert np.array_equal(npts.numpy()[:npg], gn), f"point counts step {step}"
        assert np.array_equal(points.numpy()[:npg], gpts), f"contact slots step {step}"
        prev = npg
        # stage 3
        imp = DeviceArray.from_numpy(np.zeros((max(npg, 1), 8, 3), np.float32))
        res3 = st.solve(body, solver, pa, pb, npg, npts, points, sc.dt, sc.iterations, imp)
        w.solve()
        w.sync()
        synchronize()
        _, _, gv, gw = w.state()
        sv = solver.numpy()
        assert np.array_equal(sv[:, 17:20], gv) and np.array_equal(sv[:, 20:23], gw), f"velocities step {step}"
        assert np.array_equal(imp.numpy()[:npg], w.impulses()), f"impulses step {step}"
        assert int(res3.numpy()[0]) == w.counts()["colors"]
        checked += int(gn.sum() > 0)
    print(f"[stage] {name}: {checked} steps with contacts, last step pairs {prev}", flush=True)
    assert checked > 10
    st.close()
    w.close()


def test_stage_broadphase_small_and_flags():
    """n < 2 leaves the pairs alone and keeps the axis; ignore-collision and both-kinematic bodies
    are in no pair; a full output reports the overflow with the true pair count.  On a stream of
    the caller's."""
    from pe_amd.hip import DeviceArray, Stream
    from pe_amd.stage import Stage
    st = Stage(64, 64)
    s = Stream()
    one = DeviceArray.from_numpy(np.array([[0, 0, 0, 0, 1, 1, 1, 0]], np.float32))
    pa = DeviceArray.full(4, -7, np.int32)
    pb = DeviceArray.full(4, -7, np.int32)
    _, _, res = st.broadphase(one, 2, 4, pa, pb, stream=s)
    s.synchronize()
    assert res.numpy().tolist() == [0, 2, 0] and pa.numpy().tolist() == [-7] * 4
    # five unit boxes overlapping in a row along x; body 1 ignores collisions, 3 and 4 kinematic
    rec = []
    for i in range(5):
        rec.append([i * 0.9, 0, 0, 1.0 if i >= 3 else 0.0, i * 0.9 + 1, 1, 1, 1.0 if i == 1 else 0.0])
    aabb = DeviceArray.from_numpy(np.array(rec, np.float32))
    pa, pb, res = st.broadphase(aabb, 0, 8, stream=s)
    s.synchronize()
    k = int(res.numpy()[0])
    got = sorted(zip(pa.numpy()[:k].tolist(), pb.numpy()[:k].tolist()))
    assert got == [(2, 3)], got        # (0,1), (1,2) ignored; (3,4) both kinematic
    big = DeviceArray.from_numpy(np.array([[0, 0, 0, 0, 1, 1, 1, 0]] * 20, np.float32))
    pa, pb, res = st.broadphase(big, 1, 50, stream=s)
    s.synchronize()
    r = res.numpy().tolist()
    assert r[0] == 190 and r[2] & 1, r   # 20 * 19 / 2 pairs found, 50 written
    st.close()
    s.close()
