"""Parity of the HIP narrowphase for cylinders, convex hulls and a triangle-mesh heightfield
(every shape pair type) against the CPU oracle, which is pinned bit-exact to the compiled
reference (test_oracle_vs_ref.py) and to golden fixtures (test_golden.py).

Tolerances (written here, applied below):
  integrate / AABB / pair list / colour classes ........... bit-exact
  contact points (normals, depths, positions) ............. 1e-5 relative (abs floor 1e-6)
  velocities after one solve .............................. 1e-5 relative (abs floor 1e-6)
  accumulated impulses after one solve .................... 1e-5 relative (floor 1e-6 x the step's largest)
  60-step trajectories .................................... position RMS <= 1e-3 m, max <= 1e-2 m,
                                                            orientation max <= 1e-3 rad
"""
import numpy as np
import pytest

from helpers import brute_force_pairs, flat_state, impulse_error, pair_set, rotation_angle

pytestmark = pytest.mark.gpu

REL = 1e-5
ABS = 1e-6


def _close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


def shape_scene(n=300, seed=9):
    from pe_amd import scenes
    return scenes.shape_mix_scene(n, seed=seed, volume=(12, 8, 12), grid=16, ground_size=(16, 16))


def test_shape_stages_match_oracle():
    """Lock-step run of 140 steps; every stage compared on every step once bodies touch."""
    import pe_amd
    from pyoracle import OracleWorld
    sc = shape_scene()
    g, o = pe_amd.World(sc), OracleWorld(sc, 32)
    T = sc.shape_type
    seen = set()
    worst = 0.0
    print(f"[shapes] n={sc.n} meshes={len(sc.meshes)}", flush=True)
    for step in range(140):
        g.integrate(); o.integrate()
        g.sync()
        assert np.array_equal(flat_state(g.state()), flat_state(o.state()).astype(np.float32)), f"state step {step}"
        gmn, gmx = g.aabbs()
        omn, omx = o.aabbs()
        assert np.array_equal(gmn, omn.astype(np.float32)) and np.array_equal(gmx, omx.astype(np.float32)), step
        g.broadphase(); o.broadphase()
        gp, op = g.pairs(), o.pairs()
        assert np.array_equal(gp, op), f"pair order step {step}"
        if step % 20 == 0:
            assert pair_set(gp) == brute_force_pairs(gmn, gmx, sc.kinematic)
        assert g.axis() == o.axis()
        g.narrowphase(); o.narrowphase()
        gn, gpts = g.contacts()
        on, _, opts = o.contacts()
        bad = np.nonzero(gn != on)[0]
        assert len(bad) == 0, f"point counts step {step}: pairs {bad[:5]} types {[tuple(T[gp[k]]) for k in bad[:5]]}"
        vm = np.arange(8)[None, :] < on[:, None]
        assert _close(gpts[vm], opts[vm]), f"contact points step {step}"
        assert _close(gpts[..., 12], opts[..., 12]), f"slot distances step {step}"
        for k in np.nonzero(on)[0]:
            seen.add(tuple(sorted((int(T[gp[k][0]]), int(T[gp[k][1]])))))
        g.solve(); o.solve()
        g.sync()
        order, cs = g.solve_order()
        gcol = np.full(len(gp), -1)
        for c in range(len(cs) - 1):
            gcol[order[cs[c]:cs[c + 1]]] = c
        assert np.array_equal(gcol, o.color_of()), f"colour classes step {step}"
        _, _, gv, gw = g.state()
        _, _, ov, ow = o.state()
        assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}"
        err, rows = impulse_error(g.impulses(), o.impulses(), on)
        worst = max(worst, err)
        assert err <= 1e-5, f"impulses step {step}: max relative error {err:.3e} over {rows} rows"
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    print(f"[shapes] contact pair types seen: {sorted(seen)}; largest relative impulse error {worst:.3e}", flush=True)
    # the heightfield against every other shape, and the polyhedral pairs among bodies
    for t in (0, 1, 2, 3):
        assert (t, 4) in seen
    # the pairs solved one wave each: by the split separating-axis test (cylinder-cylinder,
    # box-hull, hull-hull) and by the mesh faces tested over the lanes (sphere or cylinder with a
    # hull or the heightfield); the scene gives all 14 pair types contacts within these steps
    for t in ((2, 2), (0, 3), (3, 3), (1, 3), (2, 3), (1, 4), (2, 4)):
        assert t in seen
    assert len(seen) == 14


def test_shape_trajectory_60_steps():
    import pe_amd
    from pyoracle import OracleWorld
    sc = shape_scene(200, seed=5)
    g, o = pe_amd.World(sc), OracleWorld(sc, 32)
    g.step(110)
    o.step(110)
    g.sync()
    pg, rg = g.state()[:2]
    po, ro = o.state()[:2]
    d = np.linalg.norm(pg - po, axis=1)
    rms = float(np.sqrt(np.mean(d ** 2)))
    ang = rotation_angle(rg, ro)
    print(f"[shapes traj] position rms {rms:.3e} max {d.max():.3e}, orientation max {ang.max():.3e} rad", flush=True)
    assert rms <= 1e-3 and d.max() <= 1e-2
    assert ang.max() <= 1e-3


def test_shape_worlds_batched_equal_alone():
    """Shape-mix worlds stepped in one batch are bit-identical to each stepped alone."""
    import pe_amd
    ws = [shape_scene(60 + 20 * k, seed=40 + k) for k in range(4)]
    b = pe_amd.Batch(ws)
    b.step(90)
    b.sync()
    for k, sc in enumerate(ws):
        g = pe_amd.World(sc)
        g.step(90)
        g.sync()
        a = g.state()
        bb = b.world_state(k)
        for x, y in zip(a, bb):
            assert np.array_equal(x, y), f"world {k}"


def long_walk_scene():
    """Spheres and cylinders whose BVH walks find more faces than a wave has lanes: a 1 m sphere
    and a 0.8 m cylinder on a fine heightfield (0.15 m cells: hundreds of triangles under each), a
    lying cylinder, a box and a sphere on a kinematic hull of more than 64 vertices (its
    separating-axis projections read the mesh from global memory rather than a wave's LDS copy)."""
    from pe_amd import scenes
    rng = np.random.default_rng(77)
    b = scenes.SceneBuilder()
    hf = b.add_mesh(scenes.heightfield_mesh(rng, 40, 40, (6.0, 6.0), 0.05))
    b.add(scenes.SHAPE_CONCAVE, (), 1.0, (0.0, 0.0, 0.0), kin=1, mesh=hf)
    pts = rng.standard_normal((120, 3))
    big = scenes.hull_mesh(1.2 * pts / np.linalg.norm(pts, axis=1)[:, None])   # every point a vertex
    assert len(big.vertices) > 64
    vol = abs(sum(np.dot(big.vertices[f[0]], np.cross(big.vertices[f[1]], big.vertices[f[2]])) for f in big.faces)) / 6.0
    b.add(scenes.SHAPE_CONVEX, (), 1000.0 * vol, (-1.6, 1.25, 0.0), kin=1, mesh=b.add_mesh(big))
    b.add(scenes.SHAPE_SPHERE, (1.0,), 4188.8, (1.5, 1.1, 1.5))
    b.add(scenes.SHAPE_CYLINDER, (0.8, 0.8), 1608.5, (1.5, 0.5, -1.5))
    lying = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])   # axis along -x
    b.add(scenes.SHAPE_CYLINDER, (0.5, 1.6), 1256.6, (-1.6, 3.0, 0.0), lying)
    b.add(scenes.SHAPE_BOX, (0.6, 0.6, 0.6), 216.0, (-1.0, 3.2, 0.8))
    b.add(scenes.SHAPE_SPHERE, (0.5,), 523.6, (-1.6, 3.2, -1.05))
    return b.scene(name="long_walks")


def test_mesh_walks_longer_than_a_wave_match_oracle():
    """The wave routines hand a walk's faces to the lanes 64 at a time; walks of hundreds of faces
    take several batches and must add their points in the sequential walk's order."""
    import pe_amd
    from pyoracle import OracleWorld
    sc = long_walk_scene()
    g, o = pe_amd.World(sc), OracleWorld(sc, 32)
    T = sc.shape_type
    seen = set()
    print(f"[long walks] n={sc.n} meshes={len(sc.meshes)}", flush=True)
    for step in range(90):
        g.integrate(); o.integrate()
        g.broadphase(); o.broadphase()
        gp, op = g.pairs(), o.pairs()
        assert np.array_equal(gp, op), f"pair order step {step}"
        g.narrowphase(); o.narrowphase()
        gn, gpts = g.contacts()
        on, _, opts = o.contacts()
        bad = np.nonzero(gn != on)[0]
        assert len(bad) == 0, f"point counts step {step}: pairs {bad[:5]} types {[tuple(T[gp[k]]) for k in bad[:5]]}"
        vm = np.arange(8)[None, :] < on[:, None]
        assert _close(gpts[vm], opts[vm]), f"contact points step {step}"
        for k in np.nonzero(on)[0]:
            seen.add(tuple(sorted((int(T[gp[k][0]]), int(T[gp[k][1]])))))
        g.solve(); o.solve()
        g.sync()
        _, _, gv, gw = g.state()
        _, _, ov, ow = o.state()
        assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}"
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    print(f"[long walks] contact pair types seen: {sorted(seen)}", flush=True)
    for t in ((1, 4), (2, 4), (0, 3), (1, 3), (2, 3)):
        assert t in seen

