"""Pins the oracle (C++ restatement) to the reference engine compiled from its own sources
(oracle/_ref).  FP32 and FP64: every stage bit-exact.
The reference sorts with std::sort, whose order among equal keys is unspecified; where its
pair order differs from the restatement's (ties), the reference's order is fed to the oracle
so both continue from the same pair list.  Skipped where oracle/_ref is not built."""
import numpy as np
import pytest

from helpers import flat_state, ortho_units, pair_set, reference_sample_scene, small_rain
from pyoracle import OracleWorld, RefWorld

pytestmark = pytest.mark.ref


def _run(sc, bits, steps, exact):
    o = OracleWorld(sc, bits, color_order=False)
    r = RefWorld(sc, bits, threads=4)
    im_o, li_o = o.mass_props()
    im_r, li_r = r.mass_props()
    assert np.array_equal(im_o, im_r) and np.array_equal(li_o, li_r)
    eq = (lambda a, b: np.array_equal(a, b)) if exact else (lambda a, b: np.allclose(a, b, rtol=1e-12, atol=1e-12))
    reorders = 0
    for step in range(steps):
        o.integrate(); r.integrate()
        assert eq(flat_state(o.state()), flat_state(r.state())), f"integrate step {step}"
        assert eq(np.concatenate(o.aabbs(), 1), np.concatenate(r.aabbs(), 1)), f"aabb step {step}"
        o.broadphase(); r.broadphase()
        po, pr = o.pairs(), r.pairs()
        assert pair_set(po) == pair_set(pr), f"pair set step {step}"
        assert o.axis() == r.axis()
        if not np.array_equal(po, pr):
            reorders += 1
            o.set_pairs(pr)
        o.narrowphase(); r.narrowphase()
        (no, so, pto), (nr, sr, ptr) = o.contacts(), r.contacts()
        assert np.array_equal(no, nr) and np.array_equal(so, sr), f"point counts step {step}"
        vm = ptr[..., 0] < 1e29
        assert np.array_equal(pto[..., 0] < 1e29, vm)
        assert eq(pto[vm], ptr[vm]) and eq(pto[..., 12], ptr[..., 12]), f"contacts step {step}"
        o.solve(); r.solve()
        if exact:
            assert eq(flat_state(o.state()), flat_state(r.state())), f"solve step {step}"
        else:
            o.set_state(*r.state())
    return reorders


@pytest.mark.parametrize("name", ["tower", "rain"])
def test_fp32_bit_exact(name):
    from pe_amd import scenes
    sc = scenes.tower_scene() if name == "tower" else small_rain(250, seed=21)
    _run(sc, 32, 60 if name == "tower" else 120, exact=True)


@pytest.mark.parametrize("name", ["tower", "rain"])
def test_fp64_bit_exact(name):
    from pe_amd import scenes
    sc = scenes.tower_scene() if name == "tower" else small_rain(250, seed=21)
    _run(sc, 64, 60 if name == "tower" else 120, exact=True)


def test_reference_pair_set_matches_brute_force():
    from helpers import brute_force_pairs
    sc = small_rain(200, seed=4)
    r = RefWorld(sc, 32, threads=2)
    for _ in range(50):
        r.step(1)
    r.integrate()
    r.broadphase()
    mn, mx = r.aabbs()
    assert pair_set(r.pairs()) == brute_force_pairs(mn, mx, sc.kinematic)


def _shape_mix():
    from pe_amd import scenes
    return scenes.shape_mix_scene(300, seed=9, volume=(12, 8, 12), grid=16, ground_size=(16, 16))


@pytest.mark.parametrize("bits", [32, 64])
def test_shape_meshes_match_reference(bits):
    """Box and cylinder meshes, convex setMesh (volume, inertia, centring, BVH face order,
    unique edges) and the concave mesh, body by body."""
    sc = _shape_mix()
    o = OracleWorld(sc, bits, color_order=False)
    r = RefWorld(sc, bits, threads=2)
    seen = set()
    for i in range(sc.n):
        key = (int(sc.shape_type[i]), int(sc.mesh_index[i]))
        if key in seen and key[1] >= 0:
            continue
        seen.add(key)
        a, b = o.poly(i), r.poly(i)
        assert (a is None) == (b is None)
        if a is None:
            continue
        for k in ("v", "fn", "edges"):
            assert np.array_equal(a[k], b[k]), (i, k)
        assert a["faces"] == b["faces"], i
        if sc.shape_type[i] >= 3:
            assert a["volume"] == b["volume"] and np.array_equal(a["inertia"], b["inertia"]), i


@pytest.mark.parametrize("bits", [32, 64])
def test_cylinder_mesh_setmesh_order_matches_reference(bits):
    """The default cylinder mesh handed to setMesh as a convex mesh (the reference's own
    convex_mesh_shape test case): vertices, faces in the post-BVH order, normals, unique edges,
    volume and inertia equal the compiled reference's, bit for bit."""
    from pe_amd import scenes
    from pe_amd.scenes import SHAPE_CONVEX, SHAPE_CYLINDER, SceneBuilder
    b = SceneBuilder()
    b.add(SHAPE_CYLINDER, (0.5, 1.0), 1.0, (0, 0, 0), kin=1)
    P = OracleWorld(b.scene(dt=0.0, gravity=(0, 0, 0)), bits, color_order=False).poly(0)
    b = SceneBuilder()
    b.add(SHAPE_CONVEX, (), 1.0, (0, 0, 0), kin=1,
          mesh=b.add_mesh(scenes.Mesh(vertices=P["v"], faces=P["faces"], normals=P["fn"])))
    sc = b.scene(dt=0.0, gravity=(0, 0, 0))
    a, r = OracleWorld(sc, bits, color_order=False).poly(0), RefWorld(sc, bits, threads=2).poly(0)
    for k in ("v", "fn", "edges"):
        assert np.array_equal(a[k], r[k]), k
    assert a["faces"] == r["faces"]
    assert a["volume"] == r["volume"] and np.array_equal(a["inertia"], r["inertia"])


@pytest.mark.parametrize("bits", [32, 64])
def test_shape_mix_bit_exact(bits):
    """Boxes, spheres, cylinders and convex hulls over a heightfield: all 14 shape pair
    types, every stage bit-exact for 150 steps."""
    _run(_shape_mix(), bits, 150, exact=True)


@pytest.mark.parametrize("bits", [32, 64])
def test_reference_sample_scene_bit_exact(bits):
    """SampleScenes/bomb_crash_tower_smaller.json (tests/golden/): a ring of 16 boxes hit by a 70 m/s
    sphere, restitution 0.5, angular damping 0.8 -- the restitution and damping paths the synthetic
    scenes leave at zero.  150 steps, through the impact (about step 60).  The file's sleep
    thresholds are not applied: the World's defaults (0) keep every body awake in both engines."""
    import os
    sc = reference_sample_scene(os.path.join(os.path.dirname(__file__), "golden", "bomb_crash_tower_smaller.json"))
    assert sc.n == 17 and np.any(sc.restitution > 0) and np.any(sc.angular_damping > 0)
    _run(sc, bits, 150, exact=True)


def test_adapter_host_library_exports_its_entry_points():
    """The reference's World with the adapter (built beside the reference, linked to libpe_amd.so)
    loads without a GPU and exports the swap and checked-step entry points."""
    import ctypes
    import os
    from pyoracle import HERE, ref_amd_available
    if not ref_amd_available():
        pytest.skip("oracle/_ref/libperef32_amd.so not built")
    import pe_amd
    pe_amd.lib()
    lib = ctypes.CDLL(os.path.join(HERE, "_ref", "libperef32_amd.so"))
    for name in ("pref_use_amd_stages", "pref_step_checked", "pref_world_create", "pref_step"):
        assert hasattr(lib, name), name


@pytest.mark.ref
@pytest.mark.parametrize("bits", [32, 64])
def test_joints_bit_exact(bits):
    """Ball, hinge (angle limits, motor) and slider joints of the reference's constraints demo, under
    gravity with contacts: the oracle in the reference's sweep order (contacts, then joints in
    World::addConstraint order) equals the compiled reference bit for bit after every step
    (BallJointConstraint / HingeJointConstraint / SliderJointConstraint, sequential_impulse_solver.cpp:66-68)."""
    from pe_amd import scenes
    sc = scenes.joint_demo_scene()
    r = RefWorld(sc, bits, threads=2)
    o = OracleWorld(sc, bits, color_order=False)
    moved = 0.0
    p0 = r.state()[0].copy()
    for step in range(300):
        r.integrate(); o.integrate()
        r.broadphase(); o.broadphase()
        o.set_pairs(r.pairs())
        r.narrowphase(); o.narrowphase()
        r.solve(); o.solve()
        rs, os_ = r.state(), o.state()
        for a, b in zip(rs, os_):
            assert np.array_equal(a, b), f"step {step}"
        o.set_state(*rs)
    moved = float(np.abs(r.state()[0] - p0).max())
    assert moved > 1.0   # the pendulum swings, the arms turn, the slider slides


@pytest.mark.ref
def test_joint_chains_and_ignore_flags_bit_exact():
    """384 ball joints in hanging chains with neighbour ignore pairs, plus bodies flagged
    setIgnoreCollision in a falling cluster: the oracle's pairs (as a set: the reference's
    multi-threaded sort orders ties its own way) and, on the reference's pair order, its state
    after every step equal the compiled reference's."""
    from pe_amd import scenes
    chains = scenes.joint_chain_scene()
    rain = small_rain(200, seed=3)
    flags = np.zeros(rain.n, np.int32)
    flags[np.random.default_rng(2).choice(np.arange(1, rain.n), 10, replace=False)] = 1
    rain.ignore_collision = flags
    rain.ignore_pairs = [(5, 6), (7, 9), (10, 11), (3, 12)]
    for sc, steps in ((chains, 40), (rain, 40)):
        r = RefWorld(sc, 32, threads=2)
        o = OracleWorld(sc, 32, color_order=False)
        flagged = set(np.nonzero(sc.ignore_collision)[0].tolist()) if sc.ignore_collision is not None else set()
        ign = {(min(a, b), max(a, b)) for a, b in sc.ignore_pairs}
        for step in range(steps):
            r.integrate(); o.integrate()
            r.broadphase(); o.broadphase()
            rp = pair_set(r.pairs())
            assert pair_set(o.pairs()) == rp, f"{sc.name} pairs step {step}"
            assert not (rp & ign) and not any(a in flagged or b in flagged for a, b in rp)
            o.set_pairs(r.pairs())
            r.narrowphase(); o.narrowphase()
            r.solve(); o.solve()
            rs = r.state()
            for a, b in zip(rs, o.state()):
                assert np.array_equal(a, b), f"{sc.name} step {step}"
            o.set_state(*rs)


def callback_invocations(ab, ev, reporting):
    """The collision callback calls a step's events make (World::execCollisionCallbacks,
    world.cpp:153-158): A's callbacks get (A, B, pos + nor * depth, -nor, -vel), B's (B, A, pos, nor,
    vel), in FP32 as the FP32 reference computes pos + nor * depth."""
    out = []
    for (a, b), e in zip(ab, ev):
        pos, nor, vel = e[0:3].astype(np.float32), e[3:6].astype(np.float32), e[7:10].astype(np.float32)
        dep = np.float32(e[6])
        if a in reporting:
            out.append((int(a), int(b)) + tuple((pos + nor * dep).tolist()) + tuple((-nor).tolist()) + tuple((-vel).tolist()))
        if b in reporting:
            out.append((int(b), int(a)) + tuple(pos.tolist()) + tuple(nor.tolist()) + tuple(vel.tolist()))
    return sorted(out)


@pytest.mark.ref
def test_sleep_lifetime_and_callbacks_bit_exact():
    """World::updateObjectStatus with sleep thresholds (0.01 m^2/s^2, 0.01 rad^2/s^2, 1 s, as the
    reference's compound demo sets them), a body whose lifetime ends, World::removeRigidBody of another,
    the narrowphase's static / dynamic neighbour counts, and collision callbacks on three bodies: the
    oracle's status bits, counts, sleep times, pairs, callback arguments and states equal the compiled
    reference's after every stage for 420 steps.  One reference thread: the reference counts
    neighbours with unsynchronised increments from its pool threads (simple_narrow_phase.cpp:43-62),
    which lose counts when two threads meet on a body; its threaded status loop erases the expiring
    body while iterating (world.cpp:59-62), which matches the sequential loop when that body is the
    last of the array, as here."""
    sc = small_rain(120, seed=5)
    sc.sleep = (0.01, 0.01, 1.0)
    lt = np.full(sc.n, 1e30)
    lt[-1] = 2.0
    sc.lifetime = lt
    r = RefWorld(sc, 32, threads=1)
    o = OracleWorld(sc, 32, color_order=False)
    reporting = {3, 7, 20}
    for b in reporting:
        r.set_report(b)
        o.set_report(b)
    slept, woke, calls, prev = 0, 0, 0, np.zeros(sc.n, np.int32)
    for step in range(420):
        if step == 200:
            r.remove_body(40)
            o.remove_body(40)
        r.integrate(); o.integrate()
        for a, b in zip(r.status(), o.status()):
            assert np.array_equal(a, b), f"status after integrate, step {step}"
        r.broadphase(); o.broadphase()
        assert pair_set(r.pairs()) == pair_set(o.pairs()), f"pairs step {step}"
        o.set_pairs(r.pairs())
        r.narrowphase(); o.narrowphase()
        for a, b in zip(r.status(), o.status()):
            assert np.array_equal(a, b), f"neighbour counts step {step}"
        ids, data = r.callbacks()
        got = sorted(tuple(i.tolist()) + tuple(d.tolist()) for i, d in zip(ids, data))
        assert got == callback_invocations(*o.events(), reporting), f"callbacks step {step}"
        calls += len(got)
        r.solve(); o.solve()
        for a, b in zip(r.state(), o.state()):
            assert np.array_equal(a, b), f"state step {step}"
        o.set_state(*r.state())
        st = r.status()[0] & 1
        slept += int(((st == 1) & (prev == 0)).sum())
        woke += int(((st == 0) & (prev == 1)).sum())
        prev = st
    status = r.status()[0]
    assert status[-1] & 2 and status[40] & 2 and int((status & 2).sum()) == 4
    assert slept > 0 and calls > 0
    print(f"[sleep] fell asleep {slept} times, woke {woke} times, {calls} callback calls", flush=True)


def _ray_lockstep(sc, steps, bits):
    r, o = RefWorld(sc, bits, threads=1), OracleWorld(sc, bits, color_order=False)
    for _ in range(steps):
        r.integrate(); o.integrate()
        r.broadphase()
        o.set_pairs(r.pairs())
        r.narrowphase(); o.narrowphase()
        r.solve(); o.solve()
        o.set_state(*r.state())
    return r, o


def _same_ray(a, b):
    return (a[0] == b[0] and a[1] == b[1] and a[2] == b[2] and np.array_equal(a[3], b[3])
            and np.array_equal(a[4], b[4]))


@pytest.mark.ref
@pytest.mark.parametrize("bits", [32, 64])
def test_raycast_bit_exact(bits):
    """DefaultRaycastSolver::performRaycast (default_raycast_solver.cpp:11-86) after 30 steps: the
    oracle's hit flag, body, distance, point and normal equal the compiled reference's for rays in
    every direction (some with ignore sets, some starting inside bodies) over a box rain, over the
    mixed-shape scene (whose many hulls end most queries early: a mesh body whose AABB the ray misses
    makes the query report no hit, the result keeping the earlier bodies' best), and over one hull on
    the heightfield with rays aimed down at the hull (mesh hits)."""
    from pe_amd import scenes
    rng = np.random.default_rng(1)
    mix = scenes.shape_mix_scene(160, seed=23, volume=(10, 6, 10), grid=12, ground_size=(14, 14))
    conv = np.nonzero(mix.shape_type == scenes.SHAPE_CONVEX)[0]
    keep = sorted([i for i in range(mix.n) if mix.shape_type[i] != scenes.SHAPE_CONVEX] + [int(conv[0])])
    one_hull = mix.subset(keep)
    counts = {}
    for name, sc in (("rain", small_rain(150, seed=2)), ("mix", mix), ("hull", one_hull)):
        r, o = _ray_lockstep(sc, 30, bits)
        pos = r.state()[0]
        lo, hi = pos.min(0) - 1, pos.max(0) + 1
        hull = keep.index(int(conv[0]))
        hits = partial = mesh_hits = 0
        for k in range(300):
            if name == "hull":
                tgt = pos[hull] + rng.uniform(-0.4, 0.4, 3) if k % 2 == 0 else rng.uniform(lo, hi)
                start, d = tgt + np.array([0.0, 8.0, 0.0]), np.array([0.0, -1.0, 0.0]) + rng.normal(size=3) * 0.01
                length, ignore = 20.0, []
            else:
                start = rng.uniform(lo, hi)
                d = rng.normal(size=3) if k % 3 else np.array([rng.normal() * 0.2, -1.0, rng.normal() * 0.2])
                if k % 7 == 0:
                    d = np.array([0.0, -1.0, 0.0])
                length = float(rng.uniform(0.5, 40.0))
                ignore = rng.choice(sc.n, 3, replace=False).tolist() if k % 5 == 0 else []
            d = d / np.linalg.norm(d)
            a, b = r.raycast(start, d, length, ignore), o.raycast(start, d, length, ignore)
            assert _same_ray(a, b), f"{name} ray {k}: {a} vs {b}"
            hits += a[0]
            partial += (not a[0]) and a[1] >= 0
            mesh_hits += a[0] and sc.shape_type[a[1]] >= scenes.SHAPE_CONVEX
        counts[name] = (hits, partial, mesh_hits)
    print(f"[raycast FP{bits}] (hits, ended early with a partial result, mesh hits): {counts}", flush=True)
    assert counts["rain"][0] > 50 and counts["mix"][1] > 10 and counts["hull"][2] > 5


def test_raycast_closed_forms():
    """What the geometry fixes, on one unit box and one sphere: a ray down onto the box's top face
    hits at (start height - 0.5) with normal +y (the point displaced along the EPS-raised direction); a ray through a sphere's centre hits at
    (distance to centre - r) with the normal pointing back at the start; a ray that starts past a
    sphere hits it at a negative distance (RaycastSphere keeps it, raycast_sphere.cpp:21-27); a ray
    pointing away from the box misses."""
    from pe_amd.scenes import SHAPE_BOX, SHAPE_SPHERE, SceneBuilder
    b = SceneBuilder()
    b.add(SHAPE_BOX, (1.0, 1.0, 1.0), 1.0, (0.0, 0.0, 0.0), kin=1)
    b.add(SHAPE_SPHERE, (0.5,), 1.0, (5.0, 0.0, 0.0), kin=1)
    sc = b.scene(name="ray_kat")
    o = OracleWorld(sc, 64)
    o.integrate()
    hit, body, d, p, n = o.raycast((0.25, 3.0, 0.1), (0.0, -1.0, 0.0), 10.0)
    # the hit point follows the direction with its zero components raised to PE_EPS = 1e-5
    # (raycast_box.cpp:35-38, 49): x and z move by 2.5e-5 over the 2.5 m
    assert hit and body == 0 and d == 2.5 and np.allclose(p, (0.25 + 2.5e-5, 0.5, 0.1 + 2.5e-5), rtol=0, atol=1e-12)
    assert np.array_equal(n, (0, 1, 0))
    hit, body, d, p, n = o.raycast((5.0, 4.0, 0.0), (0.0, -1.0, 0.0), 10.0, ignore=[0])
    assert hit and body == 1 and d == 3.5 and np.allclose(n, (0, 1, 0))
    hit, body, d, p, n = o.raycast((5.0, 4.0, 0.0), (0.0, 1.0, 0.0), 10.0, ignore=[0])
    assert hit and body == 1 and d == -4.5
    hit, body, d, p, n = o.raycast((0.0, 3.0, 0.0), (0.0, 1.0, 0.0), 10.0, ignore=[1])
    assert not hit and body == -1


@pytest.mark.ref
@pytest.mark.parametrize("bits", [32, 64])
def test_compound_bit_exact(bits):
    """CompoundShape (compound_shape.cpp: addShape, init, getAABB) and
    CompoundCompoundCollisionAlgorithm (compound_compound_collision_algorithm.cpp:8-48): the reference's
    compound demo (boxes with cylinder arms stacked, with a box, a sphere and a cylinder) and a pile of
    compounds with box, sphere, cylinder and convex-hull children over a heightfield.  Mass properties
    (init's centroid shift and inertia), AABBs, pair sets, contact slots (in the result's object order:
    a compound B trades places with A) and states equal the compiled reference's after every stage,
    in FP32 and FP64; raycasts down the demo's stack hit the compounds' children alike."""
    from pe_amd import scenes
    for sc, steps in ((scenes.compound_demo_scene(), 200), (scenes.compound_pile_scene(), 200)):
        r = RefWorld(sc, bits, threads=1)
        o = OracleWorld(sc, bits, color_order=False)
        for a, b in zip(r.mass_props(), o.mass_props()):
            assert np.array_equal(a, b), f"{sc.name} mass properties"
        points = compound_points = 0
        comp = np.array([i in sc.compounds for i in range(sc.n)])

        ray_hits = 0
        for step in range(steps):
            r.integrate(); o.integrate()
            for a, b in zip(r.aabbs(), o.aabbs()):
                assert np.array_equal(a, b), f"{sc.name} AABBs step {step}"
            if step == 0 and sc.name.startswith("compound_demo"):
                # the stack still stands: a vertical ray through it crosses every compound's AABB.  A
                # compound's children write the same outputs in turn and the last child's distance
                # decides (default_raycast_solver.cpp:50-78), so rays at the last child's line (z = -1)
                # hit the top compound and rays elsewhere fall through to the bodies below
                for x, z in [(x, z) for x in np.linspace(-0.3, 0.3, 7) for z in (-1.05, -1.0, -0.9, 0.05)]:
                    ray = ((x, 80.0, z), (0.0, -1.0, 0.0), 100.0)
                    a, b = r.raycast(*ray), o.raycast(*ray)
                    assert _same_ray(a, b), f"compound ray at x={x}: {a} vs {b}"
                    ray_hits += a[0] and comp[a[1]]
            r.broadphase(); o.broadphase()
            assert pair_set(r.pairs()) == pair_set(o.pairs()), f"{sc.name} pairs step {step}"
            o.set_pairs(r.pairs())
            r.narrowphase(); o.narrowphase()
            rn, rs, rp = r.contacts()
            on, os_, op = o.contacts()
            assert np.array_equal(rn, on) and np.array_equal(rs, os_) and np.array_equal(rp, op), f"{sc.name} contacts step {step}"
            points += int(rn.sum())
            pr = r.pairs()
            if len(pr):
                compound_points += int(rn[comp[pr[:, 0]] | comp[pr[:, 1]]].sum())
            r.solve(); o.solve()
            for a, b in zip(r.state(), o.state()):
                assert np.array_equal(a, b), f"{sc.name} state step {step}"
            o.set_state(*r.state())
        assert compound_points > 100
        assert ray_hits > 5 or not sc.name.startswith("compound_demo")
        print(f"[compound FP{bits}] {sc.name}: {points} contact points, {compound_points} on compounds, "
              f"{ray_hits} ray hits on compounds", flush=True)


def test_batches_refuse_joints():
    from pe_amd import scenes
    from pe_amd.batch import Batch
    with pytest.raises(ValueError, match="joints"):
        Batch([scenes.joint_demo_scene()])


def test_joint_constraints_hold_in_the_oracle():
    """What the mathematics fixes, step after step under gravity: the ball joints' anchors coincide
    (Baumgarte-stabilised: drift of millimetres), the slider's body stays on its rail's axis, and the
    limited hinge's angle (measured as HingeJointConstraint measures it, hinge_joint_constraint.cpp:
    56-60) reaches its limit and stays within it up to the stabilisation slack; without the joints the
    same bodies fall apart."""
    from pe_amd import scenes
    sc = scenes.joint_demo_scene()
    o = OracleWorld(sc, 64, color_order=True)
    h = sc.joints[3]
    worst_gap, worst_off, angles = 0.0, 0.0, []
    for step in range(300):
        o.step(1)
        pos, rot, _, _ = o.state()
        for j in sc.joints[:3]:
            pa = rot[j.body_a] @ np.array(j.anchor_a) + pos[j.body_a]
            pb = rot[j.body_b] @ np.array(j.anchor_b) + pos[j.body_b]
            worst_gap = max(worst_gap, float(np.linalg.norm(pa - pb)))
        rail, sl = sc.joints[5].body_a, sc.joints[5].body_b
        axis = rot[rail] @ np.array([0.0, 1.0, 0.0])
        d = pos[sl] - pos[rail]
        worst_off = max(worst_off, float(np.linalg.norm(d - (d @ axis) * axis)))
        ta = rot[h.body_a] @ ortho_units(h.axis_a)[0]
        tb = rot[h.body_b] @ ortho_units(h.axis_b)[0]
        wa = rot[h.body_a] @ np.array(h.axis_a)
        angles.append(float(np.arctan2(np.cross(ta, tb) @ wa, ta @ tb)))
    print(f"[joints] anchor gap {worst_gap:.2e} m, slider off-axis {worst_off:.2e} m, hinge angle "
          f"[{min(angles):.3f}, {max(angles):.3f}] rad", flush=True)
    assert worst_gap < 2e-2 and worst_off < 2e-2
    lim = np.pi / 3
    assert max(abs(a) for a in angles) > 0.9 * lim           # the arm swings up to its limit
    assert all(-lim - 0.1 <= a <= lim + 0.1 for a in angles)  # and no further than the slack
    free = scenes.joint_demo_scene()
    free.joints = []
    f = OracleWorld(free, 64)
    f.step(300)
    fp = f.state()[0]
    pa = f.state()[1][1] @ np.zeros(3) + fp[1]
    pb = f.state()[1][2] @ np.array([0, -3.3, 0]) + fp[2]
    assert np.linalg.norm(pa - pb) > 0.5
