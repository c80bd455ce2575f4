"""Pins the oracle (C++ restatement) to the reference engine compiled from its own sources
(oracle/_ref).  FP32 and FP64: every stage bit-exact.
The reference sorts with std::sort, whose order among equal keys is unspecified; where its
pair order differs from the restatement's (ties), the reference's order is fed to the oracle
so both continue from the same pair list.  Skipped where oracle/_ref is not built."""
import numpy as np
import pytest

from helpers import flat_state, pair_set, reference_sample_scene, small_rain
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
