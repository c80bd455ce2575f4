"""Scene files in the reference's JSON format (pe_amd.scene_io): the reference's own sample scene
loads to the same bodies the reference's loader builds, files round-trip, quaternion rotations
and mesh bodies load as the reference loads them, and features the device step does not model
are refused."""
import json
import math
import os

import numpy as np
import pytest

from helpers import reference_sample_scene

HERE = os.path.dirname(__file__)
SAMPLE = os.path.join(HERE, "golden", "bomb_crash_tower_smaller.json")
CYL_OBJ = os.path.join(HERE, "golden", "cylinder_default_mesh_setmesh.obj")
FIELDS = ("shape_type", "shape_param", "mass", "position", "rotation", "linear_velocity", "angular_velocity",
          "friction", "restitution", "kinematic", "linear_damping", "angular_damping", "mesh_index")


def _same(a, b):
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert tuple(a.gravity) == tuple(b.gravity)


def test_sample_scene_loads_as_the_reference_loader():
    from pe_amd.scene_io import load_scene
    sc = load_scene(SAMPLE)
    _same(sc, reference_sample_scene(SAMPLE))
    assert sc.sleep == (0.04, 0.08, 1.0)
    assert sc.n == len(json.load(open(SAMPLE))["RigidBodies"])


def test_round_trip(tmp_path):
    from pe_amd import scenes
    from pe_amd.scene_io import load_scene, save_scene
    sc = scenes.shape_mix_scene(40, seed=3, volume=(6, 4, 6), grid=4, ground_size=(8, 8))
    out = str(tmp_path / "mix.json")
    save_scene(sc, out)
    back = load_scene(out)
    _same(sc, back)
    for m0, m1 in zip(sc.meshes, back.meshes):
        assert np.array_equal(np.asarray(m0.vertices), np.asarray(m1.vertices))
        assert [list(f) for f in m0.faces] == [list(f) for f in m1.faces]
        assert np.allclose(m0.normals, m1.normals, rtol=0, atol=1e-12)


def test_quaternion_rotation_and_defaults(tmp_path):
    from pe_amd.scene_io import load_scene, quaternion_basis
    ang = 0.7
    q = [math.cos(ang / 2), 0.0, math.sin(ang / 2), 0.0]      # (w, x, y, z): ang about +y
    data = {"RigidBodies": [{"type": "box", "position": [0, 1, 0], "rotation": q}]}
    p = tmp_path / "q.json"
    p.write_text(json.dumps(data))
    sc = load_scene(str(p))
    c, s = math.cos(ang), math.sin(ang)
    assert np.allclose(sc.rotation[0], [[c, 0, s], [0, 1, 0], [-s, 0, c]], atol=1e-15)
    assert np.array_equal(sc.rotation[0], quaternion_basis(*q))
    # a RigidBody's defaults (rigidbody.cpp:91-100): mass 1, friction 0.5, restitution 0.5
    assert sc.mass[0] == 1.0 and sc.friction[0] == 0.5 and sc.restitution[0] == 0.5
    assert sc.gravity == (0.0, -9.8, 0.0)


def test_mesh_body_scaled_by_size_x(tmp_path):
    """Mesh::loadFromObj multiplies every coordinate by size.x (mesh.cpp:51)."""
    from pe_amd.scene_io import load_obj, load_scene
    data = {"RigidBodies": [{"type": "convex", "mesh": CYL_OBJ, "scale": [2.0, 5.0, 7.0], "mass": 3.0}]}
    p = tmp_path / "m.json"
    p.write_text(json.dumps(data))
    sc = load_scene(str(p))
    ref = load_obj(CYL_OBJ, 1.0)
    assert sc.mesh_index[0] == 0 and len(sc.meshes) == 1
    assert np.array_equal(np.asarray(sc.meshes[0].vertices), 2.0 * np.asarray(ref.vertices))
    assert len(sc.meshes[0].faces) == 42


def test_ignore_collision_lifetime_and_sleep_round_trip(tmp_path):
    """ignore_collision, a finite lifetime (negative and PE_REAL_MAX of either precision mean
    unlimited, simulator.cpp:126, 358-361) and the sleep thresholds load into the Scene and save back."""
    from pe_amd.scene_io import load_scene, save_scene
    rbs = [{"type": "box", "scale": [1, 1, 1], "ignore_collision": True},
           {"type": "box", "scale": [1, 1, 1], "lifetime": 5.0},
           {"type": "box", "scale": [1, 1, 1], "lifetime": -1.0},
           {"type": "box", "scale": [1, 1, 1], "lifetime": 1e30},
           {"type": "box", "scale": [1, 1, 1], "lifetime": 1e100}]
    cfg = {"sleep_linear_velocity2_threshold": 0.01, "sleep_angular_velocity2_threshold": 0.02,
           "sleep_time_threshold": 1.0}
    p = tmp_path / "s.json"
    p.write_text(json.dumps({"Configuration": cfg, "RigidBodies": rbs}))
    sc = load_scene(str(p))
    assert sc.ignore_collision.tolist() == [1, 0, 0, 0, 0]
    assert sc.lifetime.tolist() == [1e100, 5.0, 1e100, 1e100, 1e100]
    assert sc.sleep == (0.01, 0.02, 1.0)
    out = str(tmp_path / "back.json")
    save_scene(sc, out)
    back = load_scene(out)
    assert back.ignore_collision.tolist() == sc.ignore_collision.tolist()
    assert back.lifetime.tolist() == sc.lifetime.tolist() and back.sleep == sc.sleep


@pytest.mark.parametrize("extra", [{"fracturable": True}, {"type": "compound"}])
def test_unmodelled_features_are_refused(tmp_path, extra):
    from pe_amd.scene_io import load_scene
    rb = {"type": "box", "scale": [1, 1, 1]}
    rb.update(extra)
    p = tmp_path / "bad.json"
    p.write_text(json.dumps({"RigidBodies": [rb]}))
    with pytest.raises(ValueError):
        load_scene(str(p))
