"""The library's host-side shape preparation (pe_mesh_prepare: ConvexMeshShape::setMesh and
ConcaveMeshShape::setMesh in FP32) against the oracle's restatement, which is pinned to the
compiled reference (test_oracle_vs_ref.py): bit-exact vertices, face order, normals, unique
edges, volume and inertia.  Runs on the CPU."""
import numpy as np
import pytest

from pe_amd import scenes
from pe_amd.mesh import prepare_mesh
from pyoracle import OracleWorld


def _one_body_scene(mesh, concave):
    b = scenes.SceneBuilder()
    m = b.add_mesh(mesh)
    b.add(scenes.SHAPE_CONCAVE if concave else scenes.SHAPE_CONVEX, (), 1.0, (0, 0, 0), kin=1 if concave else 0, mesh=m)
    return b.scene()


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_convex_hull_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    mesh = scenes.gaussian_hull(rng, 32, 0.5)
    got = prepare_mesh(mesh, concave=False)
    ref = OracleWorld(_one_body_scene(mesh, False), 32).poly(0)
    assert np.array_equal(got["v"], ref["v"].astype(np.float32))
    assert got["faces"] == ref["faces"]
    assert np.array_equal(got["fn"], ref["fn"].astype(np.float32))
    assert np.array_equal(got["edges"], ref["edges"].astype(np.float32))
    assert got["volume"] == np.float32(ref["volume"])
    assert np.array_equal(got["inertia"], ref["inertia"].astype(np.float32))


def test_heightfield_matches_oracle():
    rng = np.random.default_rng(7)
    mesh = scenes.heightfield_mesh(rng, 24, 24, (20.0, 20.0), 1.0)
    got = prepare_mesh(mesh, concave=True)
    ref = OracleWorld(_one_body_scene(mesh, True), 32).poly(0)
    assert np.array_equal(got["v"], ref["v"].astype(np.float32))
    assert got["faces"] == ref["faces"]
    assert np.array_equal(got["fn"], ref["fn"].astype(np.float32))
    assert len(got["edges"]) == 0 == len(ref["edges"])


def test_closed_form_box_hull():
    """A box given as a triangle mesh: volume abc, centre of mass at the box centre (the
    relocation), unit-mass inertia (b^2 + c^2) / 12 etc. with zero products of inertia."""
    a, b_, c = 0.6, 0.4, 0.2
    off = np.array([1.0, -2.0, 0.5])
    pts = np.array([[x, y, z] for x in (-a / 2, a / 2) for y in (-b_ / 2, b_ / 2) for z in (-c / 2, c / 2)]) + off
    got = prepare_mesh(scenes.hull_mesh(pts), concave=False)
    assert got["volume"] == pytest.approx(a * b_ * c, rel=1e-5)
    assert np.allclose(got["relocation"], off, atol=1e-5)
    I = got["inertia"]
    assert np.allclose(np.diag(I), [(b_ ** 2 + c ** 2) / 12, (a ** 2 + c ** 2) / 12, (a ** 2 + b_ ** 2) / 12], rtol=1e-4)
    assert np.allclose(I - np.diag(np.diag(I)), 0, atol=1e-6)
    assert np.allclose(got["v"].mean(axis=0), 0, atol=1e-5)
