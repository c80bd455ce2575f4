"""The oracle against golden fixtures recorded from the compiled reference
(tests/golden/make_golden.py).  Runs without the reference: states and contact slots must
match bit for bit in FP32 and FP64.  The reference's per-step pair list is fed to the
oracle so both sweep in the same order (std::sort leaves equal keys unordered)."""
import os

import numpy as np
import pytest

from pe_amd.scenes import Mesh, Scene
from pyoracle import OracleWorld

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(HERE, name))
    p = z["params"]
    meshes, mesh_index = None, None
    if "mesh_nv" in z:
        meshes, vo, fo, io = [], 0, 0, 0
        flen = z["mesh_face_len"]
        for nv, nf in zip(z["mesh_nv"], z["mesh_nf"]):
            faces = []
            for k in range(fo, fo + nf):
                faces.append(z["mesh_face_index"][io:io + flen[k]].tolist())
                io += flen[k]
            meshes.append(Mesh(z["mesh_vertices"][vo:vo + nv], faces, z["mesh_normals"][fo:fo + nf]))
            vo += nv
            fo += nf
        mesh_index = z["mesh_index"]
    sc = Scene(z["shape_type"], z["shape_param"], z["mass"], z["position"], z["rotation"], z["linear_velocity"],
               z["angular_velocity"], z["friction"], z["restitution"], z["kinematic"], dt=float(p[0]),
               gravity=tuple(p[1:4]), iterations=int(p[4]), meshes=meshes, mesh_index=mesh_index).contiguous()
    return z, sc


@pytest.mark.parametrize("name,bits", [("tower_fp32.npz", 32), ("tower_fp64.npz", 64), ("rain120_fp32.npz", 32),
                                       ("shape_mix200_fp32.npz", 32), ("shape_mix200_fp64.npz", 64)])
def test_oracle_reproduces_reference_fixture(name, bits):
    z, sc = load(name)
    o = OracleWorld(sc, bits, color_order=False)
    counts = z["pair_counts"]
    offs = np.concatenate([[0], np.cumsum(counts)])
    cstep = int(z["contacts_step"][0])
    for s in range(1, len(counts) + 1):
        o.integrate()
        o.broadphase()
        ref_pairs = z["pairs"][offs[s - 1]:offs[s]]
        assert set(map(tuple, np.sort(o.pairs(), 1).tolist())) == set(map(tuple, np.sort(ref_pairs, 1).tolist()))
        o.set_pairs(ref_pairs)
        o.narrowphase()
        if s == cstep:
            npts, swap, pts = o.contacts()
            assert np.array_equal(npts, z["contacts_npts"]) and np.array_equal(swap, z["contacts_swap"])
            ref = z["contacts_pts"]
            vm = ref[..., 0] < 1e29
            got = pts.astype(ref.dtype)
            assert np.array_equal(got[vm], ref[vm]) and np.array_equal(got[..., 12], ref[..., 12])
        o.solve()
        key = f"state_{s}"
        if key in z:
            pos, rot, v, w = o.state()
            st = np.concatenate([pos, rot.reshape(-1, 9), v, w], axis=1)
            assert np.array_equal(st, z[key]), key
