"""Scene files in the reference engine's JSON format, so a scene saved by its simulator loads here
and a scene built here opens in the reference (src/intf/simulator.cpp: saveScene 79-250, loadScene
252-591).

    sc = load_scene("SampleScenes/bomb_crash_tower_smaller.json")
    save_scene(sc, "out.json", mesh_paths={0: "hull.obj"})

What the file holds and how it maps onto a Scene:

* ``Configuration.gravity`` -> ``Scene.gravity``.  The sleep thresholds -> ``Scene.sleep``
  (linear velocity^2, angular velocity^2, time), which a World applies (World::setSleep*Threshold;
  batched and served worlds take none).
* ``RigidBodies[]``: ``type`` box (``scale`` = full size), sphere (``scale[0]`` = radius),
  cylinder (``scale`` = radius, height), convex / concave (``mesh`` = a Wavefront .obj path,
  ``scale``); ``position``; ``rotation`` as 9 row-major basis entries or 4 quaternion entries
  (w, x, y, z); ``mass``, ``friction``, ``restitution`` (defaults 1, 0.5, 0.5 as a RigidBody's,
  rigidbody.cpp:91-100); ``kinematic``; for dynamic bodies ``linear_velocity``,
  ``angular_velocity``, ``linear_damping``, ``angular_damping``; ``ignore_collision`` ->
  ``Scene.ignore_collision``; ``lifetime`` -> ``Scene.lifetime`` (negative, or PE_REAL_MAX as saveScene
  writes it, means unlimited, as loadScene takes it).
* What the step does not model is refused with ValueError rather than dropped: compound shapes,
  fracturable bodies, damage sources.  ``name`` and ``tag`` are kept on the Scene (``Scene.names``,
  ``Scene.tags``) for saving.

Mesh files are read as the reference's Mesh::loadFromObj reads them
(SimpleViewer/include/common/mesh.cpp:36-67): ``v`` lines scaled by ``scale[0]`` on every axis
(the reference multiplies all three coordinates by size.x), ``f`` lines as 1-based vertex indices
(``i/j/k`` forms keep i), face normals normalize((p1 - p0) x (p2 - p0)).  The scene's meshes are
prepared as ConvexMeshShape::setMesh / ConcaveMeshShape::setMesh prepare them when a World or
Batch is created.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .scenes import (SHAPE_BOX, SHAPE_CONCAVE, SHAPE_CONVEX, SHAPE_CYLINDER, SHAPE_SPHERE, Mesh, Scene,
                     SceneBuilder)

_TYPES = {"box": SHAPE_BOX, "sphere": SHAPE_SPHERE, "cylinder": SHAPE_CYLINDER, "convex": SHAPE_CONVEX,
          "concave": SHAPE_CONCAVE}
_NAMES = {v: k for k, v in _TYPES.items()}


def quaternion_basis(w: float, x: float, y: float, z: float) -> np.ndarray:
    """Rotation basis of the quaternion (w, x, y, z), in the reference's evaluation order
    (Eigen's QuaternionBase::toRotationMatrix: doubled components, then their products)."""
    tx, ty, tz = 2.0 * x, 2.0 * y, 2.0 * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    return np.array([[1.0 - (tyy + tzz), txy - twz, txz + twy],
                     [txy + twz, 1.0 - (txx + tzz), tyz - twx],
                     [txz - twy, tyz + twx, 1.0 - (txx + tyy)]])


def _face_normals(mesh: Mesh) -> np.ndarray:
    """normalize((p1 - p0) x (p2 - p0)) per face, as Mesh::perFaceNormal computes it."""
    v = np.asarray(mesh.vertices, np.float64)
    out = []
    for fc in mesh.faces:
        n = np.cross(v[fc[1]] - v[fc[0]], v[fc[2]] - v[fc[0]])
        out.append(n / math.sqrt(float(n @ n)))
    return np.asarray(out, float).reshape(-1, 3)


def load_obj(path: str, scale: float = 1.0) -> Mesh:
    """A Wavefront .obj read as Mesh::loadFromObj reads it (mesh.cpp:36-67), coordinates times scale."""
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                verts.append([float(tok[1]) * scale, float(tok[2]) * scale, float(tok[3]) * scale])
            elif tok[0] == "f":
                faces.append([int(t.split("/")[0]) - 1 for t in tok[1:]])
    v = np.asarray(verts, float).reshape(-1, 3)
    for fc in faces:
        if len(fc) < 3:
            raise ValueError(f"{path}: a face with fewer than 3 vertices")
    m = Mesh(v, faces, np.zeros((len(faces), 3)))
    m.normals = _face_normals(m)
    return m


def save_obj(mesh: Mesh, path: str) -> None:
    """Vertices and faces of mesh as a Wavefront .obj (1-based indices)."""
    with open(path, "w") as f:
        for p in np.asarray(mesh.vertices, float):
            f.write("v " + " ".join(repr(float(x)) for x in p) + "\n")
        for fc in mesh.faces:
            f.write("f " + " ".join(str(int(i) + 1) for i in fc) + "\n")


def load_scene(path: str, dt: float = 1.0 / 60.0, iterations: int = 10) -> Scene:
    """A Scene from a reference scene file (Simulator::loadScene, simulator.cpp:252-591).  Mesh
    paths are taken as written, then relative to the scene file's folder."""
    with open(path) as f:
        data = json.load(f)
    if data.get("DamageSources"):
        raise ValueError("damage sources (fracture) are not part of the device step")
    cfg = data.get("Configuration", {})
    b = SceneBuilder()
    lin, ang, names, tags, ign, lives = [], [], [], [], [], []
    mesh_ids = {}
    for k, rb in enumerate(data.get("RigidBodies", [])):
        if rb.get("fracturable", False):
            raise ValueError(f"body {k}: fracturable bodies are not part of the device step")
        ign.append(1 if rb.get("ignore_collision", False) else 0)
        life = float(rb.get("lifetime", 1e100))
        # saveScene writes PE_REAL_MAX for an unlimited lifetime: 1e30 in the FP32 build, 1e100 in
        # the FP64 one (phys_general.h:36-42); loadScene takes a negative value as unlimited too
        lives.append(life if 0.0 <= life < 1e30 else 1e100)
        typ = rb.get("type")
        if typ not in _TYPES:
            raise ValueError(f"body {k}: shape type {typ!r} is not part of the device step")
        kin = 1 if rb.get("kinematic", False) else 0
        rot = rb.get("rotation")
        if rot is None:
            basis = np.eye(3)
        elif len(rot) >= 9:
            basis = np.asarray(rot[:9], float).reshape(3, 3)
        elif len(rot) >= 4:
            basis = quaternion_basis(*[float(x) for x in rot[:4]])
        else:
            raise ValueError(f"body {k}: invalid rotation")
        v = (0.0, 0.0, 0.0) if kin else tuple(rb.get("linear_velocity", (0.0, 0.0, 0.0)))
        w = (0.0, 0.0, 0.0) if kin else tuple(rb.get("angular_velocity", (0.0, 0.0, 0.0)))
        scale = rb.get("scale")
        mesh = -1
        shape = _TYPES[typ]
        if shape == SHAPE_BOX:
            param = tuple(scale[:3]) if scale and len(scale) >= 3 else (1.0, 1.0, 1.0)
        elif shape == SHAPE_SPHERE:
            param = (scale[0],) if scale else (1.0,)   # loadScene's sphere default: SphereShape(1)
        elif shape == SHAPE_CYLINDER:
            param = tuple(scale[:2]) if scale and len(scale) >= 2 else (0.5, 1.0)
        else:
            if "mesh" not in rb:
                raise ValueError(f"body {k}: mesh path not specified")
            size = float(scale[0]) if scale and len(scale) >= 3 else 1.0
            obj = rb["mesh"]
            if not os.path.exists(obj):
                obj = os.path.join(os.path.dirname(os.path.abspath(path)), rb["mesh"])
            key = (os.path.abspath(obj), size)
            if key not in mesh_ids:
                mesh_ids[key] = b.add_mesh(load_obj(obj, size))
            mesh = mesh_ids[key]
            param = ()
        b.add(shape, param, float(rb.get("mass", 1.0)), tuple(rb.get("position", (0.0, 0.0, 0.0))), rot=basis,
              v=v, w=w, mu=float(rb.get("friction", 0.5)), e=float(rb.get("restitution", 0.5)), kin=kin, mesh=mesh)
        lin.append(0.0 if kin else float(rb.get("linear_damping", 0.0)))
        ang.append(0.0 if kin else float(rb.get("angular_damping", 0.0)))
        names.append(rb.get("name", ""))
        tags.append(rb.get("tag", ""))
    sc = b.scene(gravity=tuple(float(x) for x in cfg.get("gravity", (0.0, -9.8, 0.0))), dt=dt,
                 iterations=iterations, name=os.path.splitext(os.path.basename(path))[0])
    sc.linear_damping = np.asarray(lin, float)
    sc.angular_damping = np.asarray(ang, float)
    sc = sc.contiguous()
    sc.names, sc.tags = names, tags
    sleep = (float(cfg.get("sleep_linear_velocity2_threshold", 0.0)),
             float(cfg.get("sleep_angular_velocity2_threshold", 0.0)),
             float(cfg.get("sleep_time_threshold", 0.0)))
    sc.sleep = sleep if (sleep[0] != 0.0 or sleep[1] != 0.0) else None
    if any(ign):
        sc.ignore_collision = np.asarray(ign, np.int32)
    if any(x < 1e30 for x in lives):
        sc.lifetime = np.asarray(lives, float)
    return sc


def save_scene(sc: Scene, path: str, mesh_paths=None, extra_config=None, extra=None) -> None:
    """Writes sc in the reference's scene format (Simulator::saveScene, simulator.cpp:79-250).
    mesh_paths maps a mesh index to the .obj path written for it; meshes without one are saved
    beside the scene file as <scene>_mesh<k>.obj (at scale 1).  extra_config adds keys to
    ``Configuration`` and extra adds top-level keys (the reference's loader ignores both)."""
    sc = sc.contiguous()
    mesh_paths = dict(mesh_paths or {})
    base = os.path.splitext(path)[0]
    for k, m in enumerate(sc.meshes):
        if k not in mesh_paths:
            mesh_paths[k] = f"{base}_mesh{k}.obj"
            save_obj(m, mesh_paths[k])
    sleep = getattr(sc, "sleep", None) or (0.0, 0.0, 0.0)
    ign = getattr(sc, "ignore_collision", None)
    life = getattr(sc, "lifetime", None)
    names = getattr(sc, "names", None) or [""] * sc.n
    tags = getattr(sc, "tags", None) or [""] * sc.n
    rbs = []
    for i in range(sc.n):
        t = int(sc.shape_type[i])
        kin = bool(sc.kinematic[i])
        rb = {"name": names[i], "tag": tags[i], "ignore_collision": bool(ign is not None and ign[i]),
              "kinematic": kin}
        if not kin:
            rb["linear_damping"] = float(sc.linear_damping[i])
            rb["angular_damping"] = float(sc.angular_damping[i])
            rb["linear_velocity"] = [float(x) for x in sc.linear_velocity[i]]
            rb["angular_velocity"] = [float(x) for x in sc.angular_velocity[i]]
        rb["lifetime"] = float(life[i]) if life is not None and life[i] < 1e30 else 1e100
        rb["fracturable"] = False
        rb["position"] = [float(x) for x in sc.position[i]]
        rb["rotation"] = [float(x) for x in np.asarray(sc.rotation[i]).ravel()]
        rb["mass"] = float(sc.mass[i])
        rb["friction"] = float(sc.friction[i])
        rb["restitution"] = float(sc.restitution[i])
        rb["type"] = _NAMES[t]
        p = sc.shape_param[i]
        if t == SHAPE_BOX:
            rb["scale"] = [float(p[0]), float(p[1]), float(p[2])]
        elif t == SHAPE_SPHERE:
            rb["scale"] = [float(p[0])]
        elif t == SHAPE_CYLINDER:
            rb["scale"] = [float(p[0]), float(p[1])]
        else:
            rb["mesh"] = mesh_paths[int(sc.mesh_index[i])]
            rb["scale"] = [1.0, 1.0, 1.0]
        rbs.append(rb)
    data = {"Configuration": {"gravity": [float(x) for x in sc.gravity],
                              "sleep_linear_velocity2_threshold": sleep[0],
                              "sleep_angular_velocity2_threshold": sleep[1],
                              "sleep_time_threshold": sleep[2]},
            "DamageSources": [], "RigidBodies": rbs}
    data["Configuration"].update(extra_config or {})
    data.update(extra or {})
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def save_checkpoint(world, path: str, mesh_paths=None) -> None:
    """A scene file from which load_checkpoint resumes `world` exactly where it is: the next step of
    the loaded world is bit-identical to the next step of this one.

    It is the reference's scene format (so the reference's Simulator::loadScene opens it too) with
    the bodies' current poses and velocities -- FP32 values written as the exact decimal of their
    FP64 widening -- and the meshes as the scene gave them (every load prepares them alike: setMesh
    is deterministic, so the device sees the same faces in the same order).  What a step reads from
    the steps before it beyond the body state goes under keys the reference ignores:
    ``Configuration.dt``, ``Configuration.iterations``, ``Configuration.sap_axis`` (the axis the next
    sweep sorts on, broad_phase_sweep_and_prune.cpp:37-40) and ``ContactCache`` (the pair count of
    the last narrowphase and its slots' distances, FP32 little-endian in base64: the reference keeps
    them between steps in its reused ContactResult objects, contact_result.cpp:158-164).  A
    kinematic body's velocity is not written (loadScene never sets one, simulator.cpp:350-356)."""
    import base64
    import copy
    sc = copy.copy(world.scene.contiguous())
    for k, m in enumerate(sc.meshes):
        if not np.array_equal(np.asarray(m.normals, np.float64), _face_normals(m)):
            raise ValueError(f"mesh {k}: its face normals are not the ones a scene file's loader computes from its "
                             f"vertices (Mesh::perFaceNormal), so a loaded checkpoint would not resume it exactly")
    pos, rot, v, w = world.state()
    sc.position = np.asarray(pos, np.float64)
    sc.rotation = np.asarray(rot, np.float64)
    kin = sc.kinematic.astype(bool)
    sc.linear_velocity = np.where(kin[:, None], 0.0, np.asarray(v, np.float64))
    sc.angular_velocity = np.where(kin[:, None], 0.0, np.asarray(w, np.float64))
    k, dist = world.cache()
    blob = base64.b64encode(np.ascontiguousarray(dist, dtype="<f4").tobytes()).decode("ascii")
    save_scene(sc, path, mesh_paths,
               extra_config={"dt": float(sc.dt), "iterations": int(sc.iterations), "sap_axis": int(world.axis())},
               extra={"ContactCache": {"pairs": int(k), "distances_f32": blob}})


def load_checkpoint(path: str):
    """(World, Scene) resumed from a file save_checkpoint wrote (see there); a plain scene file
    loads as a fresh world (axis 0, empty contact cache)."""
    import base64
    from .world import World
    with open(path) as f:
        data = json.load(f)
    cfg = data.get("Configuration", {})
    sc = load_scene(path, dt=float(cfg.get("dt", 1.0 / 60.0)), iterations=int(cfg.get("iterations", 10)))
    w = World(sc)
    w.set_axis(int(cfg.get("sap_axis", 0)))
    cache = data.get("ContactCache")
    if cache and int(cache.get("pairs", 0)) > 0:
        d = np.frombuffer(base64.b64decode(cache["distances_f32"]), dtype="<f4").reshape(-1, 8)
        if len(d) != int(cache["pairs"]):
            raise ValueError("ContactCache: distances do not match the pair count")
        w.set_cache(d)
    return w, sc
