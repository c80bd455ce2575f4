"""Seeded synthetic scenes for the rigid-body step.

A Scene is a host-side structure of arrays in FP64 that maps 1:1 onto pe_scene_desc
(include/pe_amd.h).  Every consumer (the HIP library, the CPU oracle, the compiled
reference) rounds these values to its own precision, so all of them start from the same
FP32 bits.

Generators (numpy PCG64, seed = 1000 + workload index unless given):
  tower_scene      -- ring tower of 20 x 10 bricks and a steel ball fired into it
                      (the shape of examples/bomb_demo/main.cpp:58-82)
  box_rain_scene   -- N boxes with U(0.2, 1.0) m edges and uniform random orientation
                      dropped onto a kinematic ground box
  batch_worlds     -- many independent worlds of log-uniform body counts, boxes and spheres
  mixed_world      -- boxes / spheres / cylinders mix over a ground box
  shape_mix_scene  -- boxes, spheres, cylinders and convex hulls (hulls of 32 Gaussian points)
                      over a triangle-mesh heightfield: every narrowphase pair type
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

SHAPE_BOX, SHAPE_SPHERE, SHAPE_CYLINDER, SHAPE_CONVEX, SHAPE_CONCAVE = 0, 1, 2, 3, 4
SHAPE_COMPOUND = 5


class MeshDesc(ctypes.Structure):
    """ctypes mirror of pe_mesh_desc."""
    _fields_ = [
        ("nv", ctypes.c_int32),
        ("vertices", ctypes.POINTER(ctypes.c_double)),
        ("nf", ctypes.c_int32),
        ("face_start", ctypes.POINTER(ctypes.c_int32)),
        ("face_index", ctypes.POINTER(ctypes.c_int32)),
        ("face_normal", ctypes.POINTER(ctypes.c_double)),
    ]


@dataclass
class Mesh:
    """A polygon mesh as handed to ConvexMeshShape::setMesh / ConcaveMeshShape::setMesh."""
    vertices: np.ndarray     # f64 [nv, 3]
    faces: list              # list of int vertex-index lists (counter-clockwise seen from outside)
    normals: np.ndarray      # f64 [nf, 3] outward unit normals

    def flat(self):
        start = np.zeros(len(self.faces) + 1, np.int32)
        start[1:] = np.cumsum([len(f) for f in self.faces])
        index = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32) for f in self.faces]), np.int32)
        return (np.ascontiguousarray(self.vertices, np.float64), start, index,
                np.ascontiguousarray(self.normals, np.float64))


class CompoundChild(ctypes.Structure):
    """ctypes mirror of pe_compound_child."""
    _fields_ = [("shape_type", ctypes.c_int32), ("mesh_index", ctypes.c_int32), ("shape_param", ctypes.c_double * 3),
                ("position", ctypes.c_double * 3), ("rotation", ctypes.c_double * 9), ("mass_ratio", ctypes.c_double)]


def compound_child(shape, param, pos=(0.0, 0.0, 0.0), rot=None, ratio=1.0, mesh=-1) -> CompoundChild:
    """CompoundShape::addShape's arguments: a child shape (SHAPE_BOX / SPHERE / CYLINDER / CONVEX with its
    parameters or scene mesh) at a local transform, with a mass ratio."""
    c = CompoundChild()
    c.shape_type, c.mesh_index = int(shape), int(mesh)
    p = list(param) + [0.0] * (3 - len(param))
    c.shape_param[:] = [float(x) for x in p[:3]]
    c.position[:] = [float(x) for x in pos]
    c.rotation[:] = [float(x) for x in (np.eye(3) if rot is None else np.asarray(rot, float)).ravel()]
    c.mass_ratio = float(ratio)
    return c


class SceneDesc(ctypes.Structure):
    """ctypes mirror of pe_scene_desc."""
    _fields_ = [
        ("n", ctypes.c_int32),
        ("shape_type", ctypes.POINTER(ctypes.c_int32)),
        ("shape_param", ctypes.POINTER(ctypes.c_double)),
        ("mass", ctypes.POINTER(ctypes.c_double)),
        ("position", ctypes.POINTER(ctypes.c_double)),
        ("rotation", ctypes.POINTER(ctypes.c_double)),
        ("linear_velocity", ctypes.POINTER(ctypes.c_double)),
        ("angular_velocity", ctypes.POINTER(ctypes.c_double)),
        ("friction", ctypes.POINTER(ctypes.c_double)),
        ("restitution", ctypes.POINTER(ctypes.c_double)),
        ("linear_damping", ctypes.POINTER(ctypes.c_double)),
        ("angular_damping", ctypes.POINTER(ctypes.c_double)),
        ("kinematic", ctypes.POINTER(ctypes.c_int32)),
        ("n_meshes", ctypes.c_int32),
        ("meshes", ctypes.POINTER(MeshDesc)),
        ("mesh_index", ctypes.POINTER(ctypes.c_int32)),
        ("ignore_collision", ctypes.POINTER(ctypes.c_int32)),
        ("lifetime", ctypes.POINTER(ctypes.c_double)),
        ("compound_start", ctypes.POINTER(ctypes.c_int32)),
        ("compound_children", ctypes.POINTER(CompoundChild)),
    ]


class CollisionEvent(ctypes.Structure):
    """ctypes mirror of pe_collision_event."""
    _fields_ = [("body_a", ctypes.c_int32), ("body_b", ctypes.c_int32), ("position", ctypes.c_float * 3),
                ("normal", ctypes.c_float * 3), ("depth", ctypes.c_float), ("velocity", ctypes.c_float * 3)]


class RayHit(ctypes.Structure):
    """ctypes mirror of pe_ray_hit."""
    _fields_ = [("hit", ctypes.c_int32), ("body", ctypes.c_int32), ("distance", ctypes.c_float),
                ("point", ctypes.c_float * 3), ("normal", ctypes.c_float * 3)]


class JointDesc(ctypes.Structure):
    """ctypes mirror of pe_joint_desc (a ball / hinge / slider joint between two bodies)."""
    _fields_ = [("type", ctypes.c_int32), ("body_a", ctypes.c_int32), ("body_b", ctypes.c_int32),
                ("use_limits", ctypes.c_int32), ("use_motor", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("anchor_a", ctypes.c_double * 3), ("anchor_b", ctypes.c_double * 3),
                ("axis_a", ctypes.c_double * 3), ("axis_b", ctypes.c_double * 3),
                ("min_angle", ctypes.c_double), ("max_angle", ctypes.c_double), ("target_speed", ctypes.c_double)]


JOINT_BALL, JOINT_HINGE, JOINT_SLIDER = 0, 1, 2


def joint(kind, a, b, anchor_a=(0, 0, 0), anchor_b=(0, 0, 0), axis_a=(1, 0, 0), axis_b=(1, 0, 0),
          limits=None, motor_speed=None) -> JointDesc:
    """A joint as World::addConstraint receives it: kind JOINT_*, bodies a and b, body-local anchors and
    axes, hinge limits (min, max) in radians and a hinge motor's target speed."""
    d = JointDesc()
    d.type, d.body_a, d.body_b = kind, a, b
    d.anchor_a[:], d.anchor_b[:] = [float(x) for x in anchor_a], [float(x) for x in anchor_b]
    d.axis_a[:], d.axis_b[:] = [float(x) for x in axis_a], [float(x) for x in axis_b]
    d.use_limits = 1 if limits is not None else 0
    d.min_angle, d.max_angle = (float(limits[0]), float(limits[1])) if limits is not None else (0.0, 0.0)
    d.use_motor = 1 if motor_speed is not None else 0
    d.target_speed = float(motor_speed) if motor_speed is not None else 0.0
    return d


class WorldParams(ctypes.Structure):
    """ctypes mirror of pe_world_params."""
    _fields_ = [("dt", ctypes.c_double), ("gravity", ctypes.c_double * 3), ("iterations", ctypes.c_int32)]


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


@dataclass
class Scene:
    shape_type: np.ndarray           # int32 [n]
    shape_param: np.ndarray          # f64 [n,4]
    mass: np.ndarray                 # f64 [n]
    position: np.ndarray             # f64 [n,3]
    rotation: np.ndarray             # f64 [n,3,3] (row-major basis)
    linear_velocity: np.ndarray      # f64 [n,3]
    angular_velocity: np.ndarray     # f64 [n,3]
    friction: np.ndarray             # f64 [n]
    restitution: np.ndarray          # f64 [n]
    kinematic: np.ndarray            # int32 [n]
    linear_damping: np.ndarray = None
    angular_damping: np.ndarray = None
    dt: float = 1.0 / 60.0
    gravity: tuple = (0.0, -9.8, 0.0)
    iterations: int = 10
    name: str = ""
    meshes: list = None              # [Mesh] of the convex / concave bodies
    mesh_index: np.ndarray = None    # int32 [n] mesh of each convex / concave body (-1 otherwise)
    joints: list = field(default_factory=list)          # [JointDesc] in World::addConstraint order
    ignore_pairs: list = field(default_factory=list)    # [(a, b)] bodies that never collide (addIgnoreCollisionId)
    ignore_collision: np.ndarray = None                 # int32 [n] 0/1 (setIgnoreCollision), or None
    lifetime: np.ndarray = None                         # f64 [n] seconds (setLifeTime; >= 1e30 never), or None
    sleep: tuple = None                                 # (lin_vel2, ang_vel2, time) World::setSleep*Threshold, or None
    compounds: dict = field(default_factory=dict)       # body -> [CompoundChild] of SHAPE_COMPOUND bodies
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n(self) -> int:
        return int(self.shape_type.shape[0])

    def n_dynamic(self) -> int:
        return int((self.kinematic == 0).sum())

    def contiguous(self) -> "Scene":
        for f in ("shape_param", "mass", "position", "rotation", "linear_velocity", "angular_velocity",
                  "friction", "restitution", "linear_damping", "angular_damping"):
            v = getattr(self, f)
            if v is not None:
                setattr(self, f, np.ascontiguousarray(v, dtype=np.float64))
        self.shape_type = np.ascontiguousarray(self.shape_type, dtype=np.int32)
        self.kinematic = np.ascontiguousarray(self.kinematic, dtype=np.int32)
        if self.linear_damping is None:
            self.linear_damping = np.zeros(self.n)
        if self.angular_damping is None:
            self.angular_damping = np.zeros(self.n)
        if self.meshes is None:
            self.meshes = []
        if self.mesh_index is None:
            self.mesh_index = np.full(self.n, -1, np.int32)
        self.mesh_index = np.ascontiguousarray(self.mesh_index, dtype=np.int32)
        return self

    def desc(self) -> SceneDesc:
        """Builds a pe_scene_desc pointing at this scene's arrays (kept alive by the scene)."""
        self.contiguous()
        d = SceneDesc(self.n, _iptr(self.shape_type), _dptr(self.shape_param), _dptr(self.mass),
                      _dptr(self.position), _dptr(self.rotation), _dptr(self.linear_velocity),
                      _dptr(self.angular_velocity), _dptr(self.friction), _dptr(self.restitution),
                      _dptr(self.linear_damping), _dptr(self.angular_damping), _iptr(self.kinematic))
        if self.meshes:
            flats = [m.flat() for m in self.meshes]
            arr = (MeshDesc * len(flats))()
            for k, (v, st, ix, fn) in enumerate(flats):
                arr[k] = MeshDesc(v.shape[0], _dptr(v), len(st) - 1, _iptr(st), _iptr(ix), _dptr(fn))
            self._keep[:] = [flats, arr]
            d.n_meshes = len(flats)
            d.meshes = arr
            d.mesh_index = _iptr(self.mesh_index)
        if self.ignore_collision is not None:
            self.ignore_collision = np.ascontiguousarray(self.ignore_collision, dtype=np.int32)
            d.ignore_collision = _iptr(self.ignore_collision)
        if self.lifetime is not None:
            self.lifetime = np.ascontiguousarray(self.lifetime, dtype=np.float64)
            d.lifetime = _dptr(self.lifetime)
        if self.compounds:
            start = np.zeros(self.n + 1, np.int32)
            kids = []
            for i in range(self.n):
                kids += list(self.compounds.get(i, []))
                start[i + 1] = len(kids)
            arr = (CompoundChild * max(len(kids), 1))(*kids)
            self._keep += [start, arr]
            d.compound_start = _iptr(start)
            d.compound_children = arr
        return d

    def params(self) -> WorldParams:
        return WorldParams(self.dt, (ctypes.c_double * 3)(*self.gravity), self.iterations)

    def subset(self, idx) -> "Scene":
        self.contiguous()
        idx = np.asarray(idx)
        return Scene(self.shape_type[idx], self.shape_param[idx], self.mass[idx], self.position[idx],
                     self.rotation[idx], self.linear_velocity[idx], self.angular_velocity[idx],
                     self.friction[idx], self.restitution[idx], self.kinematic[idx],
                     None if self.linear_damping is None else self.linear_damping[idx],
                     None if self.angular_damping is None else self.angular_damping[idx],
                     self.dt, self.gravity, self.iterations, self.name,
                     list(self.meshes), self.mesh_index[idx],
                     ignore_collision=None if self.ignore_collision is None else self.ignore_collision[idx],
                     lifetime=None if self.lifetime is None else self.lifetime[idx], sleep=self.sleep,
                     compounds={k: list(self.compounds[int(i)]) for k, i in enumerate(idx) if int(i) in self.compounds}
                     ).contiguous()

    def footprint_bodies(self) -> int:
        return self.n


def concat(scenes, name="") -> Scene:
    """Concatenates bodies of several scenes into one (same dt / gravity / iterations)."""
    scenes = [s.contiguous() for s in scenes]
    s0 = scenes[0]
    cat = lambda f: np.concatenate([getattr(s, f) for s in scenes])
    meshes, mi, off = [], [], 0
    for s in scenes:
        meshes += s.meshes
        mi.append(np.where(s.mesh_index >= 0, s.mesh_index + off, -1))
        off += len(s.meshes)
    return Scene(cat("shape_type"), cat("shape_param"), cat("mass"), cat("position"), cat("rotation"),
                 cat("linear_velocity"), cat("angular_velocity"), cat("friction"), cat("restitution"),
                 cat("kinematic"), cat("linear_damping"), cat("angular_damping"),
                 s0.dt, s0.gravity, s0.iterations, name, meshes, np.concatenate(mi)).contiguous()


# ------------------------------------------------------------------ rotations
def rot_y(angle: float) -> np.ndarray:
    """Basis of Eigen::AngleAxis(angle, UnitY)."""
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def quat_to_matrix(q: np.ndarray) -> np.ndarray:
    """Rotation matrices from unit quaternions (x, y, z, w), shape [..., 4] -> [..., 3, 3]."""
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    m = np.empty(q.shape[:-1] + (3, 3))
    m[..., 0, 0] = 1 - 2 * (y * y + z * z)
    m[..., 0, 1] = 2 * (x * y - z * w)
    m[..., 0, 2] = 2 * (x * z + y * w)
    m[..., 1, 0] = 2 * (x * y + z * w)
    m[..., 1, 1] = 1 - 2 * (x * x + z * z)
    m[..., 1, 2] = 2 * (y * z - x * w)
    m[..., 2, 0] = 2 * (x * z - y * w)
    m[..., 2, 1] = 2 * (y * z + x * w)
    m[..., 2, 2] = 1 - 2 * (x * x + y * y)
    return m


def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (Shoemake's subgroup algorithm)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=-1)


# ------------------------------------------------------------------ builders
class SceneBuilder:
    """Collects bodies one by one (the order of World::addRigidBody) into a Scene."""

    def __init__(self):
        self.rows = []
        self.meshes = []

    def add_mesh(self, mesh: "Mesh") -> int:
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add(self, shape, param, mass, pos, rot=None, v=(0, 0, 0), w=(0, 0, 0), mu=0.5, e=0.0, kin=0, mesh=-1):
        p = np.zeros(4)
        p[:len(param)] = param
        self.rows.append((shape, p, mass, np.asarray(pos, float),
                          np.eye(3) if rot is None else np.asarray(rot, float),
                          np.asarray(v, float), np.asarray(w, float), mu, e, kin, mesh))
        return len(self.rows) - 1

    def scene(self, **kw) -> Scene:
        r = self.rows
        return Scene(np.array([x[0] for x in r], np.int32), np.stack([x[1] for x in r]),
                     np.array([x[2] for x in r], float), np.stack([x[3] for x in r]),
                     np.stack([x[4] for x in r]), np.stack([x[5] for x in r]),
                     np.stack([x[6] for x in r]), np.array([x[7] for x in r], float),
                     np.array([x[8] for x in r], float), np.array([x[9] for x in r], np.int32),
                     meshes=list(self.meshes), mesh_index=np.array([x[10] for x in r], np.int32),
                     **kw).contiguous()


def tower_scene(layers: int = 20, per_layer: int = 10, seed: int = 1000) -> Scene:
    """Ring tower of layers x per_layer bricks (1.0 x 0.5 x 0.5 m, 1000 kg/m^3) on a kinematic
    50 x 1 x 50 m ground box, and a r = 0.5 m steel sphere (7800 kg/m^3) at 30 m/s."""
    b = SceneBuilder()
    b.add(SHAPE_BOX, (50.0, 1.0, 50.0), 1.0, (0.0, -0.5, 0.0), kin=1)
    L, H, W = 1.0, 0.5, 0.5
    angle = 2.0 * math.pi / per_layer
    radius = 1.25 * L / angle
    mass = 1000.0 * L * H * W
    for i in range(layers):
        off = (i % 2) * angle / 2.0
        for k in range(per_layer):
            th = off + k * angle
            pos = (radius * math.cos(th), H * (0.5 + i), radius * math.sin(th))
            b.add(SHAPE_BOX, (W, H, L), mass, pos, rot_y(-th))
    r = 0.5
    b.add(SHAPE_SPHERE, (r,), 7800.0 * 4.0 / 3.0 * math.pi * r ** 3, (0.0, 2.5, 12.0), v=(0.0, 0.0, -30.0))
    return b.scene(name=f"tower{layers}x{per_layer}")


def _rejection_centres(rng, radii, lo, hi, gap):
    """Places spheres of the given radii uniformly in the box [lo, hi] so that no two come
    within `gap` of each other (bounding-sphere rejection, grid accelerated)."""
    n = len(radii)
    cell = 2.0 * radii.max() + gap
    grid = {}
    out = np.empty((n, 3))
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    for i in range(n):
        for _ in range(10000):
            c = lo + rng.random(3) * (hi - lo)
            key = tuple((c // cell).astype(int))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            d = out[j] - c
                            if d @ d < (radii[i] + radii[j] + gap) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                out[i] = c
                grid.setdefault(key, []).append(i)
                break
        else:
            raise RuntimeError("rejection sampling failed: volume too small")
    return out


def box_rain_scene(n: int = 10000, seed: int = 1001, lo=(-20, 1, -20), hi=(20, 81, 20),
                   ground=(200.0, 1.0, 200.0), density=1000.0, mu=0.5) -> Scene:
    """n boxes, edges iid U(0.2, 1.0) m, uniform orientation, centres uniform in [lo, hi]
    with 0.05 m bounding-sphere rejection, over a kinematic ground box (top face at y = 0)."""
    rng = np.random.default_rng(seed)
    size = rng.uniform(0.2, 1.0, (n, 3))
    rot = quat_to_matrix(random_quaternions(rng, n))
    radii = 0.5 * np.linalg.norm(size, axis=1)
    pos = _rejection_centres(rng, radii, lo, hi, 0.05)
    b = SceneBuilder()
    b.add(SHAPE_BOX, ground, 1.0, (0.0, -ground[1] / 2.0, 0.0), mu=mu, kin=1)
    for i in range(n):
        b.add(SHAPE_BOX, size[i], density * size[i].prod(), pos[i], rot[i], mu=mu)
    return b.scene(name=f"box_rain{n}")


def mixed_world(n: int, seed: int, p_sphere: float = 0.5, p_cyl: float = 0.0, density=1000.0) -> Scene:
    """One small world: n bodies, each a box (edges U(0.2, 1.0)) or a sphere (r U(0.1, 0.6))
    (or a cylinder r U(0.1, 0.4), h U(0.2, 1.0) with probability p_cyl), dropped onto its own
    kinematic ground box.  The drop volume grows with n."""
    rng = np.random.default_rng(seed)
    half = 2.0 + 0.5 * math.sqrt(n)
    height = 2.0 + n / 16.0
    u = rng.random(n)
    kind = np.where(u < p_cyl, SHAPE_CYLINDER, np.where(u < p_cyl + p_sphere, SHAPE_SPHERE, SHAPE_BOX))
    size = rng.uniform(0.2, 1.0, (n, 3))
    rad = rng.uniform(0.1, 0.6, n)
    cyl_r = rng.uniform(0.1, 0.4, n)
    cyl_h = rng.uniform(0.2, 1.0, n)
    rot = quat_to_matrix(random_quaternions(rng, n))
    bound = np.where(kind == SHAPE_BOX, 0.5 * np.linalg.norm(size, axis=1),
                     np.where(kind == SHAPE_SPHERE, rad, np.sqrt(cyl_r ** 2 + 0.25 * cyl_h ** 2)))
    pos = _rejection_centres(rng, bound, (-half, 1.0, -half), (half, 1.0 + height, half), 0.05)
    b = SceneBuilder()
    g = 2.0 * half + 4.0
    b.add(SHAPE_BOX, (g, 1.0, g), 1.0, (0.0, -0.5, 0.0), kin=1)
    for i in range(n):
        if kind[i] == SHAPE_BOX:
            b.add(SHAPE_BOX, size[i], density * size[i].prod(), pos[i], rot[i])
        elif kind[i] == SHAPE_SPHERE:
            b.add(SHAPE_SPHERE, (rad[i],), density * 4.0 / 3.0 * math.pi * rad[i] ** 3, pos[i])
        else:
            b.add(SHAPE_CYLINDER, (cyl_r[i], cyl_h[i]), density * math.pi * cyl_r[i] ** 2 * cyl_h[i], pos[i], rot[i])
    return b.scene(name=f"world{n}")


def grid_world(n: int, seed: int, p_sphere: float = 0.5, density: float = 1000.0) -> Scene:
    """One world of n bodies (boxes with U(0.2, 1.0) m edges and uniform orientation, or spheres
    with r U(0.1, 0.6) m, each with probability 1/2) dropped onto its own kinematic ground box.
    Centres are jittered inside the cells of a random subset of a 1.8 m lattice, which keeps
    bounding spheres at least 0.05 m apart without a rejection loop (vectorised, so hundreds
    of thousands of bodies generate in seconds)."""
    rng = np.random.default_rng(seed)
    cell = 1.8                                   # >= 2 * max bounding radius (0.866) + 0.05
    side = max(2, int(math.ceil(math.sqrt(n / 2.0))))
    layers = int(math.ceil(n / (side * side))) + 1
    cells = rng.permutation(side * side * layers)[:n]
    cx, rest = cells % side, cells // side
    cz, cy = rest % side, rest // side
    is_sph = rng.random(n) < p_sphere
    size = rng.uniform(0.2, 1.0, (n, 3))
    rad = rng.uniform(0.1, 0.6, n)
    rot = quat_to_matrix(random_quaternions(rng, n))
    bound = np.where(is_sph, rad, 0.5 * np.linalg.norm(size, axis=1))
    slack = np.maximum(cell / 2.0 - bound - 0.025, 0.0)
    jit = rng.uniform(-1.0, 1.0, (n, 3)) * slack[:, None]
    half = side * cell / 2.0
    pos = np.stack([(cx + 0.5) * cell - half, 1.0 + (cy + 0.5) * cell, (cz + 0.5) * cell - half], axis=1) + jit
    m = n + 1
    st = np.zeros(m, np.int32)
    sp = np.zeros((m, 4))
    mass = np.ones(m)
    P = np.zeros((m, 3))
    Rm = np.tile(np.eye(3), (m, 1, 1))
    kin = np.zeros(m, np.int32)
    g = 2.0 * half + 4.0
    sp[0, :3] = (g, 1.0, g)
    P[0] = (0.0, -0.5, 0.0)
    kin[0] = 1
    st[1:] = np.where(is_sph, SHAPE_SPHERE, SHAPE_BOX)
    sp[1:, :3] = np.where(is_sph[:, None], np.stack([rad, np.zeros(n), np.zeros(n)], 1), size)
    mass[1:] = np.where(is_sph, density * 4.0 / 3.0 * math.pi * rad ** 3, density * size.prod(axis=1))
    P[1:] = pos
    Rm[1:] = np.where(is_sph[:, None, None], np.eye(3)[None], rot)
    return Scene(st, sp, mass, P, Rm, np.zeros((m, 3)), np.zeros((m, 3)), np.full(m, 0.5), np.zeros(m), kin,
                 name=f"grid_world{n}").contiguous()


def world_seed(seed: int, w: int) -> int:
    """Per-world seed: SplitMix64 of (seed, w)."""
    z = (seed * 0x9E3779B97F4A7C15 + (w + 1) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return (z ^ (z >> 31)) & 0x7FFFFFFF


def log_uniform_counts(rng, n_worlds: int, lo: int, hi: int) -> np.ndarray:
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n_worlds))).astype(int)


def batch_worlds(n_worlds: int = 4096, lo: int = 8, hi: int = 512, seed: int = 1003):
    """Independent worlds with body counts log-uniform in [lo, hi]; boxes and spheres (p = 1/2)."""
    rng = np.random.default_rng(seed)
    counts = log_uniform_counts(rng, n_worlds, lo, hi)
    return [grid_world(int(c), world_seed(seed, w)) for w, c in enumerate(counts)]


# ------------------------------------------------------------------ meshes
def hull_mesh(points: np.ndarray) -> "Mesh":
    """Convex hull of a point cloud as a triangle mesh (qhull), faces counter-clockwise seen
    from outside, normals the normalised face cross products (n / sqrt(n . n), as
    Mesh::perFaceNormal computes them when a scene file is loaded, so meshes survive a save and
    load bit for bit)."""
    from scipy.spatial import ConvexHull
    h = ConvexHull(points)
    used = np.unique(h.simplices)
    remap = -np.ones(points.shape[0], np.int64)
    remap[used] = np.arange(used.size)
    verts = points[used]
    centre = verts.mean(axis=0)
    faces, normals = [], []
    for tri in h.simplices:
        a, b, c = (int(remap[t]) for t in tri)
        n = np.cross(verts[b] - verts[a], verts[c] - verts[a])
        if n @ (verts[a] - centre) < 0:
            b, c = c, b
            n = -n
        faces.append([a, b, c])
        normals.append(n / math.sqrt(float(n @ n)))   # as a scene file's loader computes them
    return Mesh(verts, faces, np.array(normals))


def gaussian_hull(rng: np.random.Generator, npoints: int = 32, radius: float = 0.5) -> "Mesh":
    """Hull of npoints standard-normal points scaled so the farthest lies at `radius`."""
    p = rng.standard_normal((npoints, 3))
    p *= radius / np.linalg.norm(p, axis=1).max()
    return hull_mesh(p)


def heightfield_mesh(rng: np.random.Generator, nx: int = 64, nz: int = 64, size=(100.0, 100.0),
                     amplitude: float = 1.0) -> "Mesh":
    """nx x nz vertex grid over size[0] x size[1] m centred on the origin, heights U(-a, a);
    two triangles per cell, normals pointing up (+y)."""
    xs = np.linspace(-size[0] / 2.0, size[0] / 2.0, nx)
    zs = np.linspace(-size[1] / 2.0, size[1] / 2.0, nz)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = rng.uniform(-amplitude, amplitude, X.shape)
    verts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    vid = lambda i, k: i * nz + k
    faces, normals = [], []
    for i in range(nx - 1):
        for k in range(nz - 1):
            for tri in ((vid(i, k), vid(i, k + 1), vid(i + 1, k + 1)), (vid(i, k), vid(i + 1, k + 1), vid(i + 1, k))):
                a, b, c = tri
                n = np.cross(verts[b] - verts[a], verts[c] - verts[a])
                if n[1] < 0:
                    b, c = c, b
                    n = -n
                faces.append([a, b, c])
                normals.append(n / math.sqrt(float(n @ n)))   # as a scene file's loader computes them
    return Mesh(verts, faces, np.array(normals))


def shape_mix_scene(n: int = 50000, seed: int = 1002, volume=(100.0, 60.0, 100.0), grid: int = 64,
                    ground_size=None, fractions=(0.40, 0.25, 0.15, 0.20), density: float = 1000.0,
                    hull_points: int = 32, mu: float = 0.5, hull_pool: int = 0) -> Scene:
    """Mixed-shape world: boxes (edges U(0.2, 1.0) m), spheres (r U(0.1, 0.6) m), cylinders
    (r U(0.1, 0.4) m, h U(0.2, 1.0) m) and convex hulls (hull of `hull_points` Gaussian points
    scaled to 0.5 m) in the given fractions, all with uniform random orientation (spheres
    unrotated), over a kinematic grid x grid heightfield with +-1 m uniform noise.  Each hull body
    gets its own mesh, or with hull_pool > 0 one of hull_pool shared meshes (fewer meshes to
    prepare, as for many small served requests).  Centres are jittered in the cells of a 1.8 m
    lattice filling the volume from 2 m above the ground, which keeps bounding spheres 0.05 m
    apart without a rejection loop."""
    rng = np.random.default_rng(seed)
    gx, gz = ground_size if ground_size is not None else (volume[0], volume[2])
    b = SceneBuilder()
    hf = b.add_mesh(heightfield_mesh(rng, grid, grid, (gx, gz), 1.0))
    b.add(SHAPE_CONCAVE, (), 1.0, (0.0, 0.0, 0.0), kin=1, mesh=hf, mu=mu)
    cell = 1.8
    nx, nz = max(1, int(volume[0] // cell)), max(1, int(volume[2] // cell))
    ny = int(math.ceil(n / (nx * nz))) + 1
    cells = rng.permutation(nx * ny * nz)[:n]
    cx, rest = cells % nx, cells // nx
    cz, cy = rest % nz, rest // nz
    u = rng.random(n)
    f0, f1, f2, _ = np.cumsum(fractions)
    kind = np.where(u < f0, SHAPE_BOX, np.where(u < f1, SHAPE_SPHERE, np.where(u < f2, SHAPE_CYLINDER, SHAPE_CONVEX)))
    size = rng.uniform(0.2, 1.0, (n, 3))
    rad = rng.uniform(0.1, 0.6, n)
    cr = rng.uniform(0.1, 0.4, n)
    ch = rng.uniform(0.2, 1.0, n)
    rot = quat_to_matrix(random_quaternions(rng, n))
    bound = np.select([kind == SHAPE_BOX, kind == SHAPE_SPHERE, kind == SHAPE_CYLINDER],
                      [0.5 * np.linalg.norm(size, axis=1), rad, np.sqrt(cr ** 2 + 0.25 * ch ** 2)], 0.5)
    slack = np.maximum(cell / 2.0 - bound - 0.025, 0.0)
    jit = rng.uniform(-1.0, 1.0, (n, 3)) * slack[:, None]
    pos = np.stack([(cx + 0.5) * cell - nx * cell / 2.0, 2.0 + (cy + 0.5) * cell,
                    (cz + 0.5) * cell - nz * cell / 2.0], axis=1) + jit
    pool = []
    for i in range(n):
        k = kind[i]
        if k == SHAPE_BOX:
            b.add(SHAPE_BOX, size[i], density * size[i].prod(), pos[i], rot[i], mu=mu)
        elif k == SHAPE_SPHERE:
            b.add(SHAPE_SPHERE, (rad[i],), density * 4.0 / 3.0 * math.pi * rad[i] ** 3, pos[i], mu=mu)
        elif k == SHAPE_CYLINDER:
            b.add(SHAPE_CYLINDER, (cr[i], ch[i]), density * math.pi * cr[i] ** 2 * ch[i], pos[i], rot[i], mu=mu)
        else:
            if hull_pool > 0:
                if not pool:
                    for _ in range(hull_pool):
                        m = gaussian_hull(rng, hull_points, 0.5)
                        vol = abs(sum(np.dot(m.vertices[f[0]], np.cross(m.vertices[f[1]], m.vertices[f[2]]))
                                      for f in m.faces)) / 6.0
                        pool.append((b.add_mesh(m), vol))
                mid, vol = pool[int(rng.integers(hull_pool))]
            else:
                m = gaussian_hull(rng, hull_points, 0.5)
                vol = abs(sum(np.dot(m.vertices[f[0]], np.cross(m.vertices[f[1]], m.vertices[f[2]])) for f in m.faces)) / 6.0
                mid = b.add_mesh(m)
            b.add(SHAPE_CONVEX, (), density * vol, pos[i], rot[i], mu=mu, mesh=mid)
    return b.scene(name=f"shape_mix{n}")


def served_shape_world(n: int, seed: int, hull_pool: int = 8, grid: int = 16) -> Scene:
    """A served request's world in the 50k shape mix's proportions: n bodies (40 % boxes, 25 % spheres,
    15 % cylinders, 20 % convex hulls from a pool of hull_pool meshes) over a grid x grid
    heightfield sized to a footprint that stacks the lattice about four cells high."""
    side = 1.8 * max(3, int(math.ceil(math.sqrt(n / 4.0))))
    return shape_mix_scene(n, seed=seed, volume=(side, 4 * 1.8, side), grid=grid,
                           ground_size=(side + 4.0, side + 4.0), hull_pool=hull_pool)


def _axis_angle(axis, angle):
    """Basis of Eigen::AngleAxis(angle, axis) (Rodrigues)."""
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1.0 - math.cos(angle)) * (K @ K)


def joint_demo_scene(gravity=(0.0, -9.8, 0.0), motor_speed=2.0) -> Scene:
    """The joints of the reference's constraints demo (examples/constraints_demo/main.cpp:22-163): a
    two-link pendulum of ball joints (three joints, sphere and cylinder links), a hinged arm with angle
    limits and a second hinged arm driven by a motor, and a box sliding on a rail between two end
    stops; jointed bodies never collide with each other (addIgnoreCollisionId).  Under gravity, on a
    kinematic ground."""
    b = SceneBuilder()
    pi = math.pi
    b.add(SHAPE_BOX, (30.0, 10.0, 30.0), 10000.0, (0, -5, 0), kin=1, e=0.5)                          # 0 ground
    b.add(SHAPE_SPHERE, (0.5,), 1.0, (0, 0.5, 0), kin=1)                                             # 1 ball base
    s6, c6 = math.sin(pi / 6), math.cos(pi / 6)
    b.add(SHAPE_CYLINDER, (0.2, 6.0), 1.0, (-3.3 * s6, 3.3 * c6 + 0.5, 0), rot=_axis_angle((0, 0, 1), pi / 6))  # 2
    b.add(SHAPE_SPHERE, (0.4,), 1.0, (-6.6 * s6, 6.6 * c6 + 0.5, 0))                                 # 3
    b.add(SHAPE_CYLINDER, (0.2, 6.0), 1.0, (-3.3 - 6.6 * s6, 6.6 * c6 + 0.5, 0), rot=_axis_angle((0, 0, 1), pi / 2))  # 4
    R5, o5 = _axis_angle((1, 0, 0), pi / 6), np.array([4.0, 8.0, 0.0])
    b.add(SHAPE_CYLINDER, (0.2, 1.0), 1.0, o5, rot=R5, kin=1)                                         # 5 pole base
    R2, o2 = _axis_angle((0, 0, -1), pi / 2), np.array([2.0, 0.0, 0.0])
    b.add(SHAPE_CYLINDER, (0.2, 4.0), 1.0, R5 @ o2 + o5, rot=R5 @ R2)                                 # 6 arm
    R3, o3 = _axis_angle((1, 0, 0), pi / 2), np.array([4.0, 0.0, 0.75])
    b.add(SHAPE_CYLINDER, (0.2, 2.0), 1.0, R5 @ o3 + o5, rot=R5 @ R3)                                 # 7 second arm
    R0, o0 = _axis_angle((0, 1, 0), -pi / 4), np.array([-1.0, 6.0, 0.0])
    Rt = R0 @ _axis_angle((0, 0, 1), pi / 3)
    b.add(SHAPE_BOX, (0.3, 10.0, 0.3), 1.0, o0, rot=Rt, kin=1)                                       # 8 rail
    b.add(SHAPE_BOX, (1.0, 0.4, 1.0), 1.0, Rt @ np.array([0, 5.2, 0]) + o0, rot=Rt, kin=1)           # 9 stop
    b.add(SHAPE_BOX, (1.0, 0.4, 1.0), 1.0, Rt @ np.array([0, -5.2, 0]) + o0, rot=Rt, kin=1)          # 10 stop
    b.add(SHAPE_BOX, (1.0, 1.2, 1.0), 1.0, Rt @ np.array([0, 4.4, 0]) + o0, rot=Rt)                  # 11 slider
    sc = b.scene(gravity=tuple(gravity), name="joint_demo")
    sc.joints = [
        joint(JOINT_BALL, 1, 2, anchor_b=(0, -3.3, 0)),
        joint(JOINT_BALL, 3, 4, anchor_b=(0, -3.3, 0)),
        joint(JOINT_BALL, 2, 3, anchor_a=(0, 3.3, 0)),
        joint(JOINT_HINGE, 5, 6, anchor_b=(0, -2, 0), axis_a=(0, 1, 0), axis_b=(-1, 0, 0), limits=(-pi / 3, pi / 3)),
        joint(JOINT_HINGE, 6, 7, anchor_a=(0, 2, 0), anchor_b=(0, -0.75, 0), axis_a=(0, 1, 0), axis_b=(0, 0, -1),
              motor_speed=motor_speed),
        joint(JOINT_SLIDER, 8, 11, axis_a=(0, 1, 0), axis_b=(0, 1, 0)),
    ]
    sc.ignore_pairs = [(1, 2), (3, 2), (3, 4), (5, 6), (6, 7), (8, 11)]
    return sc


def joint_chain_scene(chains=48, links=8, seed=5, spacing=2.5) -> Scene:
    """Hanging chains of ball joints under gravity: each chain a kinematic anchor sphere and `links`
    capsule-like cylinders (radius 0.15, length 1) joined end to end, started swinging sideways;
    neighbouring links ignore each other (addIgnoreCollisionId), the chains' ends strike a ground
    box.  Joint colouring alternates along a chain, so a colour holds one joint of many chains."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    side = int(math.ceil(math.sqrt(chains)))
    b.add(SHAPE_BOX, (side * spacing + 10.0, 1.0, side * spacing + 10.0), 1.0, (0.0, -0.5, 0.0), kin=1)
    top = links * 1.0 + 2.0
    joints, ignore = [], []
    for c in range(chains):
        x = (c % side - side / 2) * spacing
        z = (c // side - side / 2) * spacing
        anchor = b.add(SHAPE_SPHERE, (0.2,), 1.0, (x, top, z), kin=1)
        tilt = _axis_angle((0.0, 0.0, 1.0), float(rng.uniform(0.3, 0.9)))
        prev, prev_end = anchor, np.zeros(3)
        for k in range(links):
            centre = np.array([x, top, z]) + tilt @ np.array([0.0, -(k + 0.5), 0.0])
            body = b.add(SHAPE_CYLINDER, (0.15, 1.0), 1.0, tuple(centre), rot=tilt)
            joints.append(joint(JOINT_BALL, prev, body, anchor_a=tuple(prev_end), anchor_b=(0.0, 0.5, 0.0)))
            ignore.append((prev, body))
            prev, prev_end = body, np.array([0.0, -0.5, 0.0])
    sc = b.scene(name=f"chains{chains}x{links}")
    sc.joints = joints
    sc.ignore_pairs = ignore
    return sc


def compound_demo_scene(count=10, sleep=True) -> Scene:
    """The reference's compound demo (examples/compound_demo/main.cpp:10-86): a kinematic ground box,
    a box, a sphere and a cylinder, and `count` compound bodies stacked 5 m apart -- a unit box with a
    cylinder above and one below and four horizontal cylinders around it (mass ratios 0.3, 0.2, 0.1,
    0.1 x 4) -- each placed at its stack position minus the children's centroid, as the demo does
    with CompoundShape::init's result; friction 0.5, restitution 0.5, angular damping 0.8; the demo's
    sleep thresholds (0.01, 0.01, 1 s)."""
    pi = math.pi
    kids = [compound_child(SHAPE_BOX, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), ratio=0.3),
            compound_child(SHAPE_CYLINDER, (0.2, 2.0), (0.0, 1.5, 0.0), ratio=0.2),
            compound_child(SHAPE_CYLINDER, (0.2, 1.0), (0.0, -1.0, 0.0), ratio=0.1),
            compound_child(SHAPE_CYLINDER, (0.2, 1.0), (1.0, 0.0, 0.0), _axis_angle((0, 0, 1), pi / 2), 0.1),
            compound_child(SHAPE_CYLINDER, (0.2, 1.0), (-1.0, 0.0, 0.0), _axis_angle((0, 0, 1), pi / 2), 0.1),
            compound_child(SHAPE_CYLINDER, (0.2, 1.0), (0.0, 0.0, 1.0), _axis_angle((1, 0, 0), pi / 2), 0.1),
            compound_child(SHAPE_CYLINDER, (0.2, 1.0), (0.0, 0.0, -1.0), _axis_angle((1, 0, 0), pi / 2), 0.1)]
    ratios = np.array([k.mass_ratio for k in kids])
    centroid = sum(np.array(k.position[:]) * k.mass_ratio for k in kids) / ratios.sum()
    b = SceneBuilder()
    b.add(SHAPE_BOX, (30.0, 10.0, 30.0), 10000.0, (0, -5, 0), kin=1, e=0.5)
    b.add(SHAPE_BOX, (1.0, 1.0, 1.0), 1.0, (0, 2, 0), e=0.5)
    b.add(SHAPE_SPHERE, (0.5,), 1.0, (0.1, 4, 0.1), e=0.5)
    b.add(SHAPE_CYLINDER, (0.4, 1.0), 1.0, (0, 6, 0), e=0.5)
    first = len(b.rows)
    for i in range(count):
        b.add(SHAPE_COMPOUND, (), 1.0, tuple(np.array([0.0, 10.0 + 5.0 * i, 0.0]) - centroid), e=0.5)
    sc = b.scene(name=f"compound_demo{count}")
    sc.angular_damping = np.full(sc.n, 0.8)
    sc.angular_damping[0] = 0.0
    sc.compounds = {first + i: list(kids) for i in range(count)}
    if sleep:
        sc.sleep = (0.01, 0.01, 1.0)
    return sc


def compound_pile_scene(n=24, seed=9, grid=10, size=(12.0, 12.0)) -> Scene:
    """Compound bodies over a heightfield: each a box with a sphere, a cylinder and a convex hull
    child at random offsets and orientations (mass ratios 0.4, 0.2, 0.2, 0.2), dropped in a column so
    they land on the terrain and on each other, with a few plain boxes and spheres among them."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    hf = b.add_mesh(heightfield_mesh(rng, grid, grid, size, 0.5))
    b.add(SHAPE_CONCAVE, (), 1.0, (0.0, 0.0, 0.0), kin=1, mesh=hf)
    hull = b.add_mesh(gaussian_hull(rng, 16, 0.3))
    compounds = {}
    for i in range(n):
        pos = (rng.uniform(-2, 2), 2.5 + 2.2 * i, rng.uniform(-2, 2))
        rot = quat_to_matrix(random_quaternions(rng, 1))[0]
        if i % 6 == 5:
            b.add(SHAPE_SPHERE if i % 12 == 5 else SHAPE_BOX, (0.4,) if i % 12 == 5 else (0.7, 0.5, 0.6), 1.0, pos, rot)
            continue
        k = b.add(SHAPE_COMPOUND, (), 2.0, pos, rot, e=0.2)
        r3 = quat_to_matrix(random_quaternions(rng, 3))
        compounds[k] = [compound_child(SHAPE_BOX, (0.8, 0.4, 0.6), (0.0, 0.0, 0.0), ratio=0.4),
                        compound_child(SHAPE_SPHERE, (0.3,), tuple(rng.uniform(-0.6, 0.6, 3)), ratio=0.2),
                        compound_child(SHAPE_CYLINDER, (0.15, 0.8), tuple(rng.uniform(-0.6, 0.6, 3)), r3[0], 0.2),
                        compound_child(SHAPE_CONVEX, (), tuple(rng.uniform(-0.6, 0.6, 3)), r3[1], 0.2, mesh=hull)]
    sc = b.scene(name=f"compound_pile{n}")
    sc.compounds = compounds
    return sc
