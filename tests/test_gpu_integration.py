"""The reference-side adapter (physics-engine_amd/integration: AmdBroadPhase, AmdNarrowPhase,
AmdSequentialImpulseSolver) inside the reference engine's own World, compiled from the
reference's sources (oracle/_ref/libperef32_amd.so): the stages swapped in through
World::setBroadPhase / setNarrowPhase / setConstraintSolver, World::step called by the reference.

* All three swapped: the reference's World then steps bit for bit like this library's own World
  (pe_amd.World), every step -- the reference's FP32 integrate is the device's (both pinned to the
  oracle), the stages are the same kernels -- on the reference's sample scene (restitution,
  damping, a 70 m/s impact) and on a mixed-shape scene (hull and heightfield meshes handed over as
  the reference's shapes hold them after setMesh).
* One stage swapped at a time, against the reference's own stage on the same state: the pair set
  (also past the adapter's first output capacity), the contact manifolds (1e-5 relative), the
  solved velocities (1e-5 relative; against the oracle's colour-order solve, since the adapter's
  solver sweeps in colour order)."""
import os

import numpy as np
import pytest

from helpers import flat_state, pair_set, reference_sample_scene

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

REL = 1e-5
ABS = 1e-6
HERE = os.path.dirname(__file__)


def _close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


def _need_adapter():
    from pyoracle import ref_amd_available
    if not ref_amd_available():
        pytest.skip("oracle/_ref/libperef32_amd.so not built (reference checkout absent at build time)")


def _f32(st):
    return flat_state(st).astype(np.float32)


def _scenes():
    from pe_amd import scenes
    return {"sample": reference_sample_scene(os.path.join(HERE, "golden", "bomb_crash_tower_smaller.json")),
            "shapes": scenes.shape_mix_scene(160, seed=23, volume=(10, 6, 10), grid=12, ground_size=(14, 14))}


@pytest.mark.parametrize("name,steps", [("sample", 150), ("shapes", 100)])
def test_reference_world_with_device_stages_steps_like_device_world(name, steps):
    _need_adapter()
    import pe_amd
    from pyoracle import RefWorld
    sc = _scenes()[name]
    r = RefWorld(sc, 32, threads=4, amd_stages=7)
    w = pe_amd.World(sc)
    print(f"[adapter] {name}: n={sc.n}", flush=True)
    for step in range(steps):
        r.step_checked(1)
        w.step(1)
        w.sync()
        assert np.array_equal(_f32(r.state()), flat_state(w.state())), f"{name} step {step}"
    assert r.axis() == w.axis()
    print(f"[adapter] {name}: {steps} steps bit-identical; last pairs {w.counts()['pairs']}, "
          f"points {w.counts()['points']}", flush=True)
    w.close()


def test_broadphase_alone_matches_reference():
    """AmdBroadPhase in the reference's World against BroadPhaseSweepAndPrune, same state each step;
    a 60-sphere cluster (1,770 pairs) grows the adapter's output past its first capacity."""
    _need_adapter()
    from pe_amd.scenes import SHAPE_BOX, SHAPE_SPHERE, SceneBuilder
    from pyoracle import RefWorld
    sc = _scenes()["sample"]
    rng = np.random.default_rng(3)
    b = SceneBuilder()
    b.add(SHAPE_BOX, (40.0, 1.0, 40.0), 1.0, (0.0, -0.5, 0.0), kin=1)
    for _ in range(60):
        b.add(SHAPE_SPHERE, (0.6,), 100.0, tuple(rng.uniform(-0.2, 0.2, 3) + (0, 5, 0)))
    cluster = b.scene(name="cluster60")
    for scene, steps in ((sc, 60), (cluster, 3)):
        ref, amd = RefWorld(scene, 32, threads=4), RefWorld(scene, 32, threads=4, amd_stages=1)
        for step in range(steps):
            ref.integrate(); amd.integrate()
            assert ref.broadphase() == amd.broadphase()
            assert pair_set(ref.pairs()) == pair_set(amd.pairs()), f"pair set step {step}"
            assert ref.axis() == amd.axis()
            ref.narrowphase(); ref.solve()
            amd.set_pairs(ref.pairs())   # the reference's order (its std::sort ties) for the rest
            amd.narrowphase(); amd.solve()
            amd.set_state(*ref.state())
        print(f"[adapter broadphase] {scene.name}: last pairs {len(ref.pairs())}", flush=True)


def test_narrowphase_alone_matches_reference():
    _need_adapter()
    from pyoracle import RefWorld
    for name in ("sample", "shapes"):
        sc = _scenes()[name]
        ref, amd = RefWorld(sc, 32, threads=4), RefWorld(sc, 32, threads=4, amd_stages=2)
        seen = 0
        for step in range(90):
            ref.integrate(); amd.integrate()
            ref.broadphase()
            amd.set_pairs(ref.pairs())
            ref.narrowphase(); amd.narrowphase()
            rn, rs, rp = ref.contacts()
            an, as_, ap = amd.contacts()
            assert np.array_equal(rn, an), f"{name} point counts step {step}"
            assert np.array_equal(rs, as_), f"{name} swap flags step {step}"
            vm = np.arange(8)[None, :] < rn[:, None]
            assert _close(ap[vm], rp[vm]), f"{name} contact points step {step}"
            assert _close(ap[..., 12], rp[..., 12]), f"{name} slot distances step {step}"
            seen += int(rn.sum())
            ref.solve()
            amd.set_state(*ref.state())
        print(f"[adapter narrowphase] {name}: {seen} contact points compared", flush=True)
        assert seen > 0


def test_solver_alone_matches_colour_order_oracle():
    _need_adapter()
    from pyoracle import OracleWorld, RefWorld
    sc = _scenes()["sample"]
    amd = RefWorld(sc, 32, threads=4, amd_stages=4)
    o = OracleWorld(sc, 32, color_order=True)
    worst = 0.0
    for step in range(150):
        amd.integrate(); o.integrate()
        amd.broadphase()
        o.set_pairs(amd.pairs())
        amd.narrowphase(); o.narrowphase()
        amd.solve(); o.solve()
        _, _, av, aw = amd.state()
        _, _, ov, ow = o.state()
        assert _close(av, ov) and _close(aw, ow), f"velocities step {step}"
        d = np.abs(np.concatenate([av - ov, aw - ow])) / (ABS + np.abs(np.concatenate([ov, ow])))
        worst = max(worst, float(d.max()))
        o.set_state(*amd.state())
    print(f"[adapter solver] largest relative velocity deviation {worst:.3e}", flush=True)
