"""Parity of the plain single-world HIP path against the CPU oracle (FP32 restatement of the
reference, itself pinned bit-exact to the compiled reference in test_oracle_vs_ref.py).

Tolerances (written here, applied below):
  integrate / AABB / pair list / solve order .............. bit-exact
  contact points (normals, depths, positions) ............. 1e-5 relative (abs floor 1e-6)
  velocities after one solve .............................. 1e-5 relative (abs floor 1e-6)
  accumulated impulses after one solve .................... 1e-5 relative (floor 1e-6 x the step's largest)
  60-step trajectories .................................... position RMS <= 1e-3 m, max <= 1e-2 m,
                                                            orientation max <= 1e-3 rad
"""
import numpy as np
import pytest

from helpers import (brute_force_pairs, flat_state, impulse_error, mixed_small, normal_residual, pair_set,
                     reference_sample_scene, rotation_angle, small_rain, valid_mask)

pytestmark = pytest.mark.gpu

REL = 1e-5
ABS = 1e-6
IMP_REL = 1e-5
ROT_MAX = 1e-3


def _scenes():
    from pe_amd import scenes
    return {
        "tower": scenes.tower_scene(),
        "rain300": small_rain(300),
        "mixed200": mixed_small(200),
    }


@pytest.fixture(scope="module")
def all_scenes():
    return _scenes()


def _pair(name, all_scenes):
    import pe_amd
    from pyoracle import OracleWorld
    sc = all_scenes[name]
    return sc, pe_amd.World(sc), OracleWorld(sc, 32)


def _close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


@pytest.mark.parametrize("name", ["tower", "rain300", "mixed200"])
def test_stages_match_oracle(name, all_scenes):
    sc, g, o = _pair(name, all_scenes)
    print(f"[stages] scene={name} n={sc.n}", flush=True)
    worst = 0.0
    for step in range(8):
        g.integrate(); o.integrate()
        g.sync()
        # stage 0: bit-exact state and AABBs
        assert np.array_equal(flat_state(g.state()), flat_state(o.state()).astype(np.float32)), f"state step {step}"
        gmn, gmx = g.aabbs()
        omn, omx = o.aabbs()
        assert np.array_equal(gmn, omn.astype(np.float32)) and np.array_equal(gmx, omx.astype(np.float32))
        # stage 1: pair list identical in order; equal to brute force as a set
        npg = g.broadphase()
        npo = o.broadphase()
        gp, op = g.pairs(), o.pairs()
        assert npg == npo
        assert np.array_equal(gp, op), f"pair order step {step}"
        assert pair_set(gp) == brute_force_pairs(gmn, gmx, sc.kinematic)
        assert g.axis() == o.axis()
        # stage 2: contact manifolds
        g.narrowphase(); o.narrowphase()
        gn, gpts = g.contacts()
        on, _, opts = o.contacts()
        assert np.array_equal(gn, on), f"point counts step {step}"
        vm = np.arange(8)[None, :] < on[:, None]        # the points the solver uses
        assert _close(gpts[vm], opts[vm]), f"contact points step {step}"
        assert _close(gpts[..., 12], opts[..., 12])  # distances of every slot, stale ones included
        # stage 3: solve in colour order
        g.solve(); o.solve()
        g.sync()
        order, cs = g.solve_order()
        gcol = np.full(len(gp), -1)
        for c in range(len(cs) - 1):
            gcol[order[cs[c]:cs[c + 1]]] = c
        assert np.array_equal(gcol, o.color_of()), f"colour classes step {step}"
        assert len(cs) - 1 == o.num_colors()
        _, _, gv, gw = g.state()
        _, _, ov, ow = o.state()
        assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}"
        # solved impulses (FrictionContactConstraint's accumulated impulses), same colour order
        err, rows = impulse_error(g.impulses(), o.impulses(), on)
        worst = max(worst, err)
        assert err <= IMP_REL, f"impulses step {step}: max relative error {err:.3e} over {rows} rows"
        # keep the two in lock-step on the exact same state
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    print(f"[stages] {name}: largest relative impulse error {worst:.3e}", flush=True)


@pytest.mark.parametrize("name", ["tower", "rain300"])
def test_trajectory_60_steps(name, all_scenes):
    sc, g, o = _pair(name, all_scenes)
    print(f"[traj60] scene={name} n={sc.n}", flush=True)
    g.step(60)
    o.step(60)
    g.sync()
    pg, rg, _, _ = g.state()
    po, ro, _, _ = o.state()
    d = np.linalg.norm(pg - po, axis=1)
    rms = float(np.sqrt(np.mean(d ** 2)))
    ang = rotation_angle(rg, ro)
    print(f"[traj60] {name}: position rms {rms:.3e} max {d.max():.3e}, orientation max {ang.max():.3e} rad",
          flush=True)
    assert rms <= 1e-3 and d.max() <= 1e-2
    assert ang.max() <= ROT_MAX


def test_solve_order_is_a_valid_coloring(all_scenes):
    import pe_amd
    sc = all_scenes["rain300"]
    g = pe_amd.World(sc)
    g.step(30)
    g.sync()
    order, cs = g.solve_order()
    ab = g.pairs()
    kin = sc.kinematic.astype(bool)
    for c in range(len(cs) - 1):
        bodies = ab[order[cs[c]:cs[c + 1]]].ravel()
        dyn = bodies[~kin[bodies]]
        assert len(dyn) == len(np.unique(dyn)), f"colour {c} shares a dynamic body"


def test_fp32_residual_against_fp64_oracle():
    """The FP32 device solve's normal-contact residual (measured in FP64) beside the residual of the
    FP64 oracle's own solve from the same state: both carry the 10-iteration Gauss-Seidel's
    unconverged part; FP32 rounding must not add materially to it (bound written below)."""
    import pe_amd
    from pyoracle import OracleWorld
    sc = small_rain(300)
    assert np.all(sc.restitution == 0)
    kin = sc.kinematic.astype(bool)
    g, o = pe_amd.World(sc), OracleWorld(sc, 64)
    inv_mass, linv = o.mass_props()
    g.step(60)
    r32, r64, l32, l64, rows = [], [], [], [], 0
    for step in range(30):
        g.sync()
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
        g.integrate(); g.broadphase(); g.narrowphase(); g.solve()
        o.integrate(); o.broadphase(); o.narrowphase(); o.solve()
        g.sync()
        pos, rot, v, w = g.state()
        ab, (npts, pts), imp = g.pairs(), g.contacts(), g.impulses()
        a, b_, n_ = normal_residual(ab, npts, pts, imp, rot, v, w, inv_mass, linv, kin, sc.dt)
        opos, orot, ov, ow = o.state()
        onp, _, opts = o.contacts()
        c, d, _ = normal_residual(o.pairs(), onp, opts, o.impulses(), orot, ov, ow, inv_mass, linv, kin, sc.dt)
        r32.append(a); l32.append(b_); r64.append(c); l64.append(d); rows += n_
    R32, R64 = float(np.sqrt(np.mean(np.square(r32)))), float(np.sqrt(np.mean(np.square(r64))))
    L = float(np.sqrt(np.mean(np.square(l64))))
    print(f"[residual] {rows} normal rows over 30 steps: FP32 device RMS {R32:.4e}, FP64 oracle RMS {R64:.4e} "
          f"(N s; RMS normal impulse {L:.4e}, relative {R32 / L:.3e} vs {R64 / L:.3e})", flush=True)
    assert R32 <= 1.5 * R64 + 1e-4 * L


def test_reference_sample_scene_stages():
    """The reference's SampleScenes/bomb_crash_tower_smaller.json (restitution 0.5, angular damping
    0.8, a 70 m/s sphere into a ring of boxes) in lock step with the oracle for 150 steps, every
    stage compared on every step (tolerances above), the oracle restarted from the device state
    after each step."""
    import os
    import pe_amd
    from pyoracle import OracleWorld
    sc = reference_sample_scene(os.path.join(os.path.dirname(__file__), "golden", "bomb_crash_tower_smaller.json"))
    g, o = pe_amd.World(sc), OracleWorld(sc, 32)
    print(f"[sample scene] n={sc.n}", flush=True)
    contacts = 0
    for step in range(150):
        g.integrate(); o.integrate()
        g.sync()
        assert np.array_equal(flat_state(g.state()), flat_state(o.state()).astype(np.float32)), f"state step {step}"
        g.broadphase(); o.broadphase()
        gp, op = g.pairs(), o.pairs()
        assert np.array_equal(gp, op), f"pair order step {step}"
        g.narrowphase(); o.narrowphase()
        gn, gpts = g.contacts()
        on, _, opts = o.contacts()
        assert np.array_equal(gn, on), f"point counts step {step}"
        vm = np.arange(8)[None, :] < on[:, None]
        assert _close(gpts[vm], opts[vm]), f"contact points step {step}"
        contacts += int(on.sum())
        g.solve(); o.solve()
        g.sync()
        _, _, gv, gw = g.state()
        _, _, ov, ow = o.state()
        assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}"
        err, rows = impulse_error(g.impulses(), o.impulses(), on)
        assert err <= IMP_REL, f"impulses step {step}: max relative error {err:.3e} over {rows} rows"
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    assert contacts > 0


@pytest.mark.parametrize("n,mode", [(13000, 0), (6000, 1), (6000, 2)])
def test_large_world_colouring_matches_oracle(n, mode):
    """The solver's colouring in one workgroup (claims in LDS) and in rounds over the GPU: a world
    beyond the one-workgroup limit (12 288 bodies, rounds) and a smaller one both ways give colour
    classes equal to the oracle's (bit-exact), and the solved velocities stay within 1e-5, step after
    step in lock-step."""
    import pe_amd
    from pe_amd import scenes
    from pyoracle import OracleWorld
    sc = scenes.grid_world(n, seed=77)
    g, o = pe_amd.World(sc), OracleWorld(sc, 32)
    g.set_colouring(mode)
    print(f"[large colouring] n={sc.n}, colouring mode {mode}", flush=True)
    checked = 0
    for step in range(95):
        g.integrate(); o.integrate()
        g.broadphase(); o.broadphase()
        g.narrowphase(); o.narrowphase()
        g.solve(); o.solve()
        g.sync()
        if step >= 85:
            gp = g.pairs()
            assert np.array_equal(gp, o.pairs()), f"pairs step {step}"
            order, cs = g.solve_order()
            gcol = np.full(len(gp), -1)
            for c in range(len(cs) - 1):
                gcol[order[cs[c]:cs[c + 1]]] = c
            assert np.array_equal(gcol, o.color_of()), f"colour classes step {step}"
            _, _, gv, gw = g.state()
            _, _, ov, ow = o.state()
            assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}"
            checked += 1
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    rounds = g.counts()["colour_rounds"]
    print(f"[large colouring] pairs {len(g.pairs())}, colours {len(cs) - 1}, rounds {rounds}, steps checked {checked}",
          flush=True)
    assert len(cs) - 1 >= 3
    assert (rounds > 0) == (mode != 1)
