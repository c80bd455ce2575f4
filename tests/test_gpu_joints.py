"""Joint constraints and ignore-collision on the device (pe_world_add_joints,
pe_world_add_ignore_pairs, pe_scene_desc.ignore_collision) against the oracle, whose joints and
ignore rules are pinned bit for bit to the compiled reference (test_oracle_vs_ref.py).

* Stage by stage, in lock-step on the same state: the state after integrate and the pair list
  (ignore pairs and ignore-collision bodies left out) bit-identical, the contact manifolds and the
  solved velocities within 1e-5 relative of the oracle sweeping the same colour order (contacts,
  then joints colour by colour, sequential_impulse_solver.cpp:60-81).
* Free running on the device: the constraints hold as the mathematics fixes them (anchors
  coincide, the slider stays on its rail, the limited hinge stops at its limit)."""
import numpy as np
import pytest

from helpers import flat_state, pair_set

pytestmark = pytest.mark.gpu

REL = 1e-5
ABS = 1e-6


def _close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


def _ignored(sc):
    out = {(min(a, b), max(a, b)) for a, b in sc.ignore_pairs}
    if sc.ignore_collision is not None:
        flagged = set(np.nonzero(sc.ignore_collision)[0].tolist())
        out |= {(i, j) for i in range(sc.n) for j in range(i + 1, sc.n) if i in flagged or j in flagged}
    return out


def _lockstep(sc, steps):
    import pe_amd
    from pyoracle import OracleWorld
    g, o = pe_amd.World(sc), OracleWorld(sc, 32, color_order=True)
    ign = _ignored(sc)
    worst, contacts = 0.0, 0
    for step in range(steps):
        g.integrate(); o.integrate()
        g.sync()
        assert np.array_equal(flat_state(g.state()), flat_state(o.state()).astype(np.float32)), f"state step {step}"
        assert g.broadphase() == o.broadphase()
        gp = g.pairs()
        assert np.array_equal(gp, o.pairs()), f"pair order step {step}"
        assert not (pair_set(gp) & ign), f"an ignored pair was reported, step {step}"
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
        d = np.abs(np.concatenate([gv - ov, gw - ow])) / (ABS + np.abs(np.concatenate([ov, ow])))
        worst = max(worst, float(d.max()))
        assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}: max relative deviation {d.max():.3e}"
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    g.close()
    return worst, contacts


def test_joint_demo_stages_match_oracle():
    from pe_amd import scenes
    sc = scenes.joint_demo_scene()
    print(f"[joints] demo n={sc.n} joints={len(sc.joints)}", flush=True)
    worst, contacts = _lockstep(sc, 150)
    print(f"[joints] demo: 150 steps, {contacts} contact points, largest relative velocity deviation {worst:.3e}",
          flush=True)


def test_joint_chains_match_oracle():
    """384 ball joints in 48 chains: a joint colour holds a joint of every chain, swept by many
    threads at once."""
    from pe_amd import scenes
    sc = scenes.joint_chain_scene()
    print(f"[joints] chains n={sc.n} joints={len(sc.joints)}", flush=True)
    worst, contacts = _lockstep(sc, 60)
    print(f"[joints] chains: 60 steps, {contacts} contact points, largest relative velocity deviation {worst:.3e}",
          flush=True)


def test_ignore_collision_flags_and_pairs():
    """A falling cluster where some bodies ignore every collision (setIgnoreCollision) and random
    pairs ignore each other (addIgnoreCollisionId): the pair list equals the oracle's, in order."""
    from helpers import small_rain
    sc = small_rain(240, seed=11)
    rng = np.random.default_rng(4)
    flags = np.zeros(sc.n, np.int32)
    flags[rng.choice(np.arange(1, sc.n), 12, replace=False)] = 1
    sc.ignore_collision = flags
    pairs = rng.integers(1, sc.n, size=(400, 2))
    sc.ignore_pairs = [(int(a), int(b)) for a, b in pairs if a != b]
    print(f"[ignore] n={sc.n} flagged={int(flags.sum())} ignore pairs={len(sc.ignore_pairs)}", flush=True)
    _lockstep(sc, 40)


def test_joint_constraints_hold_on_device():
    import pe_amd
    from pe_amd import scenes
    from helpers import ortho_units as _ortho
    sc = scenes.joint_demo_scene()
    g = pe_amd.World(sc)
    h = sc.joints[3]
    worst_gap, worst_off, angles = 0.0, 0.0, []
    for step in range(300):
        g.step(1)
        pos, rot, _, _ = [np.asarray(x, np.float64) for x in g.state()]
        for j in sc.joints[:3]:
            pa = rot[j.body_a] @ np.array(j.anchor_a) + pos[j.body_a]
            pb = rot[j.body_b] @ np.array(j.anchor_b) + pos[j.body_b]
            worst_gap = max(worst_gap, float(np.linalg.norm(pa - pb)))
        rail, sl = sc.joints[5].body_a, sc.joints[5].body_b
        axis = rot[rail] @ np.array([0.0, 1.0, 0.0])
        d = pos[sl] - pos[rail]
        worst_off = max(worst_off, float(np.linalg.norm(d - (d @ axis) * axis)))
        ta = rot[h.body_a] @ _ortho(h.axis_a)[0]
        tb = rot[h.body_b] @ _ortho(h.axis_b)[0]
        wa = rot[h.body_a] @ np.array(h.axis_a)
        angles.append(float(np.arctan2(np.cross(ta, tb) @ wa, ta @ tb)))
    g.close()
    print(f"[joints] device: anchor gap {worst_gap:.2e} m, slider off-axis {worst_off:.2e} m, hinge angle "
          f"[{min(angles):.3f}, {max(angles):.3f}] rad", flush=True)
    assert worst_gap < 2e-2 and worst_off < 2e-2
    lim = np.pi / 3
    assert max(abs(a) for a in angles) > 0.9 * lim
    assert all(-lim - 0.1 <= a <= lim + 0.1 for a in angles)
