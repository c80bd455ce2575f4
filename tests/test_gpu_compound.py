"""Compound bodies on the device (PE_SHAPE_COMPOUND, pe_compound_child) against the oracle, whose
CompoundShape and CompoundCompoundCollisionAlgorithm restatement is pinned bit for bit to the
compiled reference (test_oracle_vs_ref.py::test_compound_bit_exact).

Lock-step on the same state every step, the oracle sweeping the device's colour order: the state and
the AABBs after integrate and the pair list bit-identical; contact counts equal, contact points and
the solved velocities within 1e-5 relative -- the reference's compound demo (stacked boxes with
cylinder arms) and a pile of compounds with box, sphere, cylinder and convex-hull children over a
heightfield.  Raycasts down the standing demo stack equal the oracle's bit for bit."""
import numpy as np
import pytest

from helpers import flat_state

pytestmark = pytest.mark.gpu

REL = 1e-5
ABS = 1e-6


def _close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


@pytest.mark.parametrize("name,steps", [("demo", 240), ("pile", 160)])
def test_compounds_match_oracle(name, steps):
    import pe_amd
    from pe_amd import scenes
    from pyoracle import OracleWorld
    sc = scenes.compound_demo_scene() if name == "demo" else scenes.compound_pile_scene()
    g, o = pe_amd.World(sc), OracleWorld(sc, 32, color_order=True)
    comp = np.array([i in sc.compounds for i in range(sc.n)])
    print(f"[compound] {name}: n={sc.n}, {int(comp.sum())} compounds", flush=True)
    points = on_compounds = ray_hits = 0
    worst = 0.0
    for step in range(steps):
        g.integrate(); o.integrate()
        g.sync()
        assert np.array_equal(flat_state(g.state()), flat_state(o.state()).astype(np.float32)), f"state step {step}"
        gmn, gmx = g.aabbs()
        omn, omx = o.aabbs()
        assert np.array_equal(gmn, omn.astype(np.float32)) and np.array_equal(gmx, omx.astype(np.float32)), \
            f"AABBs step {step}"
        if step == 0 and name == "demo":
            xs = [(x, z) for x in np.linspace(-0.3, 0.3, 7) for z in (-1.05, -1.0, -0.9, 0.05)]
            st = np.array([(x, 80.0, z) for x, z in xs], np.float32)
            dr = np.tile(np.array([0.0, -1.0, 0.0], np.float32), (len(xs), 1))
            gh, gb, gd, gp, gn = g.raycast(st, dr, np.full(len(xs), 100.0, np.float32))
            for i, s in enumerate(st):
                oh, ob, od, op, on = o.raycast(s.astype(np.float64), (0.0, -1.0, 0.0), 100.0)
                assert bool(gh[i]) == oh and int(gb[i]) == ob and gd[i] == np.float32(od), f"ray {i}"
                assert np.array_equal(gp[i], op.astype(np.float32)) and np.array_equal(gn[i], on.astype(np.float32))
                ray_hits += oh and comp[ob]
        assert g.broadphase() == o.broadphase()
        gp_ = g.pairs()
        assert np.array_equal(gp_, o.pairs()), f"pair order step {step}"
        g.narrowphase(); o.narrowphase()
        gn_, gpts = g.contacts()
        on_, _, opts = o.contacts()
        assert np.array_equal(gn_, on_), f"point counts step {step}"
        vm = np.arange(8)[None, :] < on_[:, None]
        assert _close(gpts[vm], opts[vm]), f"contact points step {step}"
        points += int(on_.sum())
        if len(gp_):
            on_compounds += int(on_[comp[gp_[:, 0]] | comp[gp_[:, 1]]].sum())
        g.solve(); o.solve()
        g.sync()
        _, _, gv, gw = g.state()
        _, _, ov, ow = o.state()
        d = np.abs(np.concatenate([gv - ov, gw - ow])) / (ABS + np.abs(np.concatenate([ov, ow])))
        worst = max(worst, float(d.max()))
        assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}: {d.max():.3e}"
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    print(f"[compound] {name}: {steps} steps, {points} contact points ({on_compounds} on compounds), largest "
          f"relative velocity deviation {worst:.3e}, {ray_hits} ray hits on compounds", flush=True)
    assert on_compounds > 100
    if name == "demo":
        assert ray_hits > 5
    g.close()
