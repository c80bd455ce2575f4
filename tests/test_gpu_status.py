"""Sleeping, lifetimes, removal and collision reports on the device (pe_world_set_sleep,
pe_scene_desc.lifetime, pe_world_remove_body, pe_world_get_body_status, pe_world_set_collision_reports,
pe_world_get_collision_events) against the oracle, whose rules are pinned bit for bit to the compiled
reference (test_oracle_vs_ref.py::test_sleep_lifetime_and_callbacks_bit_exact).

Lock-step on the same state every step: status bits, neighbour counts, sleep times, the state after
integrate, the pair list and the collision reports bit-identical; the solved velocities within 1e-5
relative of the oracle sweeping the same colour order."""
import numpy as np
import pytest

from helpers import flat_state, small_rain

pytestmark = pytest.mark.gpu

REL = 1e-5
ABS = 1e-6


def _close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


def _scene():
    sc = small_rain(120, seed=5)
    sc.sleep = (0.01, 0.01, 1.0)
    lt = np.full(sc.n, 1e30)
    lt[-1] = 2.0
    lt[17] = 3.5
    sc.lifetime = lt
    return sc


def test_sleep_lifetime_removal_and_reports_match_oracle():
    import pe_amd
    from pyoracle import OracleWorld
    sc = _scene()
    g, o = pe_amd.World(sc), OracleWorld(sc, 32, color_order=True)
    reporting = [3, 7, 20]
    g.set_collision_reports(reporting)
    for b in reporting:
        o.set_report(b)
    print(f"[status] n={sc.n} sleep={sc.sleep} reporting={reporting}", flush=True)
    slept = events = 0
    prev = np.zeros(sc.n, np.int32)
    for step in range(420):
        if step == 200:
            g.remove_body(40)
            o.remove_body(40)
        g.integrate(); o.integrate()
        g.sync()
        gs, os_ = g.body_status(), o.status()
        assert np.array_equal(gs[0], os_[0]), f"status bits after integrate, step {step}"
        assert np.array_equal(gs[3], os_[3].astype(np.float32)), f"sleep times, step {step}"
        assert np.array_equal(flat_state(g.state()), flat_state(o.state()).astype(np.float32)), f"state step {step}"
        assert g.broadphase() == o.broadphase()
        assert np.array_equal(g.pairs(), o.pairs()), f"pair order step {step}"
        g.narrowphase(); o.narrowphase()
        gs, os_ = g.body_status(), o.status()
        assert np.array_equal(gs[1], os_[1]) and np.array_equal(gs[2], os_[2]), f"neighbour counts step {step}"
        gab, gev = g.collision_events()
        oab, oev = o.events()
        assert np.array_equal(gab, oab), f"reported pairs step {step}"
        assert np.array_equal(gev, oev.astype(np.float32)), f"reports step {step}"
        events += len(gab)
        g.solve(); o.solve()
        g.sync()
        _, _, gv, gw = g.state()
        _, _, ov, ow = o.state()
        assert _close(gv, ov) and _close(gw, ow), f"velocities step {step}"
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
        st = g.body_status()[0] & 1
        slept += int(((st == 1) & (prev == 0)).sum())
        prev = st
    flags = g.body_status()[0]
    assert flags[-1] & 2 and flags[17] & 2 and flags[40] & 2
    assert slept > 0 and events > 0
    print(f"[status] {slept} bodies fell asleep, {events} collision reports, removed "
          f"{np.nonzero(flags & 2)[0].tolist()}", flush=True)
    g.close()


def test_callbacks_are_called_with_the_reports():
    import pe_amd
    sc = _scene()
    g = pe_amd.World(sc)
    calls = []
    g.add_collision_callback(3, lambda *a: calls.append(a))
    g.step(90)
    assert calls, "body 3 touched something in 90 steps"
    me, other, pos, nor, vel = calls[-1]
    assert me == 3 and other != 3 and abs(float(np.linalg.norm(nor)) - 1.0) < 1e-5
    g.close()
