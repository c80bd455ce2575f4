"""Raycasts on the device (pe_world_raycast) against the oracle's DefaultRaycastSolver restatement,
which is pinned bit for bit to the compiled reference (test_oracle_vs_ref.py::test_raycast_bit_exact).

The device world and the oracle step in lock-step for 30 steps (the poses and AABBs after integrate
are bit-identical), then a few hundred rays go to both: hit flag, body, distance, hit point and
normal bit-identical -- boxes, spheres, cylinders, a convex hull and the heightfield, rays with
ignore sets, rays from inside bodies, and queries the reference ends early at a mesh body whose
AABB the ray misses."""
import numpy as np
import pytest

from helpers import small_rain

pytestmark = pytest.mark.gpu


def _lockstep(sc, steps):
    import pe_amd
    from pyoracle import OracleWorld
    g, o = pe_amd.World(sc), OracleWorld(sc, 32, color_order=True)
    for _ in range(steps):
        g.integrate(); o.integrate()
        g.broadphase(); o.broadphase()
        g.narrowphase(); o.narrowphase()
        g.solve(); o.solve()
        g.sync()
        o.set_state(*[np.asarray(x, np.float64) for x in g.state()])
    return g, o


def _rays(sc, pos, rng, k, aim=None):
    lo, hi = pos.min(0) - 1, pos.max(0) + 1
    out = []
    for i in range(k):
        if aim is not None and i % 2 == 0:
            s = pos[aim] + rng.uniform(-0.4, 0.4, 3) + np.array([0.0, 8.0, 0.0])
            d = np.array([0.0, -1.0, 0.0]) + rng.normal(size=3) * 0.01
            length, ig = 20.0, []
        else:
            s = rng.uniform(lo, hi)
            d = rng.normal(size=3) if i % 3 else np.array([rng.normal() * 0.2, -1.0, rng.normal() * 0.2])
            if i % 7 == 0:
                d = np.array([0.0, -1.0, 0.0])
            length = float(rng.uniform(0.5, 40.0))
            ig = rng.choice(sc.n, 3, replace=False).tolist() if i % 5 == 0 else []
        d = d / np.linalg.norm(d)
        out.append((s.astype(np.float32), d.astype(np.float32), np.float32(length), ig))
    return out


@pytest.mark.parametrize("name", ["rain", "hull", "mix"])
def test_raycast_matches_oracle(name):
    from pe_amd import scenes
    mix = scenes.shape_mix_scene(160, seed=23, volume=(10, 6, 10), grid=12, ground_size=(14, 14))
    conv = np.nonzero(mix.shape_type == scenes.SHAPE_CONVEX)[0]
    keep = sorted([i for i in range(mix.n) if mix.shape_type[i] != scenes.SHAPE_CONVEX] + [int(conv[0])])
    sc = {"rain": small_rain(150, seed=2), "hull": mix.subset(keep), "mix": mix}[name]
    print(f"[raycast] {name}: n={sc.n}", flush=True)
    g, o = _lockstep(sc, 30)
    pos = np.asarray(g.state()[0], np.float64)
    rays = _rays(sc, pos, np.random.default_rng(7), 300, aim=keep.index(int(conv[0])) if name == "hull" else None)
    hits = partial = mesh = 0
    for k, (s, d, length, ig) in enumerate(rays):
        gh, gb, gd, gp, gn = g.raycast([s], [d], [length], ig)
        oh, ob, od, op, on = o.raycast(s.astype(np.float64), d.astype(np.float64), float(length), ig)
        assert bool(gh[0]) == oh and int(gb[0]) == ob, f"ray {k}: device {(gh[0], gb[0])} oracle {(oh, ob)}"
        assert gd[0] == np.float32(od), f"ray {k} distance"
        assert np.array_equal(gp[0], op.astype(np.float32)) and np.array_equal(gn[0], on.astype(np.float32)), f"ray {k}"
        hits += oh
        partial += (not oh) and ob >= 0
        mesh += oh and sc.shape_type[ob] >= scenes.SHAPE_CONVEX
    # the same rays in one call (one workgroup per ray)
    st = np.stack([r[0] for r in rays if not r[3]])
    dr = np.stack([r[1] for r in rays if not r[3]])
    ln = np.array([r[2] for r in rays if not r[3]], np.float32)
    bh, bb, bd, _, _ = g.raycast(st, dr, ln)
    for i, (s, d, length) in enumerate(zip(st, dr, ln)):
        oh, ob, od, _, _ = o.raycast(s.astype(np.float64), d.astype(np.float64), float(length))
        assert bool(bh[i]) == oh and int(bb[i]) == ob and bd[i] == np.float32(od), f"batched ray {i}"
    print(f"[raycast] {name}: {hits} hits ({mesh} on meshes), {partial} queries ended early with a partial "
          f"result, of {len(rays)}", flush=True)
    g.close()
