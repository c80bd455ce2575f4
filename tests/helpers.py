"""Shared test helpers: scenes at test sizes, comparisons."""
import numpy as np

from pe_amd import scenes


def small_rain(n=300, seed=7):
    return scenes.box_rain_scene(n, seed=seed, lo=(-5, 1, -5), hi=(5, 11, 5), ground=(40.0, 1.0, 40.0))


def mixed_small(n=200, seed=11):
    return scenes.mixed_world(n, seed)


def flat_state(st):
    return np.concatenate([np.asarray(x, np.float64).reshape(len(x), -1) for x in st], axis=1)


def pair_set(ab):
    return set(map(tuple, np.sort(np.asarray(ab), axis=1).tolist()))


def brute_force_pairs(mn, mx, kin):
    """All (i, j) with closed-interval AABB overlap on x, y, z, not both kinematic
    (BroadPhaseBase::validateCollisionPair, broad_phase_base.cpp:5-19)."""
    n = len(mn)
    ov = np.ones((n, n), bool)
    for k in range(3):
        ov &= ~((mn[:, None, k] > mx[None, :, k]) | (mn[None, :, k] > mx[:, None, k]))
    ov &= ~(kin[:, None].astype(bool) & kin[None, :].astype(bool))
    iu = np.triu_indices(n, 1)
    sel = ov[iu]
    return set(zip(iu[0][sel].tolist(), iu[1][sel].tolist()))


def valid_mask(pts):
    return pts[..., 0] < 1e29
