"""pe_stage_broadphase_noalloc (include/pe_amd_noalloc.h), the sweep-and-prune stage for callers that
cannot allocate: through the C ABI it returns exactly what pe_stage_broadphase returns (pairs in
order, the pair count, the next axis, the capacity bits), call after call on the same fixed working
memory -- a larger input, then a smaller one, then the larger again -- and refuses inputs beyond
its fixed size before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PE_ERR_INVALID = -1


def _records(sc, steps):
    import pe_amd
    w = pe_amd.World(sc)
    w.step(steps)
    w.integrate()
    w.sync()
    aabb = w.records()[0]
    axis = w.axis()
    w.close()
    return aabb, axis


def _both(st, aabb, axis, cap, stream):
    """The allocating entry point and the no-allocation one on the same input, on one stream."""
    from pe_amd._lib import check, lib
    from pe_amd.hip import DeviceArray
    n = aabb.shape[0]
    pa0, pb0, r0 = st.broadphase(aabb, axis, cap, stream=stream)
    pa1 = DeviceArray.full(cap, -1, np.int32)
    pb1 = DeviceArray.full(cap, -1, np.int32)
    r1 = DeviceArray.full(3, -1, np.int32)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    check(lib().pe_stage_broadphase_noalloc(ctypes.c_int32(n), p(aabb), ctypes.c_int32(axis), p(pa1), p(pb1),
                                            ctypes.c_int32(cap), p(r1), stream.handle))
    stream.synchronize()
    r0, r1 = r0.numpy(), r1.numpy()
    k = min(int(r0[0]), cap)
    return (r0, pa0.numpy()[:k], pb0.numpy()[:k]), (r1, pa1.numpy()[:k], pb1.numpy()[:k])


def test_noalloc_broadphase_equals_allocating_entry_point():
    from pe_amd import scenes
    from pe_amd.hip import DeviceArray, Stream
    from pe_amd.stage import Stage
    big, ax_big = _records(scenes.box_rain_scene(3000, seed=71, lo=(-12, 1, -12), hi=(12, 31, 12),
                                                 ground=(60.0, 1.0, 60.0)), 90)
    small, ax_small = _records(scenes.shape_mix_scene(400, seed=72, volume=(10, 6, 10), grid=8, ground_size=(12, 12)), 40)
    st = Stage(4000, 64000)
    s = Stream()
    seen = []
    for aabb, axis in ((big, ax_big), (small, ax_small), (big, ax_big)):
        t = DeviceArray.from_numpy(aabb)
        (r0, a0, b0), (r1, a1, b1) = _both(st, t, axis, 64000, s)
        print(f"[noalloc] n={len(aabb)} axis {axis}: result {r0.tolist()} vs {r1.tolist()}", flush=True)
        assert np.array_equal(r0, r1)
        assert r0[0] > 0 and r0[2] == 0
        assert np.array_equal(a0, a1) and np.array_equal(b0, b1)
        seen.append(int(r0[0]))
    assert seen[0] > seen[1]          # the second call's input is the smaller one
    # an output too small for the pairs: the same overflow report and the same first pairs
    t = DeviceArray.from_numpy(small)
    (r0, a0, b0), (r1, a1, b1) = _both(st, t, ax_small, 10, s)
    assert np.array_equal(r0, r1) and r0[2] & 1 and r0[0] > 10
    assert np.array_equal(a0, a1) and np.array_equal(b0, b1)
    st.close()
    s.close()


def test_noalloc_refuses_inputs_beyond_its_working_memory():
    from pe_amd._lib import lib
    # refused on the host before any device work (no buffers needed)
    rc = lib().pe_stage_broadphase_noalloc(ctypes.c_int32(65537), None, ctypes.c_int32(0), None, None,
                                           ctypes.c_int32(0), None, None)
    assert rc == PE_ERR_INVALID
    assert b"65536" in lib().pe_last_error()
