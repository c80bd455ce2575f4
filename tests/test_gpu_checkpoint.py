"""Checkpoint and resume through the reference's scene-file format (pe_amd.scene_io.save_checkpoint /
load_checkpoint): a world stepped N steps, saved, loaded into a fresh World and stepped M more
ends bit-identical to the same world stepped N + M steps without interruption -- boxes and spheres,
and the mixed shapes over a heightfield (convex meshes re-prepared by setMesh on load).  Also the
stream-taking step (pe_world_step_async) against pe_world_step."""
import numpy as np
import pytest

from helpers import flat_state, small_rain

pytestmark = pytest.mark.gpu


def _scenes():
    from pe_amd import scenes
    return {"rain300": small_rain(300, seed=12),
            "shapes": scenes.shape_mix_scene(250, seed=19, volume=(12, 8, 12), grid=16, ground_size=(16, 16))}


@pytest.mark.parametrize("name", ["rain300", "shapes"])
def test_checkpoint_resume_is_bit_identical(name, tmp_path):
    import pe_amd
    from pe_amd.scene_io import load_checkpoint, load_scene, save_checkpoint
    sc = _scenes()[name]
    a = pe_amd.World(sc)
    a.step(60)
    a.sync()
    path = str(tmp_path / f"{name}_ckpt.json")
    save_checkpoint(a, path)
    k, _ = a.cache()
    b, sb = load_checkpoint(path)
    assert sb.n == sc.n and b.axis() == a.axis() and b.cache()[0] == k
    assert np.array_equal(flat_state(b.state()), flat_state(a.state()))
    a.step(45)
    b.step(45)
    a.sync()
    b.sync()
    assert np.array_equal(flat_state(a.state()), flat_state(b.state())), name
    # the same file is a plain scene file too (the reference's loader ignores the extra keys)
    plain = load_scene(path)
    assert plain.n == sc.n
    # without the contact cache and axis the resumed run is not the same run: both are needed
    c = pe_amd.World(plain)
    c.step(45)
    c.sync()
    print(f"[checkpoint] {name}: {k} cached pairs, axis {a.axis()}; fresh-cache run differs: "
          f"{not np.array_equal(flat_state(c.state()), flat_state(a.state()))}", flush=True)
    a.close(); b.close(); c.close()


def test_step_async_on_a_caller_stream():
    import pe_amd
    from pe_amd.hip import Stream
    sc = small_rain(200, seed=4)
    a, b = pe_amd.World(sc), pe_amd.World(sc)
    s = Stream()
    a.step(10)
    b.step(10)
    b.step_async(25, s.handle.value)
    a.step(25)
    # b's getters wait for the work enqueued on the caller's stream
    assert np.array_equal(flat_state(a.state()), flat_state(b.state()))
    b.step(5)
    b.step_async(5, 0)          # the null stream
    a.step(10)
    assert np.array_equal(flat_state(a.state()), flat_state(b.state()))
    a.close(); b.close(); s.close()
