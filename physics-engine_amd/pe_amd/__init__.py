"""pe_amd -- MI355X (gfx950) rigid-body step behind the reference engine's World interface.

    from pe_amd import scenes, World
    w = World(scenes.tower_scene())
    w.step(60)

    w = World(scene_io.load_scene("bomb_crash_tower_smaller.json"))   # the reference's scene files
"""
from . import scenes  # noqa: F401
from . import scene_io  # noqa: F401
from ._lib import PEError, check, lib, LIB_PATH  # noqa: F401
from .world import World  # noqa: F401
from .batch import Batch  # noqa: F401
from .server import Server, ServerConfig  # noqa: F401
from .mesh import prepare_mesh  # noqa: F401
from .stage import Stage  # noqa: F401
