"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU needed:
only symbol lookup, no compute calls)."""
import ctypes
import glob
import os
import re

from pe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(pe_\w+)\s*\(", text, flags=re.M))
    return names


def test_headers_list_matches_binding():
    assert declared_symbols() == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_last_error_is_callable_without_gpu():
    lib = _lib.lib()
    assert isinstance(lib.pe_last_error(), bytes)
