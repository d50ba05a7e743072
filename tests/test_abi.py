"""The C-ABI library loads and exports every symbol include/ceviche_amd.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, 'include', 'ceviche_amd.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    names = re.findall(r'^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(cv_[a-z0-9_]+)\s*\(', src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from ceviche_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from ceviche_amd import _lib
    declared = set(_declared())
    bound = set(_lib.exported_symbols())
    assert bound <= declared, bound - declared


def test_version_string():
    from ceviche_amd import _lib
    assert b'gfx950' in _lib.lib().cv_version()
