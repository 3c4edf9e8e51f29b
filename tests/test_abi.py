"""Host-side checks of the C ABI (no GPU needed): the libraries load, export
every symbol include/pdos.h and include/pdos_ext.h declare, keep the
reference's struct layouts, and reject bad input before touching the device."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import golden_data as G
import pdos_amd
from pdos_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in ("pdos.h", "pdos_ext.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "return"):
                names.add(m.group(1))
    return names


@pytest.mark.parametrize("plugin", ["indirect", "direct"])
def test_library_exports_every_declared_symbol(plugin):
    L = pdos_amd.lib(plugin)
    declared = _declared_symbols()
    assert {"pdos", "freeData", "freeSol", "pdos_ex", "pdos_batch"} <= declared
    for name in declared:
        assert hasattr(L, name), name
    assert set(pdos_amd.EXPORTS) <= declared


def test_struct_layouts_match_reference_abi():
    # pdos.h:15-41 and cones.h:8-13 with idxint = long (LP64)
    assert C.sizeof(abi.Params) == 56
    assert C.sizeof(abi.Data) == 88
    assert C.sizeof(abi.Sol) == 56
    assert C.sizeof(abi.Cone) == 32
    assert abi.Data.x.offset < abi.Data.y.offset < abi.Data.s.offset   # x, y, s order
    assert abi.Sol.x.offset < abi.Sol.s.offset < abi.Sol.y.offset      # x, s, y order


def test_null_inputs_return_null():
    L = pdos_amd.lib("indirect")
    assert not L.pdos(None, None)
    d, k, keep = G.problem("fixture01").to_c()
    assert not L.pdos(C.byref(d), None)
    assert not L.pdos(None, C.byref(k))


def test_cone_mismatch_aborts_like_the_reference_assert():
    code = (
        "import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import ctypes as C, golden_data as G, pdos_amd\n"
        "p = G.problem('fixture01'); d, k, keep = p.to_c(); k.l = k.l + 1\n"
        "pdos_amd.lib('indirect').pdos(C.byref(d), C.byref(k))\n" % (ROOT, os.path.join(ROOT, "tests"))
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, timeout=120)
    assert r.returncode != 0
    assert b"do not match" in r.stderr


def test_freedata_and_freesol_null_safe():
    L = pdos_amd.lib("indirect")
    d = C.POINTER(abi.Data)()
    k = C.POINTER(abi.Cone)()
    L.freeData(C.byref(d), C.byref(k))
    s = C.POINTER(abi.Sol)()
    L.freeSol(C.byref(s))
