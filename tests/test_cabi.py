"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/clothsim.h, its structs have the header's layout, and the
host-only entry points (version, default parameters, material table) behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import clothsimulator_amd as cs
import oracle
from clothsimulator_amd import _cabi
from clothsimulator_amd._cabi import MaterialIn, Params, material_struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "clothsim.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(cs_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    L = cs.load_library()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(cs.EXPORTED) == names


def test_struct_layout_matches_header():
    # sizes computed from the header's field lists (4-byte ints/floats, 8-byte doubles)
    assert C.sizeof(_cabi.ProximityRec) == 52
    assert C.sizeof(_cabi.ImpactRec) == 88
    assert _cabi.ImpactRec.t.offset == 24
    text = open(HEADER).read()
    for name in ("cs_params", "cs_mesh_in", "cs_material_in", "cs_stats", "cs_proximity_rec", "cs_impact_rec"):
        assert re.search(r"typedef struct\s*(%s)?\s*\{" % name, text) or ("} %s;" % name) in text, name


def test_version_and_default_params():
    L = cs.load_library()
    L.cs_version.restype = C.c_char_p
    assert L.cs_version().decode().startswith("clothsim")
    p = Params()
    L.cs_default_params(C.byref(p))
    ref = _cabi.default_params()
    for f, _ in Params._fields_:
        a, b = getattr(p, f), getattr(ref, f)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), f
        elif f not in ("dt",):
            assert a == pytest.approx(b), f


@pytest.mark.parametrize("kind", ["constant", "anisotropic"])
def test_material_table_bitexact_with_oracle(kind):
    if kind == "constant":
        m = material_struct(density=0.2, stretch=3e3, bend=2e-6, thicken=1.5)
    else:
        rows = [[900.0 + 37 * i, 11.0 * i, 1400.0 - 55 * i, 310.0 + 3 * i] for i in range(6)]
        bend = [[1e-6 * (1 + i + j) for j in range(5)] for i in range(3)]
        m = material_struct(density=0.25, stretching=rows, bending=bend, stretching_mult=1.3, bending_mult=0.7)
    t1, d1, b1 = cs.material_table(m)
    t2, d2, b2 = oracle.material_table(m)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert np.float32(d1) == np.float32(d2)
    assert np.array_equal(np.asarray(b1, np.float32).view(np.uint32), np.asarray(b2, np.float32).view(np.uint32))


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = cs.load_library()
    p = _cabi.default_params()
    ctx = C.c_void_p()
    st = L.cs_create(C.byref(p), C.byref(ctx))
    assert st != 0 and not ctx.value
