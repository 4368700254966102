"""The C-ABI library: loads, exports every entry point of include/*.h, and
validates case descriptors on the host (no GPU work is issued here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from hyperflow_amd import native
from hyperflow_amd import case as K
from hyperflow_amd import uns

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(name="hyperflow.h"):
    text = open(os.path.join(ROOT, "include", name)).read()
    return sorted(set(re.findall(r"HF_API\s+[\w\s\*]+?\b(hf_\w+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = native.lib()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(native.EXPORTS)
    learned = header_functions("hyperflow_learned.h")
    for n in learned:
        assert hasattr(lib, n), n
    assert set(learned) == set(native.EXPORTS_LEARNED)
    unstructured = header_functions("hyperflow_uns.h")
    for n in unstructured:
        assert hasattr(lib, n), n
    assert set(unstructured) == set(uns.EXPORTS_UNS)
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["hyperflow.h", "hyperflow_learned.h", "hyperflow_uns.h"]


def test_uns_desc_layout():
    assert C.sizeof(uns.hf_uns_desc) == 2 * 4 + 2 * 8 + 2 * 8 + 4 * 8


@pytest.mark.parametrize("over,msg", [
    ({"ncells": 0}, "no cells"),
    ({"cfl": 0.0}, "cfl"),
    ({"gamma": float("nan")}, "gamma"),
    ({"code": 5}, "boundary code 5"),
    ({"nan_vertex": True}, "non-finite vertex"),
])
def test_invalid_uns_descriptors_are_rejected(over, msg):
    q, c = uns.box(3, 2)
    c = np.ascontiguousarray(c, np.int32)
    if "code" in over:
        c[1, 2] = over.pop("code")
    v = np.ascontiguousarray(q, np.float64)
    if over.pop("nan_vertex", False):
        v[2, 1, 0] = np.nan
    d = uns.hf_uns_desc()
    d.ncells, d.gamma, d.cfl = len(q), 1.4, 0.5
    d.vertices = v.ctypes.data_as(C.POINTER(C.c_double))
    d.edge_types = c.ctypes.data_as(C.POINTER(C.c_int32))
    for k, val in over.items():
        setattr(d, k, val)
    h = C.c_void_p()
    rc = uns._lib().hf_uns_create(C.byref(d), C.byref(h))
    assert rc == -1
    assert msg in native.lib().hf_last_error().decode()
    assert not h.value


def test_net_desc_layout_and_param_count():
    assert C.sizeof(native.hf_net_desc) == 16 * 4 + 8
    assert native.Net.param_count() == 9 * 4 * 32 + 32 + 3 * (9 * 32 * 32 + 32) + 9 * 32 * 4 + 4


def test_struct_layout_matches_header():
    assert C.sizeof(native.hf_block_desc) == 4 * 8 + 2 * 4 + 4 * 4 + 4 * 4 + 16 * 8 + 2 * 4
    assert C.sizeof(native.hf_case_desc) == 4 * 4 + 2 * 8 + 4 * 4 + 8


def test_version_string():
    assert "gfx950" in native.version()


def _desc(**over):
    case = K.quad_case(8, 8, 0, scheme="Godunov")
    arr = native.block_descs([case])
    d = native.hf_case_desc()
    d.scheme, d.limiter, d.precision, d.dt_mode = 0, 0, 0, 0
    d.gamma, d.cfl, d.nblocks, d.tile_rows, d.tail_rows, d.fix_rows = 1.4, 0.5, 1, 0, 0, 0
    d.blocks = C.cast(arr, C.POINTER(native.hf_block_desc))
    for k, v in over.items():
        if k.startswith("block_"):
            setattr(arr[0], k[6:], v)
        else:
            setattr(d, k, v)
    return d, arr


@pytest.mark.parametrize("over,msg", [
    ({"scheme": 7}, "scheme"),
    ({"scheme": 1, "limiter": 9}, "limiter"),
    ({"precision": 3}, "precision"),
    ({"cfl": -1.0}, "cfl"),
    ({"nblocks": 0}, "no blocks"),
    ({"block_nx": 0}, "nx"),
    ({"block_x_right": -1.0}, "extents"),
    ({"block_problem": 1}, "problems"),
    ({"tail_rows": -1}, "tail_rows"),
    ({"fix_rows": -3}, "fix_rows"),
])
def test_invalid_descriptors_are_rejected(over, msg):
    d, arr = _desc(**over)
    h = C.c_void_p()
    rc = native.lib().hf_create(C.byref(d), C.byref(h))
    assert rc == -1
    assert msg in native.lib().hf_last_error().decode()
    assert not h.value


def test_internal_neighbour_validation():
    case = K.step_channel(8)
    arr = native.block_descs([case])
    arr[1].ny = 9     # shared left/right edge no longer matches
    d = native.hf_case_desc()
    d.scheme, d.limiter, d.gamma, d.cfl, d.nblocks = 1, 0, 1.4, 0.5, len(arr)
    d.blocks = C.cast(arr, C.POINTER(native.hf_block_desc))
    h = C.c_void_p()
    assert native.lib().hf_create(C.byref(d), C.byref(h)) == -1
    assert "Internal" in native.lib().hf_last_error().decode()
