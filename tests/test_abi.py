"""The C ABI library on the host (no GPU needed): it loads, exports every
entry point include/xpbdfem.h declares, its host utilities agree with the
oracle, and bad arguments come back as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
import xpbdfem as xf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xpbdfem.h")


def declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"XF_API\s+[\w\s\*]+?\b(xf_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = xf.lib()
    names = declared_entry_points()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    assert sorted(xf.exported_symbols()) == names


def test_version_string():
    assert xf.lib().xf_version().startswith(b"xpbdfem-gfx950")


def test_mesh_generator_matches_oracle():
    for dims, mirrored in (((12, 4, 4), False), ((3, 5, 2), True)):
        a = xf.tet_block(*dims, (0.1, 0.07, 0.05), (0.0, -0.2, 0.3), mirrored)
        b = po.tet_block(*dims, (0.1, 0.07, 0.05), (0.0, -0.2, 0.3), mirrored)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_minstd_order_matches_oracle():
    for n in (1, 7, 1152, 331776):
        assert np.array_equal(xf.minstd_order(n), po.minstd_order(n))


def test_normal_fill_deterministic_and_normal():
    a = xf.normal_fill(1, 0.05, 200001)
    b = xf.normal_fill(1, 0.05, 200001)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, xf.normal_fill(2, 0.05, 200001))
    assert abs(a.mean()) < 3 * 0.05 / np.sqrt(a.size)
    assert abs(a.std() / 0.05 - 1.0) < 0.01
    # fourth moment of a normal: 3 sigma^4
    assert abs(np.mean(a.astype(np.float64) ** 4) / 0.05 ** 4 - 3.0) < 0.05


def test_bad_arguments_return_status_codes():
    L = xf.lib()
    xyz = np.zeros(3, np.float32)
    tets = np.zeros(4, np.uint32)
    s = np.ones(3, np.float32)
    assert L.xf_mesh_tet_block(0, 1, 1, s, s, 0, xyz, tets) == -1
    x = np.zeros(1, np.float32)
    assert L.xf_ldlt_solve(0, x, x, x) == -1
    nc = C.c_uint32()
    order = np.zeros(1, np.uint32)
    bad = np.array([0, 1, 2, 9], np.uint32)  # node 9 does not exist
    assert L.xf_mesh_colour(4, bad, 1, 0, order, order, C.byref(nc)) == -1


def _gpu_visible():
    try:
        xf.Context(0).close()
        return True
    except xf.XfError:
        return False


def test_context_without_gpu_fails_loudly():
    if _gpu_visible():
        pytest.skip("a GPU is visible here")
    with pytest.raises(xf.XfError):
        xf.Context(0)
