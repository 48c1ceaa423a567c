"""The C ABI boundary: libmbvh.so loads, exports every entry point include/mbvh.h declares, and
reports errors through status codes (no compute calls here; a GPU is not needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mbvh_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mbvh.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint32_t|const char\*)\s+(mbvh_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("mbvh_build", "mbvh_mesh_upload", "mbvh_raycast_first", "mbvh_raycast_first_d",
                 "mbvh_closest_point", "mbvh_closest_point_d", "mbvh_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(M.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(M.EXPORTED_SYMBOLS)


def test_version_and_device_count():
    lib = M.library()
    assert lib.mbvh_abi_version() == 3
    n = M.device_count()
    assert n >= 0


def test_context_without_device_fails_cleanly():
    if M.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(M.MbvhError) as e:
        M.Context(0)
    assert e.value.code == -1
    assert "no HIP device" in str(e.value)


def test_bad_handles():
    lib = M.library()
    assert lib.mbvh_tree_info(None, None, None, None) == -3
    assert lib.mbvh_mesh_info(None, None, None, None) == -3
    assert lib.mbvh_ctx_synchronize(None) == -3
    # a query with a NULL context is rejected before touching any device
    assert lib.mbvh_raycast_first(None, None, None, None, 0, 0, 0.0, 1.0, None, None, None, None) == -3
    assert b"invalid context" in lib.mbvh_last_error()
    lib.mbvh_tree_free(None)
    lib.mbvh_mesh_free(None)
    lib.mbvh_ctx_destroy(None)


def test_build_roundtrip_through_abi():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([0, 1, 2], np.uint16)
    tree = M.build(pos, idx)
    assert len(tree.roots) == 1 and tree.roots[0].nbytes == 32
    h = tree.roots[0].view(np.uint16)
    assert h[15] == 0xFFFF and h[14] == 1
