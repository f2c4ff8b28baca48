"""The C-ABI library loads and exports every function declared in include/*.h.

No compute calls: these run without a GPU.
"""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b((?:bfs|b40c)_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for n in ["bfs_problem_create", "bfs_problem_reset", "bfs_enact", "bfs_extract_labels",
              "bfs_problem_destroy", "b40c_graph_rmat"]:
        assert n in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_symbol(lib, name):
    assert hasattr(lib, name), name


def test_status_strings(lib):
    lib.bfs_status_string.restype = ctypes.c_char_p
    assert b"overflow" in lib.bfs_status_string(3).lower()
    assert lib.bfs_strategy_frontier_type(0) == 0
    assert all(lib.bfs_strategy_frontier_type(s) == 1 for s in (1, 2, 3))


def test_invalid_arguments_fail_without_device(lib):
    import numpy as np
    from back40computing_amd import _native
    row = np.array([0, 1, 1], dtype=np.int32)
    bad_col = np.array([7], dtype=np.int32)  # out of range
    h = ctypes.c_void_p()
    rc = lib.bfs_problem_create(_native.i32ptr(row), _native.i32ptr(bad_col), 2, 1, 0, 0, ctypes.byref(h))
    assert rc == 1  # BFS_ERR_INVALID_ARG: validated on the host before any device call
    assert not h
    cfg = _native.BfsConfig()
    assert lib.bfs_config_default(ctypes.byref(cfg)) == 0
    assert cfg.wave_gather_threshold > 0 and cfg.cta_gather_threshold >= cfg.wave_gather_threshold
    assert cfg.fast_visited == 1


def test_header_constants_match_binding():
    import re
    from back40computing_amd import _native
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "..", "include", "b40c_bfs.h")) as f:
        text = f.read()
    m = re.search(r"#define BFS_FAST_VISITED_MAX (\d+)", text)
    assert m and int(m.group(1)) == _native.FAST_VISITED_MAX
    # the config struct's fields in declaration order
    body = text[text.index("typedef struct bfs_config {"):text.index("} bfs_config;")]
    names = re.findall(r"int32_t (\w+);", body)
    assert names == [f for f, _ in _native.BfsConfig._fields_]



def test_noalloc_scan_rejects_oversize_without_device(lib):
    # the allocation-free scan's 2^26-item limit is checked before any HIP call
    assert lib.b40c_scan_noalloc(0, None, None, (1 << 26) + 1, 1, None, None) == 1
    assert lib.b40c_scan_noalloc(0, None, None, -1, 1, None, None) == 1
