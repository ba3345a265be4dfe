"""The C ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pagani.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ["pg_create", "pg_corr_build", "pg_corr_csr", "pg_bc", "pg_fw", "pg_modularity",
                 "pg_cp_binary", "pg_cp_weighted", "pg_degree", "pg_strength", "pg_bfs", "pg_triangles"]:
        assert must in names


def test_library_exports_every_symbol():
    import pagani_amd
    lib = ctypes.CDLL(pagani_amd.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_status_strings():
    import pagani_amd
    lib = pagani_amd.library()
    assert lib.pg_abi_version() == 1
    assert lib.pg_status_string(0) == b"ok"
    assert lib.pg_status_string(5) == b"threshold error"


def test_create_without_device_fails_cleanly():
    import pagani_amd
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(pagani_amd.PgError):
        pagani_amd.Context(0)
