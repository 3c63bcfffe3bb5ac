"""The C-ABI library loads and exports every entry point of include/fd3d_tsn.h
(no compute: no GPU is needed)."""
import ctypes
import os

import pytest

from fd3d_tsn_amd import capi


def test_header_declares_entry_points():
    names = capi.exported_symbols()
    for must in ("fd3d_create", "fd3d_upload_medium", "fd3d_upload_fault_sw", "fd3d_upload_fault_fvw",
                 "fd3d_init_state", "fd3d_step", "fd3d_download_final", "fd3d_destroy"):
        assert must in names


def test_library_exports_all_symbols():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("HIP library not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in capi.exported_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_matches_header():
    # the ctypes mirror must list the header's fields in order
    import re
    text = open(capi.HEADER).read()
    body = text[text.index("typedef struct fd3d_config"):text.index("} fd3d_config;")]
    names = []
    for line in body.splitlines()[1:]:
        line = line.split("/*")[0].strip().rstrip(";")
        if not line:
            continue
        decl = re.sub(r"^(const\s+)?(int|float|double)\s*\*?", "", line)
        names += [n.strip().lstrip("*") for n in decl.split(",") if n.strip()]
    assert names == [f[0] for f in capi.Config._fields_]
