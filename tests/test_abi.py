"""The C ABI: the built library loads and exports every entry point that
include/specmc_hip.h declares, and the ctypes binding matches the header's
parameter lists (no compute call; no device needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "specmc_hip.h")


def _declarations():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^(?:int|long long|long|const char \*)\s*(smx_\w+)\(([^)]*)\);", text, re.M | re.S):
        args = " ".join(m.group(2).split())
        out[m.group(1)] = [] if args in ("", "void") else [a.strip() for a in args.split(",")]
    return out


def test_header_declares_the_bound_entry_points():
    from specmc_amd.runtime import SIGNATURES
    decl = _declarations()
    assert set(decl) == set(SIGNATURES)
    for name, args in decl.items():
        assert len(args) == len(SIGNATURES[name]), name
        for a, t in zip(args, SIGNATURES[name]):
            if "*" in a:
                assert t is ctypes.c_void_p, (name, a)
            elif a.startswith("long long"):
                assert t is ctypes.c_longlong, (name, a)
            elif a.startswith("long"):
                assert t is ctypes.c_long, (name, a)
            elif a.startswith("double"):
                assert t is ctypes.c_double, (name, a)
            else:
                assert a.startswith("int") and t is ctypes.c_int, (name, a)


def test_library_exports_every_declared_symbol():
    from specmc_amd import runtime
    if not os.path.exists(runtime.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(runtime.LIB_PATH)
    missing = [n for n in _declarations() if not hasattr(lib, n)]
    assert not missing, missing
    lib.smx_version.restype = ctypes.c_int
    assert lib.smx_version() >= 1


def test_dense_solve_mw_checks_its_arguments():
    """The multi-workgroup dense solve sizes its workspace and rejects a
    missing or short workspace and overlapping vectors before any launch."""
    from specmc_amd import runtime
    if not os.path.exists(runtime.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = runtime.load()
    n, nm = 3168, 5
    # header 64 B, the active list rounded to 256 B, one 8-byte word per
    # entry (FP32 work) or two (FP64 work)
    assert lib.dense_lu_solve_mw_bytes(2, nm, n) == 64 + 256 + nm * n * 8
    assert lib.dense_lu_solve_mw_bytes(0, nm, n) == 64 + 256 + nm * n * 16
    assert lib.dense_lu_solve_mw_bytes(3, nm, n) == 0
    need = lib.dense_lu_solve_mw_bytes(2, nm, n)
    fake = 1 << 40           # never dereferenced: the checks come first
    with pytest.raises(RuntimeError, match="workspace"):
        lib.dense_lu_solve_mw(2, nm, n, fake, n * n, fake, fake, fake + (1 << 30), None,
                              None, need, None)
    with pytest.raises(RuntimeError, match="workspace"):
        lib.dense_lu_solve_mw(2, nm, n, fake, n * n, fake, fake, fake + (1 << 30), None,
                              fake, need - 8, None)
    with pytest.raises(RuntimeError, match="overlap"):
        lib.dense_lu_solve_mw(2, nm, n, fake, n * n, fake, fake, fake + 8, None, fake,
                              need, None)
