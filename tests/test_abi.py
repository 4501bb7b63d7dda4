"""The C ABI library loads and exports every symbol include/mgldg.h declares
(no compute calls: no GPU is needed)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mgldg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mgldg_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    from mgldg._lib import SIGNATURES
    assert sorted(SIGNATURES) == declared_symbols()


def test_library_exports_every_declared_symbol():
    from mgldg._lib import LIB_PATH
    lib = ctypes.CDLL(LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_oracle_is_not_linked_by_the_product():
    from mgldg._lib import LIB_PATH
    data = open(LIB_PATH, "rb").read()
    assert b"orc_" not in data and b"liboracle" not in data


def test_null_handles_are_rejected():
    from mgldg._lib import lib
    L = lib()
    assert L.mgldg_ctx_create(0, None, None) == -1
    assert L.mgldg_quadtree_info(None, None) == -1
    assert L.mgldg_mg_destroy(None) == 0
