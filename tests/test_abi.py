"""The C-ABI library loads and exports every entry point include/eu3d.h declares (no GPU calls)."""
import ctypes
import os
import re

from eu3d_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "eu3d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(eu3d_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "eu3d_create" in syms and "eu3d_step" in syms and "eu3d_chem_batch" in syms
    assert set(syms) == set(api.EXPORTED), set(syms) ^ set(api.EXPORTED)


def test_library_exports_all_declared():
    lib = ctypes.CDLL(api.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_oracle_in_product():
    """The product library and package never reference the oracle."""
    pkg = os.path.join(ROOT, "eu3d_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp")):
                txt = open(os.path.join(dp, f), errors="replace").read()
                assert "import oracle" not in txt and "from oracle" not in txt and "liboracle" not in txt, f
    data = open(api.LIB_PATH, "rb").read()
    assert b"liboracle" not in data and b"orc_" not in data
