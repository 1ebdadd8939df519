"""The C-ABI library loads and exports every entry point include/eu3d.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

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


def _blocks_written(ni, nj, nk):
    """Blocks k_cfl / k_imex_fused launch, enumerated: one per 64x4 tile of (i, j) that holds an
    interior cell, per k layer."""
    return sum(1 for _ in range(nk) for bi in range(0, ni, 64) for bj in range(0, nj, 4))


def test_cfl_partials_cover_every_block_on_thin_grids():
    """The CFL partial buffer holds one double per block its kernels write. Thin grids write more
    blocks than the old n3/64 + 64 sizing held (1x1x110: 110 blocks into 108 slots)."""
    for ni, nj, nk in [(1, 1, 110), (1, 1, 200), (1, 2, 300), (1, 5, 64), (8, 1, 500), (8, 64, 64),
                       (70, 7, 5), (2000, 1, 1), (384, 128, 128), (512, 256, 128)]:
        parts, nbytes = api.grid_sizes(ni, nj, nk)
        assert parts == _blocks_written(ni, nj, nk), (ni, nj, nk)
        n3 = (ni + 4) * (nj + 4) * (nk + 4)
        assert nbytes > n3 * 8 * 150
    old = lambda ni, nj, nk: (ni + 4) * (nj + 4) * (nk + 4) // 64 + 64
    assert _blocks_written(1, 1, 110) > old(1, 1, 110) and _blocks_written(8, 64, 64) > old(8, 64, 64)


def test_grid_sizes_rejects_bad_grids():
    with pytest.raises(api.Eu3dError):
        api.grid_sizes(0, 4, 4)
    with pytest.raises(api.Eu3dError):
        api.grid_sizes(4, 4, 4, bc=1)
