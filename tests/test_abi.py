"""The C ABI library loads and exports every symbol include/*.h declares; the
synthetic generator reproduces the closed-form mesh counts.  CPU only (no
compute calls into the GPU library)."""
import os
import re

import numpy as np
import pytest

from libuipc_amd import abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mi_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header", ["mi_linsys.h", "mi_producers.h", "mi_warmstart.h"])
def test_linsys_exports_every_declared_symbol(header):
    lib = abi.linsys()
    names = declared(header)
    assert len(names) >= (25 if header == "mi_linsys.h" else 3)
    for n in names:
        assert hasattr(lib, n), n
        assert n in abi.LINSYS_PROTOS, f"{n} has no ctypes prototype"
    assert lib.mi_version() == 1


def test_synth_exports_every_declared_symbol():
    lib = abi.synth()
    for n in declared("mi_synth.h"):
        assert hasattr(lib, n), n


def test_no_oracle_in_product():
    """the product never links or imports the oracle"""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "libuipc_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "oracle" not in txt.lower().replace("the oracle", ""), f
    import subprocess

    out = subprocess.run(["ldd", abi.LINSYS_SO], capture_output=True, text=True).stdout
    assert "oracle" not in out


@pytest.mark.parametrize("grid", [2, 5, 16])
def test_generator_counts(grid):
    s = synth.fem_system(grid, 10, seed=2)
    N = grid
    assert s.block_rows == N ** 3
    assert s.tets.shape == (5 * (N - 1) ** 3, 4)
    assert s.n_triplets == N ** 3 + 10 * 5 * (N - 1) ** 3 + 100
    # positive rest volumes, every vertex used
    assert np.unique(s.tets).size == N ** 3
    # partition: at most 16 vertices per part
    assert np.bincount(s.part_ids).max() <= 16
    # kinetic blocks first: diagonal, m I
    B = s.block_rows
    assert np.array_equal(s.rows[:B], np.arange(B)) and np.array_equal(s.cols[:B], np.arange(B))
    assert np.all(s.vals[:B, 0] > 0)
    # half blocks keep row <= col
    assert np.all(s.rows <= s.cols)


def test_generator_sizes_of_c2():
    sz = synth.sizes(64, 2_000_000)
    assert sz["block_rows"] == 262144
    assert sz["n_tets"] == 1_250_235
    assert sz["n_triplets"] == 262144 + 12_502_350 + 20_000_000


def test_generator_deterministic_across_threads():
    a = synth.fem_system(9, 500, seed=4, threads=1)
    b = synth.fem_system(9, 500, seed=4, threads=7)
    assert np.array_equal(a.rows, b.rows) and np.array_equal(a.vals, b.vals) and np.array_equal(a.rhs, b.rhs)


def test_generator_rejects_bad_input():
    with pytest.raises(RuntimeError):
        synth.fem_system(1, 0)
