"""Matrix Market I/O, the format of the reference's debug dumps
(utils/matrix_market.h:34-110 for matrices, :120-158 for vectors).  The CPU
tests read hand-written files through the library's host readers; the GPU
tests write the assembled matrix and read it back."""
import numpy as np
import pytest

from libuipc_amd import abi, synth
from libuipc_amd.device import LinearSystem, read_mtx_block_triplets, read_vector_mtx
from oracle import oracle as O

# A 6x6 matrix written the way export_matrix_market writes it: banner, four
# comment lines, size line, then per block its 9 scalars row by row, 1-based.
# Blocks: (0,0) = [[4,1,0],[1,5,2],[0,2,6]], (0,1) = [[0,0,-1],[0,0,0],[3,0,0]]
# and (1,1) = diag(7, 8, 9).
KNOWN = """%%MatrixMarket matrix coordinate real general
% Exported from block-sparse triplet matrix
% Block dimension: 3x3
% Total rows: 6, Total cols: 6
% Total entries (including zeros): 27
6 6 27
1 1 4
1 2 1
1 3 0
2 1 1
2 2 5
2 3 2
3 1 0
3 2 2
3 3 6
1 4 0
1 5 0
1 6 -1
2 4 0
2 5 0
2 6 0
3 4 3
3 5 0
3 6 0
4 4 7
4 5 0
4 6 0
5 4 0
5 5 8
5 6 0
6 4 0
6 5 0
6 6 9
"""


def dense_of(B, rows, cols, vals):
    A = np.zeros((3 * B, 3 * B))
    for r, c, v in zip(rows, cols, vals):
        A[3 * r:3 * r + 3, 3 * c:3 * c + 3] += v.reshape(3, 3, order="F")
    return A


def test_read_known_matrix(tmp_path):
    p = tmp_path / "A.mtx"
    p.write_text(KNOWN)
    B, rows, cols, vals = read_mtx_block_triplets(str(p))
    assert B == 2
    assert rows.tolist() == [0, 0, 1] and cols.tolist() == [0, 1, 1]
    A = dense_of(B, rows, cols, vals)
    expect = np.zeros((6, 6))
    expect[:3, :3] = [[4, 1, 0], [1, 5, 2], [0, 2, 6]]
    expect[:3, 3:] = [[0, 0, -1], [0, 0, 0], [3, 0, 0]]
    expect[3:, 3:] = np.diag([7, 8, 9])
    assert np.array_equal(A, expect)


def test_read_zero_based_and_duplicates(tmp_path):
    p = tmp_path / "d.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 3\n0 0 1.5\n0 0 2.25\n2 1 -3\n")
    B, rows, cols, vals = read_mtx_block_triplets(str(p), one_based=False)
    assert B == 1 and rows.tolist() == [0] and cols.tolist() == [0]
    m = vals[0].reshape(3, 3, order="F")
    assert m[0, 0] == 3.75 and m[2, 1] == -3 and np.count_nonzero(m) == 2


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n3 1\n1\n2\n3\n",   # not coordinate
    "%%MatrixMarket matrix coordinate real general\n4 4 1\n1 1 1\n",  # not a multiple of 3
    "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1\n",  # truncated
    "%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1\n",  # out of range
])
def test_read_rejects_malformed(tmp_path, text):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(abi.MiError) as e:
        read_mtx_block_triplets(str(p))
    assert e.value.code == abi.MI_ERR_ARG


def test_read_vector(tmp_path):
    p = tmp_path / "b.mtx"
    x = np.array([1.0, -2.5, 1e-300, 0.1, np.pi])
    lines = ["%%MatrixMarket matrix array real general", "% Exported from dense vector",
             f"% Vector size: {x.size}", f"{x.size} 1"] + ["%.17g" % v for v in x]
    p.write_text("\n".join(lines) + "\n")
    y = read_vector_mtx(str(p))
    assert np.array_equal(x, y)


def test_read_synthetic_roundtrip_through_oracle(tmp_path):
    """A file written from the oracle's assembled upper BCOO (in the dump
    layout) reads back to the same blocks, bit for bit (%.17g round-trips
    doubles)."""
    s = synth.fem_system(4, 40, seed=3)
    r, c, v, _ = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    out = ["%%MatrixMarket matrix coordinate real general",
           f"{3 * s.block_rows} {3 * s.block_rows} {9 * len(r)}"]
    for t in range(len(r)):
        m = v[t].reshape(3, 3, order="F")
        for bi in range(3):
            for bj in range(3):
                out.append("%d %d %.17g" % (3 * r[t] + bi + 1, 3 * c[t] + bj + 1, m[bi, bj]))
    p = tmp_path / "A.mtx"
    p.write_text("\n".join(out) + "\n")
    B, rows, cols, vals = read_mtx_block_triplets(str(p))
    assert B == s.block_rows
    assert np.array_equal(rows, r) and np.array_equal(cols, c) and np.array_equal(vals, v)


def parse_coordinate(path):
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real general"
    body = [ln for ln in lines[1:] if not ln.startswith("%")]
    n, m, e = (int(t) for t in body[0].split())
    ent = np.array([ln.split() for ln in body[1:]], dtype=object)
    assert len(ent) == e
    return n, m, ent


@pytest.mark.gpu
def test_write_matrix_matches_export(ctx, c1_contact, tmp_path):
    s = c1_contact
    ls = LinearSystem(ctx, s.block_rows)
    r, c, v = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    ls.assemble(r, c, v)
    er, ec, ev, _ = ls.export_bcoo()
    p = tmp_path / "A.mtx"
    ls.write_mtx(str(p))
    n, m, ent = parse_coordinate(p)
    assert n == m == s.n and len(ent) == 9 * len(er)
    i = ent[:, 0].astype(np.int64) - 1
    j = ent[:, 1].astype(np.int64) - 1
    val = ent[:, 2].astype(np.float64)
    # entry order: block by block, then row by row inside the block
    t = np.repeat(np.arange(len(er)), 9)
    bi = np.tile(np.repeat(np.arange(3), 3), len(er))
    bj = np.tile(np.tile(np.arange(3), 3), len(er))
    assert np.array_equal(i, 3 * er[t] + bi) and np.array_equal(j, 3 * ec[t] + bj)
    assert np.array_equal(val, ev[t, bi + 3 * bj])
    # read back, reassemble, same matrix
    B, rr, cc, vv = read_mtx_block_triplets(str(p))
    ls2 = LinearSystem(ctx, B)
    ls2.assemble(ctx.to_device(rr), ctx.to_device(cc), ctx.to_device(vv))
    r2, c2, v2, _ = ls2.export_bcoo()
    assert np.array_equal(r2, er) and np.array_equal(c2, ec) and np.array_equal(v2, ev)
    ls2.close()
    ls.close()


@pytest.mark.gpu
def test_write_vector_roundtrip(ctx, tmp_path):
    rng = np.random.default_rng(5)
    x = rng.standard_normal(3001) * 10.0 ** rng.integers(-200, 200, 3001)
    d = ctx.to_device(x)
    p = tmp_path / "x.mtx"
    ctx.write_vector_mtx(str(p), d)
    head = p.read_text().splitlines()[:4]
    assert head == ["%%MatrixMarket matrix array real general", "% Exported from dense vector",
                    "% Vector size: 3001", "3001 1"]
    assert np.array_equal(read_vector_mtx(str(p)), x)
