"""geometry::mesh_partition: the reference's partition (oracle/_ref, the
reference's METIS compiled in place) on the reference's own test case
(apps/tests/geometry/mesh_partition.cpp), and the device partitioner
(mi_mesh_partition) against it: the same answer where mesh_partition is not
METIS (point clouds, a single part = connected components), and METIS-class
quality (edge cut, MAS-PCG iterations) where it is."""
import time

import numpy as np
import pytest

from libuipc_amd import abi, synth
from oracle import oracle as O
from oracle import partition as OP

needs_ref = pytest.mark.skipif(not O.ref_partition_available(),
                               reason="oracle/_ref/libref_partition.so is built only where the reference tree exists")


def chain(n=12):
    return np.array([[i - 1, i] for i in range(1, n)], np.int32)


def check_valid(part, V, K):
    """mesh_partition's own validation (mesh_partition.cpp:195-229)"""
    assert part.shape == (V,) and part.min() >= 0
    sizes = np.bincount(part)
    assert sizes.max() <= K
    assert sizes.min() > 0  # ids contiguous from 0


def tet_edges(tets):
    return np.concatenate([tets[:, [i, j]] for i in range(4) for j in range(i + 1, 4)])


def cut_fraction(part, edges):
    return float(np.count_nonzero(part[edges[:, 0]] != part[edges[:, 1]])) / len(edges)


# ------------------------------------------------------------------ CPU
@needs_ref
def test_ref_partition_on_the_reference_chain_case():
    a = O.ref_mesh_partition(12, 3, edges=chain())
    b = O.ref_mesh_partition(12, 3, edges=chain())
    assert np.array_equal(a, b)
    assert a.max() >= 1
    check_valid(a, 12, 3)
    with pytest.raises(ValueError):
        O.ref_mesh_partition(12, 0, edges=chain())


@needs_ref
def test_ref_partition_fallbacks():
    # point cloud: i / K
    assert np.array_equal(O.ref_mesh_partition(50, 16), np.arange(50) // 16)
    # one part: connected components numbered by their lowest vertex
    e = np.array([[3, 5], [5, 9], [0, 7], [7, 2]], np.int32)
    assert np.array_equal(O.ref_mesh_partition(10, 16, edges=e), [0, 1, 0, 2, 3, 2, 4, 0, 5, 2])


@needs_ref
def test_ref_partition_grid16():
    t = synth.generate_topology(16, seed=5)
    p = O.ref_mesh_partition(t.block_rows, 16, tets=t.tets)
    check_valid(p, t.block_rows, 16)


def test_partition_restatement_contract():
    """the numpy restatement of the device partitioner keeps mesh_partition's
    contract: the reference's fallbacks exactly, valid parts otherwise"""
    assert np.array_equal(OP.device_partition(50, 16)[0], np.arange(50) // 16)
    e = np.array([[3, 5], [5, 9], [0, 7], [7, 2]], np.int32)
    assert np.array_equal(OP.device_partition(10, 16, edges=e)[0], [0, 1, 0, 2, 3, 2, 4, 0, 5, 2])
    p, n = OP.device_partition(12, 3, edges=chain())
    check_valid(p, 12, 3)
    assert n == p.max() + 1 >= 2
    t = synth.generate_topology(8, seed=2)
    p, n = OP.device_partition(t.block_rows, 16, tets=t.tets)
    check_valid(p, t.block_rows, 16)
    # a chain is cut into consecutive runs of at most 16
    p, n = OP.device_partition(1000, 16, edges=chain(1000))
    check_valid(p, 1000, 16)
    assert np.all(np.diff(p) >= 0)
    with pytest.raises(ValueError):
        OP.device_partition(12, 0, edges=chain())


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_device_partition_reference_chain_case(ctx):
    a, na = ctx.mesh_partition(12, 3, edges=chain())
    b, nb = ctx.mesh_partition(12, 3, edges=chain())
    assert np.array_equal(a, b) and na == nb == a.max() + 1
    assert a.max() >= 1
    check_valid(a, 12, 3)
    with pytest.raises(abi.MiError):
        ctx.mesh_partition(12, 0, edges=chain())


@pytest.mark.gpu
@needs_ref
@pytest.mark.parametrize("case", ["cloud", "components", "one_component"])
def test_device_partition_non_metis_paths_match_reference(ctx, case):
    if case == "cloud":
        V, kw = 50, {}
    elif case == "components":
        V, kw = 10, {"edges": np.array([[3, 5], [5, 9], [0, 7], [7, 2]], np.int32)}
    else:
        V, kw = 16, {"edges": chain(16)[::-1].copy()}
    p, n = ctx.mesh_partition(V, 16, **kw)
    want = O.ref_mesh_partition(V, 16, **kw)
    assert np.array_equal(p, want) and n == want.max() + 1


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["grid8", "grid16", "chain", "cloth", "random_edges"])
def test_device_partition_matches_restatement(ctx, case):
    """GPU ids equal the numpy restatement's, value for value"""
    rng = np.random.default_rng(4)
    if case.startswith("grid"):
        t = synth.generate_topology(int(case[4:]), seed=3)
        V, kw = t.block_rows, {"tets": t.tets}
    elif case == "chain":
        V, kw = 3000, {"edges": chain(3000)}
    elif case == "cloth":
        n = 40
        idx = np.arange(n * n, dtype=np.int32).reshape(n, n)
        a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, 1:].ravel()
        V, kw = n * n, {"tris": np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)])}
    else:
        V, kw = 2000, {"edges": rng.integers(0, 2000, size=(5000, 2)).astype(np.int32)}
    p, n = ctx.mesh_partition(V, 16, **kw)
    q, m = OP.device_partition(V, 16, **kw)
    assert n == m and np.array_equal(p, q)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["tris", "edges"])
def test_device_partition_surfaces_and_curves(ctx, kind):
    n = 96
    idx = np.arange(n * n, dtype=np.int32).reshape(n, n)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, 1:].ravel()
    tris = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)])
    if kind == "tris":
        p, np_ = ctx.mesh_partition(n * n, 16, tris=tris, edges=chain(n * n))  # triangles win over edges
    else:
        p, np_ = ctx.mesh_partition(n * n, 16, edges=chain(n * n))
    check_valid(p, n * n, 16)
    assert np_ == p.max() + 1
    if kind == "edges":
        # a chain is cut into consecutive runs of at most 16 (about 11 on average)
        assert np.all(np.diff(p) >= 0)
        assert np_ <= (n * n) // 10


@pytest.mark.gpu
@needs_ref
@pytest.mark.parametrize("grid", [16, 32])
def test_device_partition_quality_vs_metis(ctx, grid):
    t = synth.generate_topology(grid, seed=5)
    V = t.block_rows
    t0 = time.perf_counter()
    p, n = ctx.mesh_partition(V, 16, tets=t.tets)
    dt = time.perf_counter() - t0
    p2, _ = ctx.mesh_partition(V, 16, tets=t.tets)
    assert np.array_equal(p, p2)  # deterministic
    check_valid(p, V, 16)
    t0 = time.perf_counter()
    ref = O.ref_mesh_partition(V, 16, tets=t.tets)
    dref = time.perf_counter() - t0
    e = tet_edges(t.tets)
    cut, cut_ref = cut_fraction(p, e), cut_fraction(ref, e)
    print(f"\n[partition] grid {grid}: device {n} parts, cut {cut:.3f}, {dt * 1e3:.1f} ms | "
          f"METIS {ref.max() + 1} parts, cut {cut_ref:.3f}, {dref * 1e3:.0f} ms", flush=True)
    assert cut <= 1.2 * cut_ref
    assert n <= 1.25 * (ref.max() + 1)


@pytest.mark.gpu
@needs_ref
def test_device_partition_mas_pcg_iterations(ctx):
    """MAS-PCG on the 16^3 system with contacts: the device partition against
    METIS's and the generator's geometric one"""
    from libuipc_amd.device import LinearSystem

    s = synth.fem_system(16, 20000, seed=7)
    parts = {"geometric": s.part_ids,
             "metis": O.ref_mesh_partition(s.block_rows, 16, tets=s.tets),
             "device": ctx.mesh_partition(s.block_rows, 16, tets=s.tets)[0]}
    rows, cols, vals = ctx.to_device(s.rows), ctx.to_device(s.cols), ctx.to_device(s.vals)
    b = ctx.to_device(s.rhs)
    iters = {}
    for name, part in parts.items():
        ls = LinearSystem(ctx, s.block_rows)
        ls.assemble(rows, cols, vals)
        ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, s.n, part_ids=part, tets=s.tets)
        ls.assemble_preconditioners()
        x = ctx.alloc(s.n, np.float64)
        st = ls.solve(b, x, criterion=abi.MI_CRIT_REL_RESIDUAL, tol=1e-4, true_residual=1)
        assert st["converged"] == 1
        iters[name] = st["iters_exact"]
        ls.close()
    print(f"\n[partition] MAS-PCG iterations to 1e-4 at 16^3 + 20k contacts: {iters}", flush=True)
    assert iters["device"] <= 1.15 * iters["metis"] + 2


@pytest.mark.gpu
def test_device_partition_full_size(ctx):
    """64^3 vertex grid (1.25M tets): valid, feeds the MAS build"""
    from libuipc_amd.device import LinearSystem

    t = synth.generate_topology(64, seed=5)
    V = t.block_rows
    t0 = time.perf_counter()
    p, n = ctx.mesh_partition(V, 16, tets=t.tets)
    dt = time.perf_counter() - t0
    check_valid(p, V, 16)
    cut = cut_fraction(p, tet_edges(t.tets))
    ls = LinearSystem(ctx, V)
    h = ls.add_preconditioner(abi.MI_PRECOND_MAS, 0, 3 * V, part_ids=p, tets=t.tets, build=abi.MI_MAS_BUILD_DEVICE)
    g = ls.mas_export(h, matrices=False)
    print(f"\n[partition] 64^3: {n} parts, cut {cut:.3f}, {dt * 1e3:.1f} ms; MAS levels {g['L']}, "
          f"clusters {g['clusters']}", flush=True)
    assert n <= V // 12
