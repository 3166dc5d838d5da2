"""Pins the oracle's MAS restatement (or_mas_*) against what the mathematics
fixes: hierarchy invariants, cluster matrices = in-bank blocks of the Galerkin
operator R_L A R_L^T per level, C C^-1 = I, and the apply = the dense
sum_L P_L^T blockdiag(C_L)^-1 P_L r.  CPU only."""
import numpy as np
import pytest

from libuipc_amd import synth
from oracle import oracle as O
from tests.test_oracle import dense_from_upper

SYM = [(a, b) for a in range(16) for b in range(a, 16)]


def sym_index(a, b):
    return 16 * a - a * (a + 1) // 2 + b


def random_partition(rng, V, tets):
    """parts of <= 16 vertices, deliberately disconnected inside a part"""
    perm = rng.permutation(V)
    parts = np.empty(V, np.int32)
    k, p = 0, 0
    while k < V:
        sz = int(rng.integers(5, 17))
        parts[perm[k:k + sz]] = p
        k += sz
        p += 1
    return parts


def system(grid, contacts, seed, partition="geometric"):
    s = synth.fem_system(grid, contacts, seed=seed)
    if partition == "random":
        s.part_ids = random_partition(np.random.default_rng(seed), s.block_rows, s.tets)
    elif partition == "unpart":
        s.part_ids = s.part_ids.copy()
        s.part_ids[:: 7] = -1
    r, c, v, rp = O.convert_sym(s.block_rows, s.rows, s.cols, s.vals)
    return s, (r, c, v)


def slots_per_level(m, t):
    """slot of every vertex at every level (-1 = none)"""
    V, L = m.V, m.L
    S = np.full((L, V), -1, np.int64)
    S[0] = t["real_to_part"]
    for l in range(1, L):
        S[l] = t["coarse_tables"][:, l - 1]
    return S


def dense_clusters(A, S, level, nclusters):
    """in-bank blocks of R_L A R_L^T for every cluster, as (clusters, 136, 9)"""
    V = A.shape[0] // 3
    out = np.zeros((nclusters, 136, 9))
    sl = S[level]
    valid = np.where(sl >= 0)[0]
    # aggregate A into the level's nodes
    nodes = {int(x) for x in sl[valid]}
    idx = {x: i for i, x in enumerate(sorted(nodes))}
    P = np.zeros((len(idx), V))
    for v in valid:
        P[idx[int(sl[v])], v] = 1.0
    P3 = np.kron(P, np.eye(3))
    Ac = P3 @ A @ P3.T
    inv_idx = sorted(nodes)
    for a_i, a in enumerate(inv_idx):
        for b_i, b in enumerate(inv_idx):
            if a // 16 != b // 16 or b < a:
                continue
            blk = Ac[3 * a_i:3 * a_i + 3, 3 * b_i:3 * b_i + 3]
            out[a // 16, sym_index(a % 16, b % 16)] = blk.reshape(-1, order="F")
    return out


def expand_cluster(M):
    """136 upper blocks -> dense 48x48"""
    D = np.zeros((48, 48))
    for k, (a, b) in enumerate(SYM):
        blk = M[k].reshape(3, 3, order="F")
        D[3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
        if a != b:
            D[3 * b:3 * b + 3, 3 * a:3 * a + 3] = blk.T
    return D


@pytest.mark.parametrize("grid,partition", [(6, "geometric"), (6, "random"), (8, "geometric"), (5, "unpart")])
def test_hierarchy_invariants(grid, partition):
    s, _ = system(grid, 0, 1, partition)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    t = m.tables()
    V, L = m.V, m.L
    # level count from compute_num_levels on max(map_nodes, V)
    n, sz = 1, -(-max(m.N, V) // 16) * 16
    while sz > 16:
        sz //= 16
        n += 1
        sz = -(-sz // 16) * 16
    assert L == min(n + 1, 6)
    ls = t["level_sizes"]
    assert m.level1_begin == -(-max(-(-m.N // 16) * 16, V) // 16) * 16
    for l in range(1, L + 1):
        assert ls[l, 1] % 16 == 0  # bank-aligned level regions
    for l in range(1, L):
        assert ls[l + 1, 1] == ls[l, 1] + -(-ls[l, 0] // 16) * 16
    assert m.total == ls[L, 1]
    # every partitioned vertex has exactly one node per level, inside that level
    S = slots_per_level(m, t)
    for v in range(V):
        if t["real_to_part"][v] < 0:
            assert (s.part_ids[v] < 0)
            continue
        assert t["part_to_real"][t["real_to_part"][v]] == v
        for l in range(1, L):
            x = S[l, v]
            assert ls[l, 1] <= x < ls[l, 1] + ls[l, 0]
    # a level-(l+1) node's children are one in-bank component of level l
    gn = t["going_next"]
    for l in range(1, L - 1):
        for x in range(ls[l, 1], ls[l, 1] + ls[l, 0]):
            p = gn[x]
            assert ls[l + 1, 1] <= p < ls[l + 1, 1] + ls[l + 1, 0]
        kids = {}
        for x in range(ls[l, 1], ls[l, 1] + ls[l, 0]):
            kids.setdefault(gn[x], set()).add(x // 16)
        assert all(len(b) == 1 for b in kids.values())


def _align16(n):
    return -(-n // 16) * 16


def mesh_edges(V, tets=None, tris=None, edges=None):
    """undirected vertex pairs of the simplices (FEMMASPreconditioner::do_init's
    adjacency, fem_mas_preconditioner.cu:163-203)"""
    pairs = [np.zeros((0, 2), np.int64)]
    for simp, k in ((tets, 4), (tris, 3), (edges, 2)):
        if simp is None:
            continue
        S = np.asarray(simp, np.int64).reshape(-1, k)
        for a in range(k):
            for b in range(a + 1, k):
                pairs.append(S[:, [a, b]])
    E = np.concatenate(pairs)
    E = E[(E[:, 0] != E[:, 1]) & (E.min(1) >= 0) & (E.max(1) < V)]
    return np.unique(np.sort(E, axis=1), axis=0)


def brute_force_hierarchy(V, part_ids, E):
    """The MAS hierarchy from its definition, independently of or_mas_create
    (mas_preconditioner_engine.cu:57-466, 1958-2078; fem_mas_preconditioner.cu:291-356):
      level 0: partition p owns bank_align(|p|) slots from the running offset,
               its vertices in ascending id order;
      level l+1: the connected components, inside each 16-node bank of level l,
               of the graph whose edges are the mesh edges mapped to level-l
               nodes; numbered in order of each component's lowest member
               (= by bank, then lowest lane); level 1 begins at
               bank_align(max(bank_align(N), V)), level l+1 at
               begin_l + bank_align(count_l);
      going_next: a vertex's level-1 node at its vertex id, and every level-l
               node's level-(l+1) node at the node's slot.
    Returns (N, L, node_abs[L + 1][V] (-1 = none), level_sizes (L + 1, 2),
    going_next)."""
    part_ids = np.asarray(part_ids)
    live = np.where(part_ids >= 0)[0]
    psize = np.bincount(part_ids[live], minlength=int(part_ids.max()) + 1)
    poff = np.concatenate([[0], np.cumsum(-(-psize // 16) * 16)])
    N = int(poff[-1])
    slot = np.full(V, -1, np.int64)
    for p in range(psize.size):
        verts = np.where(part_ids == p)[0]  # ascending
        slot[verts] = poff[p] + np.arange(verts.size)
    # compute_num_levels(max(N, V)), capped at 6 levels
    L, sz = 1, _align16(max(N, V))
    while sz > 16:
        sz = _align16(sz // 16)
        L += 1
    L = min(L + 1, 6)
    rel = [slot]  # node index of every vertex per level, relative to the level's begin
    begin, count = [0], [N]
    Ep = E[(slot[E[:, 0]] >= 0) & (slot[E[:, 1]] >= 0)]
    for l in range(L):  # level l -> l + 1 (the last one only sizes the slot total)
        a, b = rel[l][Ep[:, 0]], rel[l][Ep[:, 1]]
        inb = (a // 16 == b // 16) & (a != b)
        a, b = a[inb], b[inb]
        nodes = np.unique(rel[l][live])
        lab = np.arange(max(int(nodes.max()) + 1 if nodes.size else 0, 1), dtype=np.int64)
        while True:  # min-label propagation: each node ends at its component's lowest member
            m = np.minimum(lab[a], lab[b])
            new = lab.copy()
            np.minimum.at(new, a, m)
            np.minimum.at(new, b, m)
            if np.array_equal(new, lab):
                break
            lab = new
        comps = np.unique(lab[nodes])
        nxt = np.full(V, -1, np.int64)
        nxt[live] = np.searchsorted(comps, lab[rel[l][live]])
        rel.append(nxt)
        count.append(comps.size)
        begin.append(_align16(max(_align16(N), V)) if l == 0 else begin[l] + _align16(count[l]))
    node_abs = [np.where(r >= 0, r + bg, -1) for r, bg in zip(rel, begin)]
    ls = np.zeros((L + 1, 2), np.int64)
    for l in range(1, L + 1):
        ls[l] = (count[l], begin[l])
    gsz = _align16(max(N, V)) * L
    gn = np.full(gsz, -1, np.int64)
    # a fine vertex's entry sits at its vertex id (build_level1 writes
    # going_next[part_to_real[slot]]; level1_begin >= V keeps it apart from the
    # coarse entries), a coarse node's at its absolute slot
    gn[live] = node_abs[1][live]
    for l in range(1, L):
        gn[node_abs[l][live]] = node_abs[l + 1][live]
    return N, L, node_abs, ls, gn


def check_hierarchy(m, t, part_ids, E):
    """asserts the tables of or_mas_create equal the brute force"""
    N, L, node_abs, ls, gn = brute_force_hierarchy(m.V, part_ids, E)
    assert (m.N, m.L) == (N, L)
    assert np.array_equal(t["real_to_part"], node_abs[0])
    assert np.array_equal(t["level_sizes"][1:], ls[1:]), (t["level_sizes"], ls)
    for l in range(1, L):
        assert np.array_equal(t["coarse_tables"][:, l - 1], node_abs[l]), f"level {l}"
    assert np.array_equal(t["going_next"][:gn.size], gn)
    assert m.total == ls[L, 1]


def _hierarchy_case(grid, partition, seed, simplices):
    s, _ = system(grid, 0, seed, partition)
    kw = {"tets": s.tets}
    if simplices == "faces+edges":
        T = np.asarray(s.tets).reshape(-1, 4)
        half = T.shape[0] // 2
        kw = {"tets": T[:half], "tris": T[half:, :3], "edges": T[half:, [0, 3]]}
    E = mesh_edges(s.block_rows, kw.get("tets"), kw.get("tris"), kw.get("edges"))
    return s, kw, E


@pytest.mark.parametrize("grid,partition,simplices", [
    (6, "geometric", "tets"), (8, "unpart", "tets"), (16, "geometric", "tets"), (16, "random", "tets"),
    (12, "unpart", "faces+edges"), (24, "random", "tets"), (24, "geometric", "faces+edges"),
    (40, "random", "tets")])
def test_hierarchy_brute_force(grid, partition, simplices):
    """every level of or_mas_create against in-bank connected components of
    the fine edges mapped to the level below"""
    s, kw, E = _hierarchy_case(grid, partition, 11, simplices)
    m = O.Mas(s.block_rows, s.part_ids, **kw)
    t = m.tables()
    check_hierarchy(m, t, s.part_ids, E)
    if grid >= 24:
        assert m.L >= 5
    if (grid, partition) == (40, "random"):
        assert m.L == 6  # the capped level count
    if partition == "random":
        # the random parts are disconnected: some partitions split at level 1
        assert t["level_sizes"][1, 0] > s.part_ids.max() + 1


@pytest.mark.parametrize("mistake", ["NO_CLOSURE", "DROP_NEIGHBOURS", "COARSE_SELF"])
def test_hierarchy_brute_force_detects_planted_mistakes(mistake):
    """the brute force fails on oracle builds with a planted mistake in the
    bank closure, the neighbour compaction or the coarse connect mask"""
    s, kw, E = _hierarchy_case(16, "random", 11, "tets")
    m = O.Mas(s.block_rows, s.part_ids, library=O.mutant(mistake), **kw)
    t = m.tables()
    with pytest.raises(AssertionError):
        check_hierarchy(m, t, s.part_ids, E)


def test_geometric_partition_levels():
    """2x2x4 blocks of the grid are connected: one level-1 node per part"""
    s, _ = system(16, 0, 1)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    t = m.tables()
    assert m.L == 4 and m.N == 4096
    assert t["level_sizes"][1, 0] == 256
    assert np.all(t["prefix_orig"] == 1)


@pytest.mark.parametrize("grid,contacts,partition", [(5, 40, "geometric"), (6, 60, "random"), (5, 30, "unpart")])
def test_scatter_is_galerkin(grid, contacts, partition):
    s, (r, c, v) = system(grid, contacts, 3, partition)
    A = dense_from_upper(s.block_rows, r, c, v)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    H = m.hessians()
    t = m.tables()
    S = slots_per_level(m, t)
    want = np.zeros_like(H)
    for l in range(m.L):
        want += dense_clusters(A, S, l, H.shape[0])
    assert np.allclose(H, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("fp32", [False, True])
def test_inverse(fp32):
    s, (r, c, v) = system(6, 50, 4, "random")
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v, inv_fp32=fp32)
    H, I = m.hessians(), m.inverses()
    eps = np.finfo(np.float32 if fp32 else np.float64).eps
    for k in range(H.shape[0]):
        D = expand_cluster(H[k])
        d = np.diag(D).copy()
        D[np.diag_indices(48)] = np.where(d == 0, 1.0, d)  # zero diagonal -> 1
        Inv = expand_cluster(I[k])
        err = np.abs(D @ Inv - np.eye(48)).max()
        # backward-stable inverse of an SPD matrix: ||D X - I|| <~ n cond(D) eps
        assert err < 8 * 48 * np.linalg.cond(D) * eps, (k, err)
        if not fp32:
            assert err < 1e-9


@pytest.mark.parametrize("partition", ["geometric", "random", "unpart"])
def test_apply_is_additive_schwarz(partition):
    s, (r, c, v) = system(6, 50, 5, partition)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    t = m.tables()
    S = slots_per_level(m, t)
    I = m.inverses()
    rng = np.random.default_rng(0)
    rv = rng.standard_normal(s.n)
    z = m.apply(rv)
    # dense: z = sum_L P_L^T blockdiag(C_L^-1) P_L r (+ 3x3 Jacobi for unpartitioned)
    V = m.V
    Rs = np.zeros((m.total, 3))
    for l in range(m.L):
        for vv in range(V):
            if S[l, vv] >= 0:
                Rs[S[l, vv]] += rv[3 * vv:3 * vv + 3]
    Zs = np.zeros_like(Rs)
    for k in range(I.shape[0]):
        Zs[16 * k:16 * k + 16] = (expand_cluster(I[k]) @ Rs[16 * k:16 * k + 16].reshape(-1)).reshape(16, 3)
    want = np.zeros(s.n)
    A = dense_from_upper(s.block_rows, r, c, v)
    for vv in range(V):
        if S[0, vv] < 0:
            blk = A[3 * vv:3 * vv + 3, 3 * vv:3 * vv + 3]
            want[3 * vv:3 * vv + 3] = np.linalg.solve(blk, rv[3 * vv:3 * vv + 3])
            continue
        for l in range(m.L):
            want[3 * vv:3 * vv + 3] += Zs[S[l, vv]]
    assert np.allclose(z, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    # positive on an SPD system
    assert rv @ z > 0


def test_one_level_is_block_jacobi():
    s, (r, c, v) = system(6, 40, 6, "random")
    A = dense_from_upper(s.block_rows, r, c, v)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    t = m.tables()
    rv = np.random.default_rng(1).standard_normal(s.n)
    z = m.apply(rv, active_levels=1)
    want = np.zeros(s.n)
    for bank in range(m.N // 16):
        verts = [int(x) for x in t["part_to_real"][16 * bank:16 * bank + 16] if x >= 0]
        ix = np.concatenate([np.arange(3 * x, 3 * x + 3) for x in verts])
        want[ix] = np.linalg.solve(A[np.ix_(ix, ix)], rv[ix])
    assert np.allclose(z, want, rtol=1e-9, atol=1e-10 * np.abs(want).max())


def test_single_vertex_clusters_are_3x3_jacobi():
    s, (r, c, v) = system(5, 20, 7)
    parts = np.arange(s.block_rows, dtype=np.int32)  # one vertex per part
    m = O.Mas(s.block_rows, parts, s.tets)
    m.assemble(r, c, v)
    rv = np.random.default_rng(2).standard_normal(s.n)
    z1 = m.apply(rv, active_levels=1)
    D = O.diag3(r, c, v, 0, s.block_rows)
    assert np.allclose(z1, O.diag3_apply(D, rv), rtol=1e-10, atol=1e-14)


def test_mas_pcg_beats_diag():
    s, (r, c, v) = system(12, 2000, 8)
    m = O.Mas(s.block_rows, s.part_ids, s.tets)
    m.assemble(r, c, v)
    D = O.diag3(r, c, v, 0, s.block_rows)
    _, rm = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=3, blk_cnt=s.block_rows, mas=m)], criterion=1, tol=1e-6)
    _, rd = O.pcg(s.block_rows, r, c, v, s.rhs, [dict(kind=1, blk_cnt=s.block_rows, data=D)], criterion=1, tol=1e-6)
    assert rm["converged"] and rd["converged"]
    assert rm["iters_exact"] < rd["iters_exact"]
    assert rm["rel_res_true"] < 1e-5
