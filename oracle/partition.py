"""TEST INFRASTRUCTURE ONLY.  numpy restatement of the device mesh partitioner
(mi_mesh_partition, libuipc_amd/csrc/partition.hip), so that the GPU result can
be checked value for value.  The partitioner stands where
geometry::mesh_partition (src/geometry/mesh_partition.cpp:77-230) calls METIS;
the reference's own METIS answer comes from oracle/_ref (ref_mesh_partition in
oracle.py) and is compared on quality, not ids.

Same steps as the device code: handshake matching rounds on the cluster graph
(score w (1/|c| + 1/|d|) in float32, ties by a symmetric 64-bit pair hash,
then by the lower target id; mutual proposals merge into the lower id), then
clusters in order of their lowest vertex packed greedily into parts of at
most K vertices."""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _pair_hash(c: np.ndarray, d: np.ndarray, seed: int) -> np.ndarray:
    lo = np.minimum(c, d).astype(np.uint64)
    hi = np.maximum(c, d).astype(np.uint64)
    with np.errstate(over="ignore"):
        x = ((lo << np.uint64(32)) | hi) + np.uint64(0x9E3779B97F4A7C15) * np.uint64(seed + 1)
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x & np.uint64(0xFFFFFFFF)


def _adjacency(V: int, simplices: np.ndarray) -> np.ndarray:
    k = simplices.shape[1]
    e = np.concatenate([simplices[:, [i, j]] for i in range(k) for j in range(i + 1, k)])
    e = e[(e[:, 0] != e[:, 1]) & (e.min(1) >= 0) & (e.max(1) < V)]
    e = np.concatenate([e, e[:, ::-1]])
    return np.unique(e.astype(np.int64), axis=0)


def device_partition(V: int, K: int = 16, *, tets=None, tris=None, edges=None) -> tuple[np.ndarray, int]:
    """(part ids, number of parts), as mi_mesh_partition computes them"""
    if K <= 0:
        raise ValueError("part_max_size must be greater than zero")
    if V == 0:
        return np.zeros(0, np.int32), 0
    simp = None
    for a, k in ((tets, 4), (tris, 3), (edges, 2)):
        if a is not None and np.asarray(a).size:
            simp = np.asarray(a, np.int64).reshape(-1, k)
            break
    e = _adjacency(V, simp) if simp is not None else np.zeros((0, 2), np.int64)
    if len(e) == 0:
        return (np.arange(V) // K).astype(np.int32), (V - 1) // K + 1
    if 1 + (V - 1) // K <= 1:
        # connected components numbered by their lowest vertex
        lab = np.arange(V)
        while True:
            m = lab.copy()
            np.minimum.at(m, e[:, 0], lab[e[:, 1]])
            m = np.minimum(m, m[m])
            if np.array_equal(m, lab):
                break
            lab = m
        roots = np.unique(lab)
        return np.searchsorted(roots, lab).astype(np.int32), len(roots)
    cv = np.arange(V)
    sz = np.ones(V, np.int64)
    cs = np.arange(V)
    for rnd in range(64):
        c, d = cv[e[:, 0]], cv[e[:, 1]]
        keep = c != d
        pairs, w = np.unique(np.stack([c[keep], d[keep]], 1), axis=0, return_counts=True)
        c, d = pairs[:, 0], pairs[:, 1]
        ok = sz[c] + sz[d] <= K
        c, d, w = c[ok], d[ok], w[ok]
        sc, sd = sz[c].astype(np.float32), sz[d].astype(np.float32)
        ssum = (sz[c] + sz[d]).astype(np.float32)
        score = (w.astype(np.float32) * (ssum / (sc * sd))).astype(np.float32)
        packed = (score.view(np.uint32).astype(np.uint64) << np.uint64(32)) | _pair_hash(c, d, rnd)
        best = np.zeros(V, np.uint64)
        np.maximum.at(best, c, packed)
        hit = packed == best[c]
        target = np.full(V, 0xFFFFFFFF, np.int64)
        np.minimum.at(target, c[hit], d[hit])
        live = (sz > 0) & (target < V)
        mutual = np.zeros(V, bool)
        mutual[live] = target[target[live]] == cs[live]
        lo = mutual & (cs < target)
        hi = mutual & (cs > target)
        sz_new = sz.copy()
        sz_new[lo] = sz[lo] + sz[target[lo]]
        sz_new[hi] = 0
        redirect = cs.copy()
        redirect[hi] = target[hi]
        cv = redirect[cv]
        sz = sz_new
        if not lo.any():
            break
    part_of = np.full(V, -1, np.int64)
    parts, fill = 0, K
    for c in np.nonzero(sz)[0]:
        if fill + sz[c] > K:
            parts += 1
            fill = 0
        fill += sz[c]
        part_of[c] = parts - 1
    return part_of[cv].astype(np.int32), parts
