"""Sweep blocks on the host (xf_mesh_sweep_blocks; no reference counterpart:
the reference sweeps its MINSTD order serially, Geo.cpp:759-777).

The partition must make every colour (launch colour q, phase j) node-disjoint
and every launch colour's blocks node-disjoint; those two properties are what
make a block launch equal the serial sweep in the reported order. The block of
a tet is checked against an independent restatement: breadth-first levels
from scipy's csgraph (from a source joined to the nodes within 1e-4 of the
range of the least d . p, float32 arithmetic; an unreached component restarts
at its node of least d . p, ties by index), the direction pair chosen by the
fewest tets of one block on one node (first on ties, stopping at a quarter of
the largest valence), and the pair of minimum levels per tet. On uniform
Freudenthal blocks the x / (y - z) pair gives blocks of 6 phases: 24 phases
per substep, the node valence, the least any serial-equivalent schedule can
have."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

import workloads
import xpbdfem as xf


def node_disjoint(tets, group):
    """True if no two tets of one group share a node."""
    t = tets.reshape(-1, 4).astype(np.int64)
    pairs = np.stack([np.repeat(group.astype(np.int64), 4), t.ravel()], 1)
    return len(np.unique(pairs, axis=0)) == len(pairs)


def check_invariants(tets, sb):
    t = tets.reshape(-1, 4)
    B, P, L = sb["block"], sb["phase"], sb["launch"]
    nb = L.size
    assert np.array_equal(np.unique(B), np.arange(nb))  # no empty block
    assert np.all(np.diff(L.astype(np.int64)) >= 0)  # blocks numbered by launch colour first
    assert L.max() <= 3
    # phases of a block: node-disjoint, numbered 0 .. count - 1
    assert node_disjoint(t, B.astype(np.int64) * 64 + P)
    cnt = np.bincount(B)
    top = np.zeros(nb, np.int64)
    np.maximum.at(top, B, P.astype(np.int64))
    used = np.zeros((nb, 64), bool)
    used[B, P] = True
    assert np.array_equal(used.sum(1), top + 1)
    assert sb["max_phases"] == top.max() + 1
    # blocks of one launch colour share no node
    for q in range(4):
        sel = L[B] == q
        if not sel.any():
            continue
        nodes = t[sel].ravel().astype(np.int64)
        blk = np.repeat(B[sel].astype(np.int64), 4)
        pairs = np.unique(np.stack([nodes, blk], 1), axis=0)
        assert len(np.unique(pairs[:, 0])) == len(pairs)
    # hence every colour (q, j) is node-disjoint
    assert node_disjoint(t, L[B].astype(np.int64) * 64 + P)
    return cnt


DIRS = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, -1), (0, 1, 1), (1, 0, -1), (1, 0, 1), (1, -1, 0), (1, 1, 0)]
PAIRS = [(0, 1), (0, 3), (0, 4), (0, 2), (1, 2), (1, 5), (1, 6), (2, 7), (2, 8)]


def scipy_levels(xyz, tets, d):
    p = xyz.reshape(-1, 3).astype(np.float32)
    t = tets.reshape(-1, 4).astype(np.int64)
    n = p.shape[0]
    f = np.float32
    v = (f(d[0]) * p[:, 0] + f(d[1]) * p[:, 1]) + f(d[2]) * p[:, 2]
    used = np.zeros(n, bool)
    used[t.ravel()] = True
    lo, hi = v[used].min(), v[used].max()
    tol = f(lo + f(f(1e-4) * f(hi - lo)))
    seeds = np.where(used & (v <= tol))[0]
    rows = np.repeat(t, 4, axis=1).ravel()
    cols = np.tile(t, (1, 4)).ravel()
    rows = np.concatenate([rows, np.full(seeds.size, n)])  # source node n joined to the seeds
    cols = np.concatenate([cols, seeds])
    A = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(n + 1, n + 1))
    A = ((A + A.T) > 0).astype(np.float64)
    lev = np.full(n, -1, np.int64)
    dist = csgraph.shortest_path(A, unweighted=True, indices=[n])[0][:n]
    ok = np.isfinite(dist)
    lev[ok] = dist[ok].astype(np.int64) - 1
    for u in sorted(np.where(used & (lev < 0))[0], key=lambda u: (v[u], u)):
        if lev[u] < 0:
            dist = csgraph.shortest_path(A[:n, :n], unweighted=True, indices=[u])[0]
            ok = np.isfinite(dist) & (lev < 0)
            lev[ok] = dist[ok].astype(np.int64)
    return lev


def chosen_levels(xyz, tets):
    t = tets.reshape(-1, 4)
    n = xyz.size // 3
    valence = np.bincount(t.ravel(), minlength=n).max()
    floor_bound = (valence + 3) // 4
    cache = {}
    best, best_bound = None, None
    for a, b in PAIRS:
        for k in (a, b):
            if k not in cache:
                cache[k] = scipy_levels(xyz, tets, DIRS[k])
        k1, k2 = cache[a][t].min(1), cache[b][t].min(1)
        key = k1 * (1 << 32) + k2
        pairs = np.stack([np.repeat(key, 4), t.ravel().astype(np.int64)], 1)
        bound = np.unique(pairs, axis=0, return_counts=True)[1].max()
        if best_bound is None or bound < best_bound:
            best, best_bound = (a, b), bound
        if best_bound <= floor_bound:
            break
    return cache[best[0]], cache[best[1]], best


def check_against_scipy(xyz, tets, sb):
    t = tets.reshape(-1, 4)
    l1, l2, pair = chosen_levels(xyz, tets)
    assert np.all(l1[t].max(1) - l1[t].min(1) <= 1) and np.all(l2[t].max(1) - l2[t].min(1) <= 1)
    k1, k2 = l1[t].min(1), l2[t].min(1)
    q = (k1 & 1) | ((k2 & 1) << 1)
    keys = sorted(set(zip(q.tolist(), k1.tolist(), k2.tolist())))
    index = {k: i for i, k in enumerate(keys)}
    want = np.array([index[k] for k in zip(q.tolist(), k1.tolist(), k2.tolist())])
    assert np.array_equal(sb["block"], want)
    assert np.array_equal(sb["launch"], np.array([k[0] for k in keys]))
    return pair


def test_structured_beam_blocks():
    """12x4x4 cells: levels along x and y - z, so a block holds the cells of
    one x index on one (y - z) diagonal; 96 blocks of 6 phases (the 24 tets
    of a node split 6 per block over 4 launch colours)."""
    wl = workloads.beam_neo_12x4x4()
    sb = xf.mesh_sweep_blocks(wl["xyz"], wl["tets"])
    check_invariants(wl["tets"], sb)
    assert check_against_scipy(wl["xyz"], wl["tets"], sb) == (0, 3)
    assert sb["launch"].size == 12 * (4 + 4) and sb["max_phases"] == 6
    assert np.all(np.bincount(sb["launch"]) == 24)
    p = np.rint(wl["xyz"].reshape(-1, 3) / 0.1).astype(np.int64)[wl["tets"].reshape(-1, 4)]
    cell_x = p[:, :, 0].min(1)
    for b in range(sb["launch"].size):
        assert np.ptp(cell_x[sb["block"] == b]) == 0


@pytest.mark.parametrize("cells", [(96, 24, 24), (24, 8, 6)])
def test_big_beam_six_phases(cells):
    W, H, D = cells
    xyz, tets = xf.tet_block(W, H, D, (0.1, 0.1, 0.1), (0.0, -0.5 * H * 0.1, -0.5 * D * 0.1))
    sb = xf.mesh_sweep_blocks(xyz, tets)
    check_invariants(tets, sb)
    assert sb["launch"].size == W * (H + D)
    assert sb["max_phases"] == 6


def test_mirrored_block_has_no_six_phase_pair():
    """Mirrored blocks alternate the cell diagonal: every pair puts 12 tets of
    one block on some node, and the first pair (x, y) is kept."""
    xyz, tets = xf.tet_block(6, 4, 4, (0.1, 0.1, 0.1), (0.0, 0.0, 0.0), mirrored=True)
    sb = xf.mesh_sweep_blocks(xyz, tets)
    check_invariants(tets, sb)
    assert check_against_scipy(xyz, tets, sb) == (0, 1)
    assert sb["max_phases"] >= 12


@pytest.mark.parametrize("colour_order", [xf.COLOUR_NATURAL, xf.COLOUR_MINSTD])
def test_relabelled_mesh_and_visit_orders(colour_order):
    """Node numbering and tet order permuted, mirrored block: the partition
    depends on the mesh, not its numbering; phases follow the visit order."""
    rng = np.random.default_rng(3)
    xyz, tets = xf.tet_block(7, 5, 4, (0.1, 0.13, 0.09), (0.3, -0.2, 0.1), mirrored=True)
    n = xyz.size // 3
    perm = rng.permutation(n).astype(np.uint32)
    inv = np.argsort(perm).astype(np.uint32)
    xyz2 = xyz.reshape(-1, 3)[perm].ravel().copy()
    t2 = inv[tets.reshape(-1, 4)][rng.permutation(tets.size // 4)].ravel().astype(np.uint32)
    sb = xf.mesh_sweep_blocks(xyz2, t2, colour_order)
    check_invariants(t2, sb)
    check_against_scipy(xyz2, t2, sb)
    again = xf.mesh_sweep_blocks(xyz2, t2, colour_order)
    for k in ("block", "phase", "launch"):
        assert np.array_equal(sb[k], again[k])


def test_components_restart_levels():
    """Two boxes in one mesh, the second out of reach of the first's search
    and rotated: its levels restart at its node of least coordinate."""
    xa, ta = xf.tet_block(4, 3, 3, (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    xb, tb = xf.tet_block(3, 4, 2, (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    c, s = np.cos(0.4), np.sin(0.4)
    pb = xb.reshape(-1, 3) @ np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float32).T + np.float32([2.0, -0.05, 0.3])
    xyz = np.concatenate([xa, pb.astype(np.float32).ravel()])
    tets = np.concatenate([ta, tb + xa.size // 3]).astype(np.uint32)
    sb = xf.mesh_sweep_blocks(xyz, tets)
    check_invariants(tets, sb)
    check_against_scipy(xyz, tets, sb)


def test_unused_nodes_and_errors():
    xyz, tets = xf.tet_block(3, 2, 2, (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    extra = np.concatenate([xyz, np.float32([-5.0, -5.0, -5.0])])  # an unused node of least x and y
    sb = xf.mesh_sweep_blocks(extra, tets)
    check_invariants(tets, sb)
    check_against_scipy(extra, tets, sb)
    bad = tets.copy()
    bad[5] = xyz.size // 3 + 7
    with pytest.raises(xf.XfError):
        xf.mesh_sweep_blocks(xyz, bad)
