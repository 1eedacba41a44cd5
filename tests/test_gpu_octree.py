"""GPU octree construction (dwg_octree_build / dwg_octree_leaves) against the
oracle's recursive octreeDWG restatement: node counts, breadth-first leaf
cells (centres and widths bit-exact), term sets of the Barnes-Hut traversal
on degenerate inputs (duplicates, a single point, shallow trees with many
points per leaf), and the input normalisation of dwg_launch.
"""
import numpy as np
import pytest

import oracle as O
from dwg_amd import gpu, pipeline, shapes

pytestmark = pytest.mark.gpu


def _cloud(n, seed, dup=0):
    rng = np.random.default_rng(seed)
    P, N = shapes.sphere(n, seed)
    P = P + rng.normal(0, 0.003, P.shape)
    if dup:
        P = np.concatenate([P, P[:dup]])  # exact duplicates split to max_depth
        N = np.concatenate([N, N[:dup]])
    return P.astype(np.float64), N


@pytest.mark.parametrize("n,max_depth,leaf_size,dup", [(5000, 10, 1, 0), (5000, 10, 1, 300), (20000, 4, 1, 0),
                                                       (3000, 10, 8, 50), (1, 10, 1, 0), (2, 10, 1, 1)])
def test_octree_matches_oracle(ctx, n, max_depth, leaf_size, dup):
    P, N = _cloud(n, 3 + n % 7, dup)
    t = gpu.Octree(ctx, P, max_depth=max_depth, leaf_size=leaf_size)
    o = O.Octree(P, max_depth, leaf_size)
    info = t.info()
    assert info["nodes"] == o.num_nodes
    ch, cnt, depth = o.topology()
    assert info["leaves"] == int(((ch[:, 0] == -1) & (cnt > 0)).sum())
    assert info["depth"] == int(depth.max())
    if leaf_size == 1:
        for min_num in (0, 2):
            L, _ = O.octree_leaves(P, max_depth, min_num)
            cen, wid = t.leaves(min_num)
            assert cen.shape == L.reshape(-1, 3).shape
            assert np.array_equal(cen, L.reshape(-1, 3))
    # Barnes-Hut term sets: decisions use float CM/R of both trees
    A = np.full(len(P), 2e-6)
    t.update(N.astype(np.float32), A.astype(np.float32))
    o.precompute(N, A)
    rng = np.random.default_rng(n)
    Q = (P[rng.integers(0, len(P), 200)] + rng.normal(0, 0.05, (200, 3))).astype(np.float32)
    terms = t.terms(Q, 2.3).cpu().numpy()
    _, ref_terms = o.eval(Q.astype(np.float64), 2.3, kind="plain", count_terms=True)
    assert np.array_equal(terms, ref_terms)


def test_octree_leaf_widths_halve(ctx):
    P, _ = _cloud(4000, 9)
    t = gpu.Octree(ctx, P, max_depth=10)
    cen, wid = t.leaves(0)
    w0 = (P.max(0) - P.min(0)).max()
    lev = np.log2(w0 / wid)
    assert np.allclose(lev, np.round(lev)) and lev.min() >= 1 and lev.max() <= 10


def test_normalisation(ctx):
    P, _ = shapes.torus(30_000, 4)
    P = (P * 3.0 + np.array([5.0, -2.0, 1.0])).astype(np.float32)
    r = pipeline.reconstruct(ctx, P, grid_reso=48, final_reso=48, total_iter=1, source_mode=1, want_mesh=False)
    c = P.astype(np.float64).mean(0)
    assert np.allclose(r.center, c, rtol=0, atol=1e-12 * np.abs(c).max())
    s = (2 * np.abs(P.astype(np.float64) - np.asarray(r.center))).max()
    assert r.scale == s


def _octant_paths(P, max_depth):
    # each point's octant path from the root cell (centre of the bounding
    # box, width its largest extent), 3 bits a level, child bit x | y << 1 |
    # z << 2 set where the coordinate is ≥ the cell's centre: the split rule
    # of octree.cpp:171-335 in FP64, the arithmetic the build's
    # octree_codes_kernel uses
    lo, hi = P.min(0), P.max(0)
    c = np.tile((lo + hi) / 2, (len(P), 1))
    w = np.full(len(P), (hi - lo).max())
    k = np.zeros(len(P), np.uint64)
    for _ in range(max_depth):
        o = (P >= c).astype(np.uint64) << np.arange(3, dtype=np.uint64)
        o = o.sum(1, dtype=np.uint64)
        k = (k << np.uint64(3)) | o
        q = w / 4
        for d in range(3):
            c[:, d] += np.where((o >> np.uint64(d)) & np.uint64(1), q, -q)
        w = w / 2
    return k


@pytest.mark.parametrize("n,max_depth,dup", [(50_000, 10, 5000), (20_000, 21, 300), (300_000, 10, 0), (1, 10, 0)])
def test_octree_point_order_is_stable_sort_by_path(ctx, n, max_depth, dup):
    # the build's radix sort: points ordered by octant path, equal paths
    # (duplicates, or points sharing a cell at max_depth) in caller order —
    # numpy's stable argsort of the same paths
    P, _ = _cloud(n, 40 + n % 11, dup)
    t = gpu.Octree(ctx, P, max_depth=max_depth)
    ref = np.argsort(_octant_paths(P, max_depth), kind="stable")
    assert np.array_equal(t.point_order(), ref)
