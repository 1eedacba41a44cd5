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
