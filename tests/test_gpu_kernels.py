"""GPU parity of each libdwg_hip kernel against the CPU oracle (C ABI calls).

Tolerances (FP32 kernels vs FP64 oracle):
  * direct / tree sums: |Δw| ≤ 2e-5 + 1e-4·|w| on unit-scale sums
  * gradients: |Δ∇w| ≤ 1e-4·max|∇w|
  * tree term counts, kNN indices and distances, MC triangle counts: exact
"""
import numpy as np
import pytest
import torch

import oracle as O
from mc_order import to_brick_order
from dwg_amd import gpu, shapes

pytestmark = pytest.mark.gpu


def _sphere(n, seed):
    P, N = shapes.sphere(n, seed)
    return P, N, np.full(n, 4 * np.pi / n, np.float32)


def _queries(P, m, seed, lo=-1.5, hi=1.5):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(lo, hi, (m, 3)), P[: m // 4]]).astype(np.float32)


def _close(a, b, atol=2e-5, rtol=1e-4):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    err = np.abs(a - b)
    assert np.all(err <= atol + rtol * np.abs(b)), f"max err {err.max()} at {err.argmax()}"


@pytest.mark.parametrize("kahan", [False, True])
def test_direct_plain(ctx, kahan):
    P, N, A = _sphere(2000, 1)
    Q = _queries(P, 400, 2)
    w = ctx.gwn_direct(P, N, A, Q, kahan=kahan).cpu().numpy()
    ref = O.direct(P.astype(float), N.astype(float), A.astype(float), Q.astype(float))
    _close(w, ref)


def test_direct_screened_radii(ctx):
    P, N, A = _sphere(2000, 3)
    P = P * 0.5
    A = np.full(len(P), 2e-4, np.float32)
    Q = _queries(P, 400, 4, -0.8, 0.8)
    rad = np.random.default_rng(5).uniform(0.0015, 0.015, len(P)).astype(np.float32)
    for sigma in [0.0, 10.0, 20.0]:
        for radii in [None, rad]:
            w = ctx.gwn_direct(P, N, A, Q, sigma=sigma, screened=True, radii=radii, kahan=True).cpu().numpy()
            ref = O.direct(P.astype(float), N.astype(float), A.astype(float), Q.astype(float), kind="screened",
                           sigma=sigma, radii=None if radii is None else radii.astype(float))
            scale = np.abs(ref).max()
            _close(w, ref, atol=2e-5 * max(scale, 1e-3), rtol=2e-4)


def test_direct_gradient(ctx):
    P, N, A = _sphere(1500, 6)
    rng = np.random.default_rng(7)
    Q = rng.uniform(-1.4, 1.4, (300, 3)).astype(np.float32)
    Q = Q[np.abs(np.linalg.norm(Q, axis=1) - 1) > 0.05]
    w, g = ctx.gwn_direct(P, N, A, Q, grad=True)
    rw, rg = O.direct_grad(P.astype(float), N.astype(float), A.astype(float), Q.astype(float))
    _close(w.cpu().numpy(), rw)
    g = g.cpu().numpy()
    assert np.abs(g - rg).max() <= 1e-4 * np.abs(rg).max()


def _tree_pair(ctx, P, N, A, leaf_size=1, max_depth=10):
    t = gpu.Octree(ctx, P.astype(np.float64), A.astype(np.float64), max_depth=max_depth, leaf_size=leaf_size)
    t.update(N, A)
    o = O.Octree(P.astype(np.float64), max_depth, leaf_size)
    o.precompute(N.astype(np.float64), A.astype(np.float64))
    return t, o


@pytest.mark.parametrize("leaf_size", [1, 8])
def test_tree_terms_and_values(ctx, leaf_size):
    P, N, A = _sphere(4000, 8)
    t, o = _tree_pair(ctx, P, N, A, leaf_size)
    info = t.info()
    assert info["nodes"] == o.num_nodes
    Q = _queries(P, 800, 9)
    for beta in [1.0, 2.3, 4.0]:
        terms = t.terms(Q, beta).cpu().numpy()
        ref_w, ref_terms = o.eval(Q.astype(float), beta, kind="plain", count_terms=True)
        assert np.array_equal(terms, ref_terms)
        w = t.eval(Q, beta, screened=False).cpu().numpy()
        _close(w, ref_w)
        sw = t.eval(Q, beta, sigma=10.0, screened=True).cpu().numpy()
        sref = o.eval(Q.astype(float), beta, kind="screened", sigma=10.0)
        _close(sw, sref, atol=2e-5 * np.abs(sref).max(), rtol=2e-4)


def _grid_points_f32(origin, h, res):
    i = np.arange(res[0], dtype=np.float64)
    j = np.arange(res[1], dtype=np.float64)
    k = np.arange(res[2], dtype=np.float64)
    o = np.asarray(origin, np.float32).astype(np.float64)
    hh = float(np.float32(h))
    # fma(i, h, o) in float: one rounding of the exact value
    X = (i * hh + o[0]).astype(np.float32)
    Y = (j * hh + o[1]).astype(np.float32)
    Z = (k * hh + o[2]).astype(np.float32)
    G = np.stack(np.meshgrid(X, Y, Z, indexing="ij"), -1)
    return np.ascontiguousarray(G.transpose(2, 1, 0, 3).reshape(-1, 3))


def test_tree_grid_matches_points(ctx):
    P, N, A = _sphere(3000, 10)
    P = P * 0.5
    A = np.full(len(P), 2e-6, np.float32)
    t, o = _tree_pair(ctx, P, N, A)
    origin, h, res = O.voxel_grid(P.astype(float), 0.05, 23)
    grid = gpu.make_grid(origin, h, res)
    G = _grid_points_f32(grid.origin, grid.h, res)
    wg = t.eval_grid(grid, 2.3, sigma=20.0).cpu().numpy()
    wp = t.eval(G, 2.3, sigma=20.0).cpu().numpy()
    # same per-query term set; the wave-union traversal may order the adds differently
    _close(wg, wp, atol=1e-6 * np.abs(wp).max(), rtol=1e-5)
    ref = o.eval(G.astype(float), 2.3, kind="screened", sigma=20.0)
    _close(wg, ref, atol=2e-5 * np.abs(ref).max(), rtol=2e-4)
    assert t.grid_terms(grid, 2.3) == o.eval(G.astype(float), 2.3, kind="plain", count_terms=True)[1].sum()


def test_tree_grid_deep_octree(ctx):
    """The grid walk on an octree of the deepest depth the walk accepts (21
    levels: a tight cluster of nearly coincident points splits to the
    depth limit), where its stack is sized by the tree's depth: per-query
    term counts equal the oracle's, values match the explicit-point walk
    and the FP64 oracle, with the four points per lane of the grid bricks
    and partial bricks at every grid edge."""
    P, N, A = _sphere(2000, 40)
    P = 0.5 * P
    rng = np.random.default_rng(41)
    C = (P[7] + rng.normal(0, 1e-6, (300, 3))).astype(np.float32)
    P = np.concatenate([P, C]).astype(np.float32)
    N = np.concatenate([N, np.repeat(N[7:8], len(C), 0)]).astype(np.float32)
    A = np.full(len(P), 2e-6, np.float32)
    t, o = _tree_pair(ctx, P, N, A, max_depth=21)
    assert t.info()["depth"] >= 19
    origin, h, res = O.voxel_grid(P.astype(float), 0.05, 21)
    grid = gpu.make_grid(origin, h, res)
    G = _grid_points_f32(grid.origin, grid.h, res)
    for beta in (1.0, 2.3):
        wg = t.eval_grid(grid, beta, sigma=20.0).cpu().numpy()
        wp = t.eval(G, beta, sigma=20.0).cpu().numpy()
        _close(wg, wp, atol=1e-6 * np.abs(wp).max(), rtol=1e-5)
        ref = o.eval(G.astype(float), beta, kind="screened", sigma=20.0)
        _close(wg, ref, atol=2e-5 * np.abs(ref).max(), rtol=2e-4)
        assert t.grid_terms(grid, beta) == o.eval(G.astype(float), beta, kind="plain", count_terms=True)[1].sum()


def test_tree_bf16_far_field_error(ctx):
    P, N, A = _sphere(20000, 11)
    t, o = _tree_pair(ctx, P, N, A)
    Q = _queries(P, 1000, 12)
    ref = O.direct(P.astype(float), N.astype(float), A.astype(float), Q.astype(float))
    w32 = t.eval(Q, 2.3, screened=False).cpu().numpy()
    w16 = t.eval(Q, 2.3, screened=False, bf16=True).cpu().numpy()
    e32 = np.abs(w32 - ref).mean()
    e16 = np.abs(w16 - ref).mean()
    # bf16 moments add at most a small fraction of the order-0 truncation error
    assert e16 < 1.25 * e32 + 1e-4


@pytest.mark.parametrize("k", [1, 10, 15, 16, 32])
def test_knn_bit_exact(ctx, k):
    P, _, _ = _sphere(20000, 13)
    P = P * 0.5
    rng = np.random.default_rng(14)
    Q = np.concatenate([P[:500] + rng.normal(0, 0.01, (500, 3)).astype(np.float32),
                        rng.uniform(-0.7, 0.7, (100, 3)).astype(np.float32), P[1000:1100]]).astype(np.float32)
    idx = gpu.KnnIndex(ctx, P, k_hint=k)
    got, d2 = idx.query(Q, k, d2=True)
    ref, rd2 = O.knn_brute(P, Q, k, return_d2=True)
    assert np.array_equal(got.cpu().numpy(), ref)
    assert np.array_equal(d2.cpu().numpy(), rd2)


def test_knn_ties_skip_and_mask(ctx):
    L = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    L = (L * 0.01).astype(np.float32)
    rng = np.random.default_rng(15)
    skip = (rng.uniform(size=len(L)) < 0.3).astype(np.uint8)
    qmask = (rng.uniform(size=len(L)) < 0.5).astype(np.uint8)
    idx = gpu.KnnIndex(ctx, L, k_hint=10)
    got = idx.query(L, 10, skip=torch.from_numpy(skip), qmask=torch.from_numpy(qmask)).cpu().numpy()
    ref = O.knn_brute(L, L, 10, skip=skip)
    sel = qmask == 1
    assert np.array_equal(got[sel], ref[sel])
    assert np.all(got[~sel] == -1)


def test_marching_cubes_matches_oracle(ctx):
    res = np.array([30, 27, 25], np.int32)
    origin = np.array([-1.1, -0.9, -1.0])
    h = 0.08
    G = O.grid_points(origin, h, res)
    f = (np.linalg.norm(G - [0.1, 0.0, -0.05], axis=1) + 0.2 * np.sin(3 * G[:, 0]) * np.cos(2 * G[:, 1])).astype(np.float32)
    grid = gpu.make_grid(origin, h, res)
    for iso in [0.5, 0.75]:
        tris, cent, nrm = gpu.marching_cubes(ctx, torch.from_numpy(f), grid, iso, tris=True)
        G32 = _grid_points_f32(grid.origin, grid.h, res).astype(np.float64)
        rt, rc, rn = O.marching_cubes(f.astype(np.float64), res, np.asarray(grid.origin, np.float64), float(np.float32(h)), iso)
        order = to_brick_order(f, res, np.float32(iso))
        rt, rc, rn = rt[order], rc[order], rn[order]
        assert len(rt) == tris.shape[0] > 0
        assert np.abs(tris.cpu().numpy() - rt).max() < 2e-5
        assert np.abs(cent.cpu().numpy() - rc).max() < 2e-5
        assert np.abs(nrm.cpu().numpy() - rn).max() < 2e-3
        assert G32.shape[0] == f.shape[0]


def _rows_sorted(c, n):
    """(centroid, normal) rows as one array, sorted row by row on their bit patterns."""
    rows = np.concatenate([c, n], axis=1).astype(np.float32).view(np.uint32)
    return rows[np.lexsort(rows.T[::-1])]


@pytest.mark.parametrize("res,iso", [((30, 27, 25), 0.5), ((70, 41, 37), 0.75), ((9, 9, 40), 0.5)])
def test_marching_cubes_any_order_same_triangles(ctx, res, iso):
    # the one-pass variant writes the ordered passes' triangles bit for bit,
    # in another order; the grids cross brick edges in x, y and z (bricks of
    # 8 x 8 x 16 cubes), a first call with a short capacity reports the count
    res = np.array(res, np.int32)
    h = 2.2 / float(res.max())
    origin = np.array([0.1, 0.0, -0.05]) - 0.5 * h * (res - 1)  # centred on the surface below
    G = O.grid_points(origin, h, res)
    f = (np.linalg.norm(G - [0.1, 0.0, -0.05], axis=1) + 0.2 * np.sin(3 * G[:, 0]) * np.cos(2 * G[:, 1])).astype(np.float32)
    grid = gpu.make_grid(origin, h, res)
    ft = torch.from_numpy(f).cuda()
    _, c0, n0 = gpu.marching_cubes(ctx, ft, grid, iso)
    T = c0.shape[0]
    assert T > 100
    c1, n1 = gpu.marching_cubes_any_order(ctx, ft, grid, iso, capacity=T // 3)
    assert c1.shape[0] == T
    assert np.array_equal(_rows_sorted(c0.cpu().numpy(), n0.cpu().numpy()),
                          _rows_sorted(c1.cpu().numpy(), n1.cpu().numpy()))
    # exact capacity, then no triangles at all
    c2, n2 = gpu.marching_cubes_any_order(ctx, ft, grid, iso, capacity=T)
    assert np.array_equal(_rows_sorted(c0.cpu().numpy(), n0.cpu().numpy()),
                          _rows_sorted(c2.cpu().numpy(), n2.cpu().numpy()))
    c3, _ = gpu.marching_cubes_any_order(ctx, ft, grid, float(f.max()) + 1.0, capacity=16)
    assert c3.shape[0] == 0


def test_radii_and_transfer(ctx):
    P, _, _ = _sphere(20000, 16)
    P = P * 0.5
    idx = gpu.KnnIndex(ctx, P, k_hint=10)
    r = gpu.radii(ctx, idx, k=10).cpu().numpy()
    assert np.allclose(r, O.radii(P, 10), rtol=1e-6, atol=1e-9)
    src_n = np.random.default_rng(17).normal(size=P.shape).astype(np.float32)
    Q = (P[:3000] * 1.01).astype(np.float32)
    got = gpu.transfer_normals(ctx, idx, src_n, Q, k=10).cpu().numpy()
    nb = O.knn_brute(P, Q, 10)
    m = src_n[nb].mean(1)
    want = m / (np.linalg.norm(m, axis=1, keepdims=True) + 1e-8)
    assert np.abs(got - want).max() < 1e-5


def test_diffuse_normals_matches_restatement(ctx):
    P, _, _ = _sphere(5000, 18)
    P = P * 0.5
    rng = np.random.default_rng(19)
    T = 20000
    cent = (P[rng.integers(0, len(P), T)] + rng.normal(0, 0.01, (T, 3))).astype(np.float32)
    tn = rng.normal(size=(T, 3)).astype(np.float32)
    tn /= np.linalg.norm(tn, axis=1, keepdims=True)
    n0 = rng.normal(size=P.shape).astype(np.float32)
    idx = gpu.KnnIndex(ctx, P, k_hint=15)
    nrm = torch.from_numpy(n0.copy()).cuda()
    nz = gpu.diffuse_normals(ctx, idx, cent, tn, nrm)
    got = nrm.cpu().numpy()
    nb = O.knn_brute(P, cent, 15)
    acc = np.zeros((len(P), 3))
    np.add.at(acc, nb.ravel(), np.repeat(tn.astype(np.float64), 15, axis=0))
    l = np.linalg.norm(acc, axis=1)
    zero = l < 1e-10
    want = np.where(zero[:, None], n0, acc / (l[:, None] + 1e-8))
    assert nz == zero.sum()
    if nz > 10:
        zb = O.knn_brute(P, P, 10, skip=zero.astype(np.uint8))
        s = want[zb].sum(1)
        fill = s / (np.linalg.norm(s, axis=1, keepdims=True) + 1e-8)
        want = np.where(zero[:, None], fill, want)
    assert np.abs(got - want).max() < 1e-5
    # bit-reproducible
    nrm2 = torch.from_numpy(n0.copy()).cuda()
    gpu.diffuse_normals(ctx, idx, cent, tn, nrm2)
    assert torch.equal(nrm, nrm2)


def test_small_reductions(ctx):
    v = np.random.default_rng(20).normal(size=1_000_003).astype(np.float32)
    assert ctx.reduce_sum(v) == pytest.approx(v.astype(np.float64).sum(), abs=1e-6)
    P, _, _ = _sphere(1000, 21)
    lo, hi = ctx.bounds(P)
    assert np.array_equal(lo, P.min(0)) and np.array_equal(hi, P.max(0))
    c = (lo + hi) / 2
    g = gpu.gauss_map(ctx, P, c).cpu().numpy()
    assert np.allclose(g, O.gauss_map(P.astype(float)), atol=1e-6)


def test_pca_normals_match_oracle(ctx):
    rng = np.random.default_rng(63)
    P, _ = shapes.torus(20_000, 64)
    P = (P + rng.normal(0, 0.002, P.shape)).astype(np.float32)
    lo, hi = P.min(0), P.max(0)
    vp = (lo + hi) / 2
    idx = gpu.KnnIndex(ctx, P, k_hint=10)
    got = gpu.pca_normals(ctx, idx, 10, vp).cpu().numpy().astype(np.float64)
    ref = O.pca_normals(P, 10, vp)
    assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-6)
    c = (got * ref).sum(1)
    # same line (FP64 on both sides; only near-degenerate covariances may differ)
    assert np.quantile(np.abs(c), 0.001) > 0.9999
    # same side of the viewpoint wherever the side is well defined
    t = vp - P
    sure = np.abs((ref * t).sum(1)) > 1e-3 * np.linalg.norm(t, axis=1)
    assert np.all(c[sure] > 0)


def test_point_areas_match_oracle(ctx):
    # noisy sphere with 10% of the normals flipped (the filter drops them),
    # against the oracle's Qhull-triangulated restatement
    rng = np.random.default_rng(71)
    P, N = shapes.sphere(6000, 72)
    P = (0.5 * P + rng.normal(0, 0.001, P.shape)).astype(np.float32)
    N = N.astype(np.float32)
    N[rng.uniform(size=len(N)) < 0.1] *= -1
    idx = gpu.KnnIndex(ctx, P, k_hint=15)
    got = gpu.point_areas(ctx, idx, 15, N).cpu().numpy().astype(np.float64)
    nb = O.knn_brute(P, P, 15)
    ref = O.point_areas(P.astype(np.float64), nb, N.astype(np.float64))
    close = np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-12
    assert close.mean() > 0.999, (close.mean(), np.abs(got - ref).max())
    assert np.all(got >= 0)
    # without normals every neighbour is kept
    a = gpu.point_areas(ctx, idx, 15).cpu().numpy().astype(np.float64)
    ref = O.point_areas(P.astype(np.float64), nb, None)
    assert (np.abs(a - ref) <= 1e-5 * ref + 1e-12).mean() > 0.999
    # a quasi-uniform (Fibonacci) sphere has every Delaunay neighbour among the
    # 15 nearest, so the cells tile the sphere: mean area = 4πr²/n
    n = 6000
    t = np.arange(n) + 0.5
    phi, th = np.arccos(1 - 2 * t / n), np.pi * (1 + 5 ** 0.5) * t
    F = (0.5 * np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)).astype(np.float32)
    a = gpu.point_areas(ctx, gpu.KnnIndex(ctx, F, k_hint=15), 15).cpu().numpy()
    assert abs(a.mean() / (np.pi / n) - 1) < 0.002


def test_point_areas_hex_lattice(ctx):
    h = 0.01
    ij = np.array([(i, j) for i in range(-30, 31) for j in range(-30, 31)], np.float64)
    xy = np.stack([h * (ij[:, 0] + 0.5 * ij[:, 1]), h * (np.sqrt(3) / 2) * ij[:, 1]], 1)
    R, _ = np.linalg.qr(np.random.default_rng(73).normal(size=(3, 3)))
    P = (np.concatenate([xy, np.zeros((len(xy), 1))], 1) @ R.T).astype(np.float32)
    idx = gpu.KnnIndex(ctx, P, k_hint=15)
    a = gpu.point_areas(ctx, idx, 15).cpu().numpy()
    inner = np.abs(ij).max(1) <= 25
    assert np.allclose(a[inner], np.sqrt(3) / 2 * h * h, rtol=1e-4)
