"""Exact kNN at sizes that take the tiled (LDS-staged) query path (≥ 4096
queries), bit-exact against the oracle's brute-force ranking by
(float d², index): k up to 32, queries outside the point bounds, a dense
cluster that overflows one LDS stage, exact duplicates (ties), a skipped
region large enough to send queries to the per-lane fallback, and a query
mask; a sparse query set (fewer than one query per 8 points) that skips
the tiles for the pyramid walk; and far queries of a dense query set, which
share one pyramid walk per wave.
"""
import numpy as np
import pytest
import torch

import oracle as O
from dwg_amd import gpu, shapes

pytestmark = pytest.mark.gpu


def _noisy_sphere(n, seed, sigma=0.002):
    rng = np.random.default_rng(seed)
    P, _ = shapes.sphere(n, seed)
    return (0.5 * P + rng.normal(0, sigma, P.shape)).astype(np.float32)


@pytest.mark.parametrize("k", [4, 10, 15, 17, 32])
def test_knn_tiled_bit_exact(ctx, k):
    P = _noisy_sphere(30_000, 20 + k)
    rng = np.random.default_rng(k)
    Q = np.concatenate([P[rng.integers(0, len(P), 5000)] + rng.normal(0, 0.004, (5000, 3)),
                        rng.uniform(-0.6, 0.6, (1500, 3)),      # inside the box, off the surface
                        rng.uniform(-1.5, 1.5, (500, 3))]).astype(np.float32)  # partly outside the box
    idx = gpu.KnnIndex(ctx, P, k_hint=k)
    got, d2 = idx.query(Q, k, d2=True)
    ref, rd2 = O.knn_brute(P, Q, k, return_d2=True)
    assert np.array_equal(got.cpu().numpy(), ref)
    assert np.array_equal(d2.cpu().numpy(), rd2)


def test_knn_tiled_dense_cluster_and_duplicates(ctx):
    rng = np.random.default_rng(41)
    P = _noisy_sphere(20_000, 40)
    blob = (rng.normal(0, 0.002, (6000, 3)) + [0.5, 0.0, 0.0]).astype(np.float32)  # thousands per cell
    P = np.concatenate([P, blob, P[:800]]).astype(np.float32)  # exact duplicates tie on d²
    Q = np.concatenate([P[::3], blob[:2000] + rng.normal(0, 0.001, (2000, 3)).astype(np.float32)]).astype(np.float32)
    idx = gpu.KnnIndex(ctx, P, k_hint=15)
    got, d2 = idx.query(Q, 15, d2=True)
    ref, rd2 = O.knn_brute(P, Q, 15, return_d2=True)
    assert np.array_equal(got.cpu().numpy(), ref)
    assert np.array_equal(d2.cpu().numpy(), rd2)


def test_knn_tiled_skip_region_and_mask(ctx):
    P = _noisy_sphere(25_000, 42)
    rng = np.random.default_rng(43)
    # a whole cap is skipped (as zero normals cluster): its queries must
    # reach far beyond the staged tiles
    skip = (P[:, 0] > 0.3).astype(np.uint8)
    skip |= (rng.uniform(size=len(P)) < 0.1).astype(np.uint8)
    qmask = (rng.uniform(size=len(P)) < 0.7).astype(np.uint8)
    idx = gpu.KnnIndex(ctx, P, k_hint=10)
    got = idx.query(P, 10, skip=torch.from_numpy(skip), qmask=torch.from_numpy(qmask)).cpu().numpy()
    ref = O.knn_brute(P, P, 10, skip=skip)
    sel = qmask == 1
    assert np.array_equal(got[sel], ref[sel])
    assert np.all(got[~sel] == -1)


def test_knn_far_queries(ctx):
    # two small dense spheres far apart: queries in the empty space between
    # them start many coarse cells away from any point and close only after
    # wide shells (the wave fallback)
    rng = np.random.default_rng(45)
    S, _ = shapes.sphere(20_000, 46)
    P = np.concatenate([0.05 * S[:10_000] + [-0.8, 0, 0], 0.05 * S[10_000:] + [0.8, 0.3, -0.2]]).astype(np.float32)
    Q = np.concatenate([rng.uniform(-0.85, 0.85, (6000, 3)),
                        rng.normal(0, 0.01, (2000, 3)) + [0.0, 0.1, 0.0]]).astype(np.float32)
    for k in (10, 15):
        idx = gpu.KnnIndex(ctx, P, k_hint=k)
        ctx.profile_reset()
        ctx.profile(True)
        got, d2 = idx.query(Q, k, d2=True)
        ctx.profile(False)
        # enough far queries to take the pyramid walk (≥ 2048 after the tiles)
        assert ctx.counter("knn_query:tiles3:redo") >= 2048
        ref, rd2 = O.knn_brute(P, Q, k, return_d2=True)
        assert np.array_equal(got.cpu().numpy(), ref)
        assert np.array_equal(d2.cpu().numpy(), rd2)


@pytest.mark.parametrize("k", [10, 16])
def test_knn_sparse_queries(ctx, k):
    # 6,000 queries over 60,000 points, as the zero fill queries the non-zero
    # sources: near and far queries and a skipped cap, one wave per query
    P = _noisy_sphere(60_000, 50 + k)
    rng = np.random.default_rng(60 + k)
    Q = np.concatenate([P[rng.integers(0, len(P), 3000)] + rng.normal(0, 0.003, (3000, 3)),
                        rng.uniform(-0.7, 0.7, (2500, 3)),
                        rng.uniform(-1.5, 1.5, (500, 3))]).astype(np.float32)
    skip = (P[:, 2] > 0.35).astype(np.uint8)
    idx = gpu.KnnIndex(ctx, P, k_hint=k)
    ctx.profile_reset()
    ctx.profile(True)
    got, d2 = idx.query(Q, k, d2=True, skip=torch.from_numpy(skip))
    ctx.profile(False)
    assert ctx.counter("knn_query:sparse") == len(Q)
    ref, rd2 = O.knn_brute(P, Q, k, skip=skip, return_d2=True)
    assert np.array_equal(got.cpu().numpy(), ref)
    assert np.array_equal(d2.cpu().numpy(), rd2)


@pytest.mark.parametrize("k", [10, 15, 16])
def test_knn_far_queries_wave_shared(ctx, k):
    # as many queries as points, most of them on a dense sheet far from every
    # point (a spurious surface piece's centroids): the far ones share one
    # pyramid walk per wave
    rng = np.random.default_rng(70 + k)
    S, _ = shapes.sphere(20_000, 71)
    P = np.concatenate([0.05 * S[:10_000] + [-0.8, 0, 0], 0.05 * S[10_000:] + [0.8, 0.3, -0.2]]).astype(np.float32)
    r = 0.3 * np.sqrt(rng.uniform(size=20_000))
    phi = rng.uniform(0, 2 * np.pi, 20_000)
    sheet = np.stack([rng.normal(0, 0.002, 20_000), r * np.cos(phi), r * np.sin(phi)], 1)
    Q = np.concatenate([sheet, P[rng.integers(0, len(P), 4000)] + rng.normal(0, 0.003, (4000, 3))]).astype(np.float32)
    idx = gpu.KnnIndex(ctx, P, k_hint=k)
    ctx.profile_reset()
    ctx.profile(True)
    got, d2 = idx.query(Q, k, d2=True)
    ctx.profile(False)
    assert ctx.counter("knn_query:far_shared") >= 2048
    ref, rd2 = O.knn_brute(P, Q, k, return_d2=True)
    assert np.array_equal(got.cpu().numpy(), ref)
    assert np.array_equal(d2.cpu().numpy(), rd2)


@pytest.mark.parametrize("k", [10, 15])
def test_knn_denormal_and_zero_distances(ctx, k):
    # a cluster near the origin with points 1e-21 apart (and duplicates):
    # the squared distances are float denormals or 0, so the keys read as
    # doubles are denormals, which the candidate lists order with FP64
    # max/min; both the path of dense query sets (as many queries as points)
    # and that of fewer queries
    rng = np.random.default_rng(81 + k)
    S = _noisy_sphere(20_000, 82)
    C = np.zeros((300, 3), np.float32)
    C[:, 0] = (np.arange(300) * 1e-21).astype(np.float32)
    C = np.concatenate([C, C[:50]])
    P = np.concatenate([S, C]).astype(np.float32)
    Qc = C[rng.integers(0, len(C), 16_000)].copy()
    Qc[:, 1] += (rng.uniform(-3e-22, 3e-22, len(Qc))).astype(np.float32)
    Qs = S[rng.integers(0, len(S), 6000)] + rng.normal(0, 0.002, (6000, 3)).astype(np.float32)
    idx = gpu.KnnIndex(ctx, P, k_hint=k)
    for Q in (np.concatenate([Qc, Qs]).astype(np.float32), np.concatenate([Qc[:4000], Qs[:1000]]).astype(np.float32)):
        got, d2 = idx.query(Q, k, d2=True)
        ref, rd2 = O.knn_brute(P, Q, k, return_d2=True)
        assert np.array_equal(got.cpu().numpy(), ref)
        assert np.array_equal(d2.cpu().numpy(), rd2)
        tiny = rd2[: len(Q) - len(Q) // 5]
        assert np.mean((tiny > 0) & (tiny < np.finfo(np.float32).tiny)) > 0.5
        assert np.any(tiny == 0)
