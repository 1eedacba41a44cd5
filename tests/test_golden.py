"""The oracle against its committed golden vectors (tests/golden): every
function of the path recomputed from the stored inputs must give the stored
outputs, so a change to the oracle cannot pass unnoticed; plus the
invariants the vectors themselves must satisfy."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import mlp as M

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "dwg_golden.npz"))


def test_sums_and_gradient():
    P, N, A, Q = (G[k].astype(np.float64) for k in ("gwn_P", "gwn_N", "gwn_A", "gwn_Q"))
    assert np.allclose(O.direct(P, N, A, Q, kind="plain"), G["gwn_w_plain"], rtol=1e-12, atol=1e-15)
    assert np.allclose(O.direct(P, N, A, Q, kind="screened", sigma=10.0), G["gwn_w_screened"], rtol=1e-12, atol=1e-15)
    w, g = O.direct_grad(P, N, A, Q)
    assert np.allclose(w, G["gwn_grad_w"], rtol=1e-12, atol=1e-15)
    assert np.allclose(g, G["gwn_grad"], rtol=1e-12, atol=1e-15)
    # the plain sum is the same with or without the gradient
    assert np.allclose(G["gwn_grad_w"], G["gwn_w_plain"], rtol=1e-12, atol=1e-15)


def test_tree():
    P, N, A, Q = (G[k].astype(np.float64) for k in ("gwn_P", "gwn_N", "gwn_A", "gwn_Q"))
    t = O.Octree(P, 10, 1)
    t.precompute(N, A)
    assert t.num_nodes == int(G["tree_nodes"])
    w, terms = t.eval(Q, 2.3, kind="plain", count_terms=True)
    assert np.array_equal(terms, G["tree_terms"])
    assert np.allclose(w, G["tree_w_plain"], rtol=1e-12, atol=1e-15)
    assert np.allclose(t.eval(Q, 2.3, kind="screened", sigma=20.0), G["tree_w_screened"], rtol=1e-12, atol=1e-15)
    # a Barnes-Hut sum never takes more terms than there are sources
    assert terms.max() <= len(P) and terms.min() > 0


def test_knn_radii_voxel():
    idx, d2 = O.knn_brute(G["knn_P"], G["knn_Q"], 10, return_d2=True)
    assert np.array_equal(idx, G["knn_idx"]) and np.array_equal(d2, G["knn_d2"])
    assert np.all(np.diff(G["knn_d2"], axis=1) >= 0)
    assert np.allclose(O.radii(G["knn_P"].astype(np.float64), 10), G["radii_k10"], rtol=1e-12)
    o, h, r = O.voxel_grid(G["gwn_P"].astype(np.float64), 0.05, 32)
    assert np.allclose(o, G["voxel_origin"], rtol=0, atol=1e-15) and h == G["voxel_h"]
    assert np.array_equal(np.asarray(r), G["voxel_res"]) and G["voxel_res"].max() == 32


def test_marching_cubes():
    t, c, n = O.marching_cubes(G["mc_field"].astype(np.float64), G["mc_res"], G["mc_origin"], float(G["mc_h"]),
                               float(G["mc_iso"]))
    assert np.array_equal(t, G["mc_tris"]) and np.array_equal(c, G["mc_centroids"])
    assert np.array_equal(n, G["mc_normals"])
    assert np.allclose(G["mc_tris"].mean(1), G["mc_centroids"], atol=1e-12)


def test_gauss_map_and_loop():
    assert np.array_equal(O.gauss_map(G["gauss_P"].astype(np.float64)), G["gauss_N"])
    P = G["loop_P"].astype(np.float64)
    n, wm, nt, nz = O.dwg_iterate(P, O.gauss_map(P), np.full(len(P), 2e-6), G["loop_res"], G["loop_origin"],
                                  float(G["loop_h"]), int(G["loop_iters"]), sigma=10.0)
    assert np.allclose(wm, G["loop_w_mean"], rtol=1e-12)
    assert np.array_equal(nt, G["loop_tris"]) and np.array_equal(nz, G["loop_zeros"])
    assert np.allclose(n, G["loop_normals"], rtol=0, atol=1e-12)


def test_mlp():
    th = G["mlp_theta"].astype(np.float64)
    X = G["mlp_X"].astype(np.float64)
    y, _ = M.forward(th, X)
    assert np.allclose(y, G["mlp_y"], rtol=1e-12, atol=1e-14)
    loss, grad = M.loss_grad(th, X, G["mlp_T"])
    assert loss == pytest.approx(float(G["mlp_loss"]), rel=1e-12)
    assert np.allclose(grad, G["mlp_grad"], rtol=1e-10, atol=1e-14)
    # BF16 path close to the exact one
    assert np.abs(G["mlp_y_bf16"] - G["mlp_y"]).max() < 0.05 * max(1.0, np.abs(G["mlp_y"]).max())
