"""The HIP path against the committed golden vectors (tests/golden), written
by the FP64 oracle.  Tolerances as in the kernel tests: FP32 sums
|Δw| ≤ 2e-5·max|w| + 2e-4·|w|; gradients 1e-4·max|∇w|; term counts, kNN
indices and distances, triangle counts exact; the DWG loop's mean source WN
1e-3 relative and its triangle counts 1%; the MLP forward within BF16
rounding of the BF16-emulated oracle."""
import os

import numpy as np
import pytest
import torch

from dwg_amd import gpu, pipeline
from mc_order import to_brick_order

pytestmark = pytest.mark.gpu

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "dwg_golden.npz"))


def _close(a, b, rel=2e-4):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err = np.abs(a - b)
    assert np.all(err <= 2e-5 * np.abs(b).max() + rel * np.abs(b)), f"max err {err.max()}"


def test_direct_sums(ctx):
    P, N, A, Q = G["gwn_P"], G["gwn_N"], G["gwn_A"], G["gwn_Q"]
    _close(ctx.gwn_direct(P, N, A, Q).cpu().numpy(), G["gwn_w_plain"])
    _close(ctx.gwn_direct(P, N, A, Q, sigma=10.0, screened=True, kahan=True).cpu().numpy(), G["gwn_w_screened"])
    w, g = ctx.gwn_direct(P, N, A, Q, grad=True)
    _close(w.cpu().numpy(), G["gwn_grad_w"])
    assert np.abs(g.cpu().numpy() - G["gwn_grad"]).max() <= 1e-4 * np.abs(G["gwn_grad"]).max()


def test_tree(ctx):
    P, N, A, Q = G["gwn_P"], G["gwn_N"], G["gwn_A"], G["gwn_Q"]
    t = gpu.Octree(ctx, P.astype(np.float64), A.astype(np.float64), max_depth=10, leaf_size=1)
    t.update(N, A)
    assert t.info()["nodes"] == int(G["tree_nodes"])
    assert np.array_equal(t.terms(Q, 2.3).cpu().numpy(), G["tree_terms"])
    _close(t.eval(Q, 2.3, screened=False).cpu().numpy(), G["tree_w_plain"])
    _close(t.eval(Q, 2.3, sigma=20.0, screened=True).cpu().numpy(), G["tree_w_screened"])


def test_knn_and_radii(ctx):
    idx = gpu.KnnIndex(ctx, G["knn_P"], k_hint=10)
    got, d2 = idx.query(G["knn_Q"], 10, d2=True)
    assert np.array_equal(got.cpu().numpy(), G["knn_idx"])
    assert np.array_equal(d2.cpu().numpy(), G["knn_d2"])
    r = gpu.radii(ctx, idx, k=10).cpu().numpy()
    assert np.allclose(r, G["radii_k10"], rtol=1e-6, atol=1e-9)


def test_marching_cubes(ctx):
    grid = gpu.make_grid(G["mc_origin"], float(G["mc_h"]), G["mc_res"])
    tris, cent, nrm = gpu.marching_cubes(ctx, torch.from_numpy(G["mc_field"]), grid, float(G["mc_iso"]), tris=True)
    order = to_brick_order(G["mc_field"], G["mc_res"], np.float32(G["mc_iso"]))
    assert tris.shape[0] == len(G["mc_tris"])
    assert np.abs(tris.cpu().numpy() - G["mc_tris"][order]).max() < 2e-5
    assert np.abs(cent.cpu().numpy() - G["mc_centroids"][order]).max() < 2e-5
    assert np.abs(nrm.cpu().numpy() - G["mc_normals"][order]).max() < 2e-3


def test_gauss_map(ctx):
    P = G["gauss_P"]
    lo, hi = ctx.bounds(P)
    g = gpu.gauss_map(ctx, P, (lo + hi) / 2).cpu().numpy()
    assert np.allclose(g, G["gauss_N"], atol=1e-6)


def test_loop(ctx):
    P = G["loop_P"]
    it = int(G["loop_iters"])
    r = pipeline.reconstruct(ctx, P, want_mesh=False, source_mode=1, normalize_input=0, grid_reso=32, final_reso=32,
                             total_iter=it, input_normal=3, beta=2.3, leaf_size=1)
    assert r.grid_points == int(np.prod(G["loop_res"]))
    assert np.allclose(r.w_mean, G["loop_w_mean"], rtol=1e-3)
    assert np.all(np.abs(r.iter_tris - G["loop_tris"]) <= 0.01 * G["loop_tris"])
    cos = (r.normals * G["loop_normals"]).sum(1) / np.linalg.norm(r.normals, axis=1)
    assert np.median(cos) > 0.999


def test_mlp_forward(ctx):
    net = gpu.Mlp(ctx, 0)
    net.set_params(G["mlp_theta"])
    from oracle import mlp as M
    X = torch.from_numpy(M.bf16_bits(G["mlp_X"]).view(np.int16)).cuda()
    y = net.forward(X, normalize=False).cpu().numpy()
    ref = G["mlp_y_bf16"]
    assert np.abs(y - ref).max() < 2e-2 * max(1.0, np.abs(ref).max())
