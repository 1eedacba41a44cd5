"""Writes tests/golden/dwg_golden.npz: small seeded inputs and the FP64
oracle's outputs for every function on the DWG path (winding-number sums and
gradient, Barnes-Hut term counts and values, kNN, voxel grid, marching
cubes, radii, Gauss map, the DWG loop, the MLP forward/loss/gradient).

The reference ships no tests or fixtures, so these vectors come from the
oracle, which the CPU tests pin against closed forms and invariants; the
GPU tests compare the HIP path with these files.  Regenerate with

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from oracle import mlp as M  # noqa: E402


def sphere(n, rng, noise=0.0):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return (v * 0.5 + rng.normal(0, noise, (n, 3))).astype(np.float32), v.astype(np.float32)


def main(path=os.path.join(HERE, "dwg_golden.npz")):
    rng = np.random.default_rng(20260)
    g = {}
    # winding-number sums
    P, N = sphere(300, rng, 0.002)
    A = rng.uniform(1e-3, 3e-3, 300).astype(np.float32)
    Q = np.concatenate([rng.uniform(-0.8, 0.8, (100, 3)), P[:20] + rng.normal(0, 0.01, (20, 3))]).astype(np.float32)
    P64, N64, A64, Q64 = (x.astype(np.float64) for x in (P, N, A, Q))
    g.update(gwn_P=P, gwn_N=N, gwn_A=A, gwn_Q=Q)
    g["gwn_w_plain"] = O.direct(P64, N64, A64, Q64, kind="plain")
    g["gwn_w_screened"] = O.direct(P64, N64, A64, Q64, kind="screened", sigma=10.0)
    w, grad = O.direct_grad(P64, N64, A64, Q64)
    g["gwn_grad_w"], g["gwn_grad"] = w, grad
    # Barnes-Hut (octreeDWG, leaf size 1, depth 10, β = 2.3)
    t = O.Octree(P64, 10, 1)
    t.precompute(N64, A64)
    g["tree_nodes"] = np.int64(t.num_nodes)
    wp, terms = t.eval(Q64, 2.3, kind="plain", count_terms=True)
    g["tree_terms"], g["tree_w_plain"] = terms, wp
    g["tree_w_screened"] = t.eval(Q64, 2.3, kind="screened", sigma=20.0)
    # kNN
    P2 = (rng.normal(size=(2000, 3)) * [0.3, 0.2, 0.1]).astype(np.float32)
    Q2 = np.concatenate([P2[:300], rng.uniform(-0.6, 0.6, (400, 3))]).astype(np.float32)
    idx, d2 = O.knn_brute(P2, Q2, 10, return_d2=True)
    g.update(knn_P=P2, knn_Q=Q2, knn_idx=idx, knn_d2=d2)
    g["radii_k10"] = O.radii(P2.astype(np.float64), 10)
    # voxel grid and marching cubes
    origin, h, res = O.voxel_grid(P64, 0.05, 32)
    g.update(voxel_origin=np.asarray(origin, np.float64), voxel_h=np.float64(h), voxel_res=np.asarray(res, np.int32))
    mres = np.array([20, 18, 16], np.int32)
    morigin = np.array([-1.0, -0.9, -0.8])
    mh = 0.1
    G = O.grid_points(morigin, mh, mres)
    field = (np.linalg.norm(G - [0.05, -0.02, 0.03], axis=1) + 0.15 * np.sin(4 * G[:, 0])).astype(np.float32)
    tris, cent, nrm = O.marching_cubes(field.astype(np.float64), mres, morigin, mh, 0.6)
    g.update(mc_field=field, mc_res=mres, mc_origin=morigin, mc_h=np.float64(mh), mc_iso=np.float64(0.6),
             mc_tris=tris, mc_centroids=cent, mc_normals=nrm)
    # Gauss map and the DWG loop over a fixed source set
    g["gauss_P"] = P
    g["gauss_N"] = O.gauss_map(P64)
    P3, _ = sphere(800, rng, 0.001)
    P3_64 = P3.astype(np.float64)
    lo, hl, lres = O.voxel_grid(P3_64, 0.05, 32)
    n3, wm, nt, nz = O.dwg_iterate(P3_64, O.gauss_map(P3_64), np.full(len(P3), 2e-6), lres, lo, hl, 4, sigma=10.0)
    g.update(loop_P=P3, loop_origin=np.asarray(lo), loop_h=np.float64(hl), loop_res=np.asarray(lres, np.int32),
             loop_iters=np.int64(4), loop_w_mean=wm, loop_tris=nt, loop_zeros=nz, loop_normals=n3)
    # MLP 48→128→128→3 (GELU, tanh form)
    theta = np.concatenate([rng.uniform(-1, 1, int(np.prod(s))) / np.sqrt(s[1] if len(s) == 2 else 64)
                            for s in M.SIZES]).astype(np.float32)
    X = M.to_bf16(rng.normal(size=(64, 48)).astype(np.float32))
    T = rng.normal(size=(64, 3))
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    loss, grad = M.loss_grad(theta.astype(np.float64), X.astype(np.float64), T)
    # the forward pass as BF16 kernels compute it: BF16 weights and
    # activations, FP32-or-better accumulation
    bf = lambda a: M.to_bf16(np.asarray(a, np.float32)).astype(np.float64)  # noqa: E731
    W1, b1, W2, b2, W3, b3 = M.unpack(theta.astype(np.float64))
    h1 = bf(M.gelu(bf(X) @ bf(W1).T + b1))
    h2 = bf(M.gelu(h1 @ bf(W2).T + b2))
    g["mlp_y_bf16"] = h2 @ bf(W3).T + b3
    g.update(mlp_theta=theta, mlp_X=X, mlp_T=T, mlp_y=M.forward(theta.astype(np.float64), X.astype(np.float64))[0],
             mlp_loss=np.float64(loss), mlp_grad=grad)
    np.savez_compressed(path, **g)
    print(f"wrote {path}: {len(g)} arrays, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
