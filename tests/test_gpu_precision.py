"""Precision sweep on the headline workload (1,000,000 points, union of 32
primitives, seed 3, noise 0.001·diag, normalised as ReadPointandNormals
does; octree-leaf sources, areas 2e-6, β = 2.3, σ = 10): FP32, FP32 +
Kahan and BF16-far-field tree sums, and the direct sums, against the FP64
oracle at 2,000 source and 2,000 grid queries of the 512³ iteration grid.

The FP64 references are the oracle's tree walk with the same β (the same
term set, so the difference is arithmetic only) and its all-pairs FP64 sum
(fast_winding_number.cpp:313-323; the screened term of TurboLibiGWN.cu:
544-585), which adds the Barnes-Hut truncation.  Errors are relative to
max|w| over the queries.  Bounds (measured values in the printed lines):
  FP32 tree vs FP64 tree          ≤ 1e-5   (FP32 rounding; measured 2.4e-6 max, 2.2e-7 mean)
  FP32 + Kahan tree vs FP64 tree  ≤ 1e-5   (measured 2.4e-6 max)
  BF16 far field + Kahan          ≤ 2e-3   (BF16 moments: 2^-9 relative; measured 6.9e-4 max, 1.0e-4 mean)
  FP32 direct vs FP64 direct      ≤ 1e-5   (measured 2.0e-6 max)
  term sets (FP32 walk)           exact
and with DWG_NEAR_FP64 (the terms inside the screening radius in FP64) the
same bounds (measured: no difference at source queries of this workload).
"""
import numpy as np
import pytest
import torch

import oracle as O
from dwg_amd import gpu, shapes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload(ctx):
    P, N = shapes.union(1_000_000, 32, 3)
    c = P.astype(np.float64).mean(0)
    s = 2 * np.abs(P.astype(np.float64) - c).max()
    Pn = (P.astype(np.float64) - c) / s
    pts = gpu.Octree(ctx, Pn, None, max_depth=10)
    L, _ = pts.leaves()
    Ln = pts.leaf_normals(Pn, N.astype(np.float64)).astype(np.float64)
    Ln /= np.linalg.norm(Ln, axis=1, keepdims=True)
    A = np.full(len(L), 2e-6)
    origin, h, res = O.voxel_grid(Pn.astype(np.float32).astype(np.float64), 0.05, 512)
    rng = np.random.default_rng(11)
    gi = rng.integers(0, int(res[0]) * int(res[1]) * int(res[2]), 2000)
    G = np.stack([origin[0] + (gi % res[0]) * h, origin[1] + (gi // res[0] % res[1]) * h,
                  origin[2] + gi // (res[0] * res[1]) * h], -1)
    Q = np.concatenate([L[rng.choice(len(L), 2000, replace=False)], G]).astype(np.float32).astype(np.float64)
    print(f"precision workload: {len(Pn)} points, {len(L)} leaf sources, grid {tuple(res)}", flush=True)
    return L, Ln, A, Q


def _err(w, ref):
    scale = np.abs(ref).max()
    d = np.abs(np.asarray(w, np.float64) - ref)
    return d.max() / scale, d.mean() / scale


def test_precision_sweep_headline_workload(ctx, workload):
    L, Ln, A, Q = workload
    t = gpu.Octree(ctx, L, A, max_depth=10)
    t.update(Ln.astype(np.float32), A.astype(np.float32))
    o = O.Octree(L, 10, 1)
    o.precompute(Ln, A)
    ref_tree, ref_terms = o.eval(Q, 2.3, kind="screened", sigma=10.0, count_terms=True)
    assert np.array_equal(t.terms(Q.astype(np.float32), 2.3).cpu().numpy(), ref_terms)
    rows = {}
    for name, kw in [("fp32", dict()), ("fp32+kahan", dict(kahan=True)), ("bf16", dict(bf16=True)),
                     ("bf16+kahan", dict(kahan=True, bf16=True))]:
        w = t.eval(Q.astype(np.float32), 2.3, sigma=10.0, **kw).cpu().numpy()
        rows[name] = _err(w, ref_tree)
    # all-pairs sums: every query against every source, FP64 on the CPU
    sub = np.r_[0:250, 2000:2250]
    ref_direct = O.direct(L, Ln, A, Q[sub], kind="screened", sigma=10.0)
    wd = ctx.gwn_direct(L.astype(np.float32), Ln.astype(np.float32), A.astype(np.float32),
                        Q[sub].astype(np.float32), sigma=10.0, screened=True, kahan=True).cpu().numpy()
    rows["direct fp32+kahan"] = _err(wd, ref_direct)
    rows["fp64 tree vs fp64 direct (β = 2.3 truncation)"] = _err(ref_tree[sub], ref_direct)
    wt = t.eval(Q[sub].astype(np.float32), 2.3, sigma=10.0, kahan=True, bf16=True).cpu().numpy()
    rows["bf16+kahan tree vs fp64 direct"] = _err(wt, ref_direct)
    for k, (mx, mean) in rows.items():
        print(f"precision {k:48s} max {mx:.3e}  mean {mean:.3e}  (of max|w| = {np.abs(ref_tree).max():.4f})")
    assert rows["fp32"][0] <= 1e-5
    assert rows["fp32+kahan"][0] <= 1e-5
    assert rows["bf16+kahan"][0] <= 2e-3 and rows["bf16"][0] <= 2e-3
    assert rows["direct fp32+kahan"][0] <= 1e-5
    # the BF16 error stays well under the method's own Barnes-Hut truncation
    assert rows["bf16+kahan"][1] < 0.5 * rows["fp64 tree vs fp64 direct (β = 2.3 truncation)"][1]


def test_near_field_fp64(ctx, workload):
    # DWG_NEAR_FP64: the point terms inside a source's screening radius in
    # FP64.  On the headline workload it changes nothing measurable (at the
    # 2,000 source queries the error against the FP64 walk stays that of
    # FP32: the inputs' own FP32 rounding and the far field dominate); on a
    # dense cluster of sources all inside each other's screening radius,
    # whose large terms cancel, the FP64 sum is the more accurate one
    L, Ln, A, Q = workload
    t = gpu.Octree(ctx, L, A, max_depth=10)
    t.update(Ln.astype(np.float32), A.astype(np.float32))
    o = O.Octree(L, 10, 1)
    o.precompute(Ln, A)
    Qs = Q[:2000]
    ref = o.eval(Qs, 2.3, kind="screened", sigma=10.0)
    rows = {}
    for name, kw in [("fp32", {}), ("fp32+near64", {"near64": True}), ("bf16+kahan", {"kahan": True, "bf16": True}),
                     ("bf16+kahan+near64", {"kahan": True, "bf16": True, "near64": True})]:
        rows[name] = _err(t.eval(Qs.astype(np.float32), 2.3, sigma=10.0, **kw).cpu().numpy(), ref)
    rng = np.random.default_rng(17)
    u = rng.standard_normal((3000, 3))
    P = (0.001 * u / np.linalg.norm(u, axis=1, keepdims=True) * rng.random((3000, 1)) ** (1 / 3)).astype(np.float32)
    Nc = rng.standard_normal((3000, 3)).astype(np.float32)
    Nc /= np.linalg.norm(Nc, axis=1, keepdims=True)
    Ac = np.full(3000, 1e-3, np.float32)
    Qc = (0.0005 * rng.standard_normal((200, 3))).astype(np.float32)
    ref_c = O.direct(P.astype(np.float64), Nc.astype(np.float64), Ac.astype(np.float64), Qc.astype(np.float64),
                     kind="screened", sigma=10.0)
    for name, near in [("cluster fp32+kahan", False), ("cluster fp32+kahan+near64", True)]:
        w = ctx.gwn_direct(P, Nc, Ac, Qc, sigma=10.0, screened=True, kahan=True, near64=near).cpu().numpy()
        rows[name] = _err(w, ref_c)
    for k, (mx, mean) in rows.items():
        print(f"near-field {k:28s} max {mx:.3e}  mean {mean:.3e}")
    assert rows["fp32+near64"][0] <= 1e-5 and rows["fp32+near64"][1] <= 1.01 * rows["fp32"][1]
    assert rows["bf16+kahan+near64"][0] <= 2e-3
    assert rows["cluster fp32+kahan+near64"][1] <= rows["cluster fp32+kahan"][1]


def test_mixed_precision_loop_tracks_fp32(ctx):
    # the learned+mixed iterations' precision flags (BF16 far-field moments,
    # Kahan-compensated accumulation) against the same loop in FP32 on a
    # 200k-point union shape (octree-leaf sources, 256^3 grids, 8 iterations,
    # Gauss-map start so that only the precision differs): per-iteration mean
    # source winding number within 5e-3 relative (it is a small mean of values
    # of both signs; measured 1.8e-3), triangle counts within 1% (measured
    # 0.05%), and the final normals' error against the analytic normals
    # within 0.1° (measured 0.003°)
    from dwg_amd import metrics, pipeline
    P, N = shapes.union(200_000, 32, 7)
    kw = dict(grid_reso=256, final_reso=256, total_iter=8, input_normal=3, source_mode=0, want_mesh=False)
    fp32 = pipeline.reconstruct(ctx, P, **kw)
    mixed = pipeline.reconstruct(ctx, P, kahan=True, bf16_far=True, **kw)
    rel = np.abs(mixed.w_mean - fp32.w_mean) / np.abs(fp32.w_mean)
    a32 = metrics.normal_angle_deg(fp32.normals, N).mean()
    amx = metrics.normal_angle_deg(mixed.normals, N).mean()
    print(f"mixed vs fp32: max rel W_mean diff {rel.max():.2e}, triangles {mixed.iter_tris.tolist()} vs "
          f"{fp32.iter_tris.tolist()}, mean angle {amx:.3f} vs {a32:.3f} deg")
    assert rel.max() < 5e-3
    assert np.all(np.abs(mixed.iter_tris - fp32.iter_tris) <= 0.01 * fp32.iter_tris)
    assert abs(amx - a32) < 0.1
