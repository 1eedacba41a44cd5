"""The large configuration at full size: 5,000,000 points from a seeded union
of 128 primitives and 16 sphere shells 0.01 thick, about 5% removed in
0.05-radius holes, noise 0.001·diag (seed 5); octree-leaf sources, 1024³
grids.  Learned start + mixed precision (BF16 far field, Kahan, 12
iterations) against plain FP32 (Gauss-map start, 40 iterations), as
`bench.py --large` runs them.

Checked: the equal-error gate (mean normal angle within 0.5°, Chamfer within
1% of the plain run, both against the clean surface); the Barnes-Hut sums of
the learned run's final source normals at 2,000 source and 2,000 grid
queries against the FP64 oracle's walk over the same sources with the same
β (identical term counts; FP32 within 1e-5 of max|w|, BF16 far field within
2e-3); the point count and bits of the workload (pinned on the CPU too).
"""
import os

import numpy as np
import pytest

import oracle as O
from dwg_amd import gpu, metrics, pipeline, shapes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "dwg_amd", "weights", "dwg_mlp.bin")


def _quality(ctx, r, Ntrue, S):
    ang = metrics.normal_angle_deg(r.normals, Ntrue)
    return float(ang.mean()), metrics.chamfer(ctx, r.tris_input_frame(), S)


def test_large_config(ctx):
    P, Ntrue, S = shapes.large(5_000_000, 5, clean=True)
    assert len(P) == 4_996_750
    net = gpu.Mlp(ctx, path=WEIGHTS)
    base = dict(grid_reso=1024, final_reso=1024, beta=2.3, leaf_size=1, source_mode=0)
    learned = pipeline.reconstruct(ctx, P, mlp=net, total_iter=12, input_normal=4, kahan=True, bf16_far=True,
                                   want_source_normals=True, **base)
    plain = pipeline.reconstruct(ctx, P, total_iter=40, input_normal=3, **base)
    a_l, c_l = _quality(ctx, learned, Ntrue, S)
    a_p, c_p = _quality(ctx, plain, Ntrue, S)
    print(f"large: learned+mixed {learned.time_total_s:.2f} s (angle {a_l:.3f}, Chamfer {c_l:.4e}); plain "
          f"{plain.time_total_s:.2f} s (angle {a_p:.3f}, Chamfer {c_p:.4e}); {learned.n_sources} sources", flush=True)
    assert a_l <= a_p + 0.5 and c_l <= 1.01 * c_p
    # the sums of the learned run's final normals against the FP64 walk
    Pn = (P.astype(np.float64) - learned.center) / learned.scale
    L, _ = gpu.Octree(ctx, Pn, None, max_depth=10).leaves()
    assert len(L) == learned.n_sources
    Ln = learned.source_normals.astype(np.float64)
    A = np.full(len(L), 2e-6)
    t = gpu.Octree(ctx, L, A, max_depth=10)
    t.update(Ln.astype(np.float32), A.astype(np.float32))
    o = O.Octree(L, 10, 1)
    o.precompute(Ln, A)
    origin, h, res = O.voxel_grid(Pn.astype(np.float32).astype(np.float64), 0.05, 1024)
    rng = np.random.default_rng(13)
    gi = rng.integers(0, int(res[0]) * int(res[1]) * int(res[2]), 2000)
    G = np.stack([origin[0] + (gi % res[0]) * h, origin[1] + (gi // res[0] % res[1]) * h,
                  origin[2] + gi // (res[0] * res[1]) * h], -1)
    Q = np.concatenate([L[rng.choice(len(L), 2000, replace=False)], G]).astype(np.float32)
    ref, terms = o.eval(Q.astype(np.float64), 2.3, kind="screened", sigma=10.0, count_terms=True)
    assert np.array_equal(t.terms(Q, 2.3).cpu().numpy(), terms)
    scale = np.abs(ref).max()
    e32 = np.abs(t.eval(Q, 2.3, sigma=10.0).cpu().numpy() - ref).max() / scale
    e16 = np.abs(t.eval(Q, 2.3, sigma=10.0, kahan=True, bf16=True).cpu().numpy() - ref).max() / scale
    print(f"large: sums vs FP64 walk at {len(Q)} queries: FP32 {e32:.2e}, BF16 far field {e16:.2e} of max|w|")
    assert e32 <= 1e-5 and e16 <= 2e-3
