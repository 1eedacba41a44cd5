"""DWG reconstruction benchmark on one MI355X.

Workload (the headline metric's configuration): 1,000,000 points sampled
area-uniformly from a seeded union of 32 random spheres / ellipsoids / tori
in the unit cube (seed 3), Gaussian noise 0.001·bbox-diagonal; 512³
iteration and final grids; octree-leaf sources, β = 2.3, σ = 10 (the
reference's defaults).  A step is one whole reconstruction (setup,
iterations, final mesh) through dwg_launch, inputs already in host memory.

Two methods on the same input, same protocol (warm-up, then timed steps):
  plain FP32     Gauss-map start, 40 iterations, FP32 sums (the reference's
                 method and defaults);
  learned+mixed  learned start (MLP normal line, Gauss-map side), fewer
                 iterations, BF16 far-field moments with Kahan-compensated
                 FP32 accumulation in the iterations, FP32 final mesh.
`value` is the learned+mixed wall time; the line also carries the plain
time, the speed-up and whether learned+mixed met the equal-error gate
(mean normal angle within 0.5°, Chamfer within 1% of plain FP32).

Also reported: the roofline of the dominant kernel (algorithmic FLOPs from
exact Barnes-Hut term counts × FLOPs per term over the kernel's HIP-event
time on the library's stream); the plain method at fewer iterations, down
to the fewest that still meets the gate, and the speed-up against that run;
the direct all-pairs sums' FLOP rate on the small sphere configuration; a
CPU baseline measured on a bounded sample with the FP64 CPU restatement, and
on the same sample the error of the GPU sums (FP32, FP32 + Kahan, BF16 far
field + Kahan) against FP64.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "recon wall time @1M pts: learned+mixed vs plain FP32, equal normal error"
WEIGHTS = os.path.join(ROOT, "dwg_amd", "weights", "dwg_mlp.bin")
# iterations of the learned+mixed method, fixed on validation clouds
# (seeds 4, 5, 6) as the fewest that met the equal-error gate on all three
LEARNED_ITERS = 12
DWG_KAHAN, DWG_FAR_BF16 = 4, 8
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 vector (packed) peak
FLOPS_PER_TERM = {"screened": 27, "plain": 20}
# HBM bytes per launch of the iteration grid kernel, from the rocprofv3
# FETCH_SIZE / WRITE_SIZE runs recorded in this summary (copy-calibrated)
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "gwn_tree_grid_screened_bf16_1M_512_traffic.txt")
# and of the marching-cubes passes on the same workload
MC_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "mc_1M_512_traffic.txt")
PEAK_HBM = 8000.0  # GB/s, MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(n, seed):
    from dwg_amd import shapes
    P, N = shapes.union(n, 32, seed)
    return P, N


def surface_sample(n, seed):
    from dwg_amd import shapes
    rng = np.random.default_rng(seed)
    prims = shapes.random_primitives(rng, 32)
    S, _ = shapes.union_of(prims, n, np.random.default_rng(seed + 10_000), 0.0)
    return S


def plain_options(args):
    return dict(grid_reso=args.grid, final_reso=args.grid, total_iter=args.iters, input_normal=3, beta=2.3,
                leaf_size=1, source_mode=0)


def learned_options(args):
    return dict(plain_options(args), total_iter=args.learned_iters, input_normal=4,
                precision=DWG_KAHAN | DWG_FAR_BF16)


def run_once(ctx, P, opts, profile=False, count_terms=False, mlp=None):
    from dwg_amd import pipeline
    if profile:
        ctx.profile_reset()
        ctx.profile(True)
    r = pipeline.reconstruct(ctx, P, count_terms=1 if count_terms else 0, mlp=mlp, **opts)
    if profile:
        ctx.profile(False)
    return r


KERNELS = ["gwn_tree_grid_screened", "gwn_tree_screened", "octree_pack", "octree_ec_scan", "octree_ec", "mc_count",
           "mc_emit", "mc_fused", "knn_query_centroids", "knn_query_zero_fill", "knn_query_radii", "knn_query_transfer",
           "knn_query_patches", "knn_build", "diffuse_nonzero_index", "diffuse_normalize", "diffuse_zero_fill", "radii",
           "gauss_map", "mlp_init", "mlp_forward"]


def large_config(args):
    """The large configuration, one step of each method after one warm-up
    run of each: 5M points from a seeded union of 128 primitives and 16
    sphere shells 0.01 thick, about 5% removed in 0.05-radius holes, noise
    0.001·diag (seed 5); 1024^3 grids; learned+mixed (learned start,
    LEARNED_ITERS iterations, BF16 far field + Kahan) against plain FP32 (40
    iterations); quality against the same points before the noise."""
    from dwg_amd import shapes
    t0 = time.perf_counter()
    P, Ntrue, S = shapes.large(5_000_000, 5, clean=True,
                               progress=lambda have, n: log(f"sampling: {have} of {n} ({time.perf_counter() - t0:.0f} s)"))
    log(f"large workload: {len(P)} points in {time.perf_counter() - t0:.0f} s")
    import torch
    from dwg_amd import gpu
    ctx = gpu.Context(0)
    net = gpu.Mlp(ctx, path=args.weights)
    a = argparse.Namespace(**dict(vars(args), grid=1024))
    plain, learned = plain_options(a), learned_options(a)
    run_once(ctx, P, dict(plain, total_iter=2))
    run_once(ctx, P, dict(learned, total_iter=2), mlp=net)
    plain_t, plain_r = timed(ctx, P, plain, 1, "plain")
    learned_t, learned_r = timed(ctx, P, learned, 1, "learned+mixed", mlp=net)
    q_plain, q_learned = quality(ctx, plain_r, Ntrue, S), quality(ctx, learned_r, Ntrue, S)
    passed = gate(q_learned, q_plain)
    print(json.dumps({"config": "large: 5M points, 128 primitives + 16 shells, 5% holes, 1024^3 grids",
                      "n_points": int(len(P)), "n_sources": int(learned_r.n_sources),
                      "learned_mixed_s": round(learned_t[0], 3), "plain_fp32_s": round(plain_t[0], 3),
                      "speedup_vs_plain": round(plain_t[0] / learned_t[0], 3), "equal_error": passed,
                      "learned_iterations": args.learned_iters, "quality": q_learned, "plain_quality": q_plain}),
          flush=True)


def direct_sums(ctx=None, reps=5):
    """FLOP rate of the direct all-pairs sums (plain and screened) for the
    small sphere configuration: 10,000 area-uniform sphere points (seed 1),
    queries at the 128^3 points of a grid over the bounding box padded by
    10%.  Per kernel: mean HIP-event time on the library's stream over
    `reps` launches after one untimed launch."""
    import torch
    from dwg_amd import gpu, shapes
    P, N = shapes.sphere(10_000, 1)
    A = np.full(len(P), 4 * np.pi / len(P), np.float32)
    lo, hi = P.min(0), P.max(0)
    pad = 0.1 * (hi - lo)
    axes = [np.linspace(lo[d] - pad[d], hi[d] + pad[d], 128, dtype=np.float32) for d in range(3)]
    Q = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    ctx = ctx or gpu.Context(0)
    Pt, Nt, At, Qt = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (P, N, A, Q))
    out = []
    for screened, flops in ((False, FLOPS_PER_TERM["plain"]), (True, FLOPS_PER_TERM["screened"])):
        name = "gwn_direct_screened" if screened else "gwn_direct_plain"
        ctx.gwn_direct(Pt, Nt, At, Qt, screened=screened, kahan=False)
        torch.cuda.synchronize()
        ctx.profile_reset()
        ctx.profile(True)
        for _ in range(reps):
            ctx.gwn_direct(Pt, Nt, At, Qt, screened=screened, kahan=False)
        ctx.profile(False)
        tot, cnt = ctx.profile_get(name)
        ms = tot / cnt
        pairs = len(P) * len(Q)
        tf = pairs * flops / (ms * 1e-3) / 1e12
        out.append({"kernel": f"gwn_direct_kernel<{'screened' if screened else 'plain'}>",
                    "workload": "10,000 sphere sources x 128^3 grid queries", "pairs": pairs,
                    "flops_per_pair": flops, "avg_launch_ms": round(ms, 3),
                    "tpairs_per_s": round(pairs / (ms * 1e-3) / 1e12, 3), "tflops": round(tf, 2),
                    "frac_of_157.3": round(tf / PEAK_FP32, 4), "frac_of_78.6_non_packed": round(tf / 78.6, 4)})
    return out


def measured_traffic(path=TRAFFIC_SUMMARY, key="traffic_bytes_per_launch"):
    """`<key>: bytes` from a committed counter summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def hbm_rates(ktimes):
    """HBM GB/s of the grid walk and the marching-cubes passes: the bytes per
    launch that the FETCH_SIZE/WRITE_SIZE counters gave on this workload
    (committed summaries) over this run's mean launch times (HIP events)."""
    out = {}
    # the iterations run the one-pass marching cubes (mc_fused); the ordered
    # count and emit passes run in the final pass (recorded as "<kernel>:final")
    for name, tname, path, key in [
            ("gwn_tree_grid_screened", "gwn_tree_grid_screened", TRAFFIC_SUMMARY, "traffic_bytes_per_launch"),
            ("mc_fused", "mc_fused", MC_TRAFFIC_SUMMARY, "mc_fused_bytes_per_launch"),
            ("mc_count", "mc_count:final", MC_TRAFFIC_SUMMARY, "mc_count_bytes_per_launch"),
            ("mc_emit", "mc_emit:final", MC_TRAFFIC_SUMMARY, "mc_emit_bytes_per_launch")]:
        b = measured_traffic(path, key)
        if b is None or tname not in ktimes:
            continue
        ms = ktimes[tname][0] / ktimes[tname][1]
        gbs = b / (ms * 1e-3) / 1e9
        out[name] = {"bytes_per_launch": b, "avg_launch_ms": round(ms, 3), "gb_per_s": round(gbs, 1),
                     "frac_of_8000": round(gbs / PEAK_HBM, 4)}
    return out


def cpu_baseline(P, grid_reso, iters, normals=None, budget_s=20.0):
    """FP64 CPU restatement (the oracle) on a bounded sample of the same work:
    octree + moments over all sources, then screened Barnes-Hut sums at a
    random subset of the source and grid queries; the per-query rates are
    scaled to the full reconstruction (iters × (sources + grid) queries).
    Also returns the sample (points, normals, areas, queries and their FP64
    sums) so that the GPU sums can be checked against it."""
    import oracle as O
    cores = O.threads()
    c = P.astype(np.float64).mean(0)
    Pn = P.astype(np.float64) - c
    Pn /= 2 * np.abs(Pn).max()
    N0 = O.gauss_map(Pn) if normals is None else np.asarray(normals, np.float64)
    t0 = time.perf_counter()
    tree = O.Octree(Pn, 10, 1)
    A = np.full(len(Pn), 2e-6)
    tree.precompute(N0, A)
    t_build = time.perf_counter() - t0
    origin, h, res = O.voxel_grid(Pn, 0.05, grid_reso)
    G = int(res[0]) * int(res[1]) * int(res[2])
    rng = np.random.default_rng(0)
    # grow the sample until it takes a measurable fraction of the budget
    m = 2000
    while True:
        qs = Pn[rng.integers(0, len(Pn), m)]
        gi = rng.integers(0, G, m)
        gq = np.stack([origin[0] + (gi % res[0]) * h, origin[1] + ((gi // res[0]) % res[1]) * h,
                       origin[2] + (gi // (res[0] * res[1])) * h], -1)
        t1 = time.perf_counter()
        w_src = tree.eval(qs, 2.3, kind="screened", sigma=10.0)
        t_src = time.perf_counter() - t1
        t1 = time.perf_counter()
        w_grid = tree.eval(gq, 2.3, kind="screened", sigma=10.0)
        t_grid = time.perf_counter() - t1
        if t_src + t_grid > budget_s / 4 or m >= 2_000_000:
            break
        m *= 4
    per_iter = t_src / m * len(Pn) + t_grid / m * G
    total = t_build + (iters + 1) * per_iter
    base = {"value": round(total, 1), "unit": "s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle: octree+moments over {len(Pn)} points ({t_build:.1f} s) and screened "
                      f"Barnes-Hut sums at {m} source + {m} grid queries, scaled to {iters + 1} evaluations of "
                      f"{len(Pn)} sources + {G} grid points (MC/kNN/diffusion not counted)"}
    sample = {"P": Pn, "N": N0, "A": A, "Q": np.concatenate([qs, gq]), "w": np.concatenate([w_src, w_grid])}
    return base, sample


def precision_check(ctx, sample, n_direct=256):
    """Error of the GPU sums against the FP64 oracle on the CPU-baseline
    sample: the same octree (the 1M points as sources, the plain run's final
    normals, areas 2e-6), β = 2.3, σ = 10, at its source and grid queries,
    for FP32, FP32 + Kahan and BF16 far field + Kahan (the learned+mixed
    iterations' mode); and against the FP64 all-pairs sum at n_direct of them
    (which adds the Barnes-Hut truncation).  Errors relative to max|w|."""
    import oracle as O
    from dwg_amd import gpu
    P, N, A, Q, ref = sample["P"], sample["N"], sample["A"], sample["Q"], sample["w"]
    t = gpu.Octree(ctx, P, A, max_depth=10)
    t.update(N.astype(np.float32), A.astype(np.float32))
    Qf = Q.astype(np.float32)
    scale = float(np.abs(ref).max())
    sub = np.random.default_rng(1).choice(len(Q), n_direct, replace=False)
    ref_d = O.direct(P, N, A, Qf[sub].astype(np.float64), kind="screened", sigma=10.0)
    out = {"reference": f"FP64 oracle tree (same beta, same term set) at {len(Q)} queries; FP64 direct sum at "
                        f"{n_direct} of them", "max_abs_w": round(scale, 6)}
    for name, kw in (("fp32", {}), ("fp32_kahan", {"kahan": True}), ("bf16_far_kahan", {"kahan": True, "bf16": True})):
        w = t.eval(Qf, 2.3, sigma=10.0, **kw).cpu().numpy().astype(np.float64)
        d = np.abs(w - ref) / scale
        dd = np.abs(w[sub] - ref_d) / scale
        out[name] = {"vs_fp64_tree_max": float(f"{d.max():.3e}"), "vs_fp64_tree_mean": float(f"{d.mean():.3e}"),
                     "vs_fp64_direct_mean": float(f"{dd.mean():.3e}")}
    out["fp64_tree_vs_fp64_direct_mean"] = float(f"{(np.abs(ref[sub] - ref_d) / scale).mean():.3e}")
    return out


def gate(q, q_plain):
    return bool(q["mean_angle_deg"] <= q_plain["mean_angle_deg"] + 0.5 and q["chamfer"] <= 1.01 * q_plain["chamfer"])


def plain_gate_scan(ctx, P, args, Ntrue, S, q_plain):
    """Plain FP32 at fewer iterations (even counts only: the normals' sign
    alternates each iteration), from iters − 2 down until one misses the
    equal-error gate against the full plain run.  Returns the fewest count
    that passed (the full count if none), its one-step wall time, and the
    scan."""
    scan = []
    best = (args.iters, None)
    for k in range(args.iters - 2, 1, -2):
        t, r = timed(ctx, P, dict(plain_options(args), total_iter=k), 1, f"plain {k} iterations")
        q = quality(ctx, r, Ntrue, S)
        ok = gate(q, q_plain)
        scan.append({"iterations": k, "seconds": round(t[0], 3), "mean_angle_deg": q["mean_angle_deg"],
                     "chamfer": q["chamfer"], "gate": ok})
        if not ok:
            break
        best = (k, t[0])
    return best, scan


def quality(ctx, r, Ntrue, S):
    from dwg_amd import metrics
    ang = metrics.normal_angle_deg(r.normals, Ntrue)
    cd = metrics.chamfer(ctx, r.tris_input_frame(), S)
    return {"mean_angle_deg": round(float(ang.mean()), 3), "median_angle_deg": round(float(np.median(ang)), 3),
            "flipped_frac": round(float((ang > 90).mean()), 5), "chamfer": round(cd, 7),
            "triangles": int(r.tris.shape[0]) if r.tris is not None else 0}


def timed(ctx, P, opts, steps, label, mlp=None):
    import torch
    times, last = [], None
    for s in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        last = run_once(ctx, P, opts, mlp=mlp)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"{label} step {s}: {times[-1]:.3f} s (setup {last.time_setup_s:.3f}, iterations {last.time_iter_s:.3f}, "
            f"final {last.time_final_s:.3f})")
    return times, last


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--iters", type=int, default=40)
    ap.add_argument("--learned-iters", type=int, default=LEARNED_ITERS)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--no-gate-scan", action="store_true",
                    help="skip the plain runs at fewer iterations that find the fewest meeting the gate")
    ap.add_argument("--large", action="store_true",
                    help="only run the large configuration: 5M points (128 primitives + 16 thin shells, 5%% in "
                         "holes), 1024^3 grids, learned+mixed vs plain FP32; prints its own JSON line")
    ap.add_argument("--direct-sums", action="store_true",
                    help="only time the direct (all-pairs) sums of the small sphere configuration: 10,000 "
                         "sources, 128^3 grid queries")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return
    if args.direct_sums:
        for line in direct_sums():
            print(json.dumps(line), flush=True)
        return
    if args.large:
        large_config(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the reconstruction "
              "runs on a single GPU; sharding it by grid slabs across GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)
    import torch
    from dwg_amd import gpu

    log(f"generating workload: {args.n} points, seed {args.seed}")
    P, Ntrue = workload(args.n, args.seed)
    ctx = gpu.Context(0)
    if not os.path.exists(args.weights):
        raise SystemExit(f"learned-initialisation weights not found: {args.weights}")
    net = gpu.Mlp(ctx, path=args.weights)
    plain, learned = plain_options(args), learned_options(args)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        run_once(ctx, P, plain)
        run_once(ctx, P, learned, mlp=net)
        log(f"warmup {w}: {time.perf_counter() - t0:.2f} s")
    torch.cuda.synchronize()
    plain_t, plain_r = timed(ctx, P, plain, args.steps, "plain")
    learned_t, learned_r = timed(ctx, P, learned, args.steps, "learned+mixed", mlp=net)
    value = float(np.mean(learned_t))
    plain_s = float(np.mean(plain_t))
    # profiled run of the measured method: per-kernel HIP-event times on the
    # library's stream and exact term counts
    prof = run_once(ctx, P, learned, profile=True, count_terms=True, mlp=net)
    ktimes = {}
    for k in KERNELS + [k + ":final" for k in KERNELS]:
        ms, cnt = ctx.profile_get(k)
        if cnt:
            ktimes[k] = (ms, cnt)
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    log("kernel times (ms total, launches): " + ", ".join(f"{k}={v[0]:.1f}/{v[1]}" for k, v in ktimes.items()))
    g_ms, g_cnt = ktimes["gwn_tree_grid_screened"]
    # final-pass launches are recorded as "<kernel>:final", so this average
    # covers the iteration grid evaluations only
    avg_ms = g_ms / g_cnt
    achieved = prof.terms_grid * FLOPS_PER_TERM["screened"] / (avg_ms * 1e-3) / 1e12
    roofline = {"kernel": "gwn_tree_grid_screened (gwn_tree_kernel<screened, BF16 far field, grid>)",
                "bound": "valu", "pipe": "fp32 VALU (instruction issue)",
                "achieved": round(achieved, 2), "peak": PEAK_FP32, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32, 4), "traffic": measured_traffic(),
                "terms_per_launch": int(prof.terms_grid), "flops_per_term": FLOPS_PER_TERM["screened"],
                "avg_launch_ms": round(avg_ms, 3), "launches": int(g_cnt), "dominant_by_time": dom,
                "hbm": hbm_rates(ktimes)}
    q_plain = q_learned = passed = None
    at_gate = {}
    if not args.no_quality:
        S = surface_sample(2_000_000, args.seed)
        q_plain, q_learned = quality(ctx, plain_r, Ntrue, S), quality(ctx, learned_r, Ntrue, S)
        passed = gate(q_learned, q_plain)
        if not args.no_gate_scan:
            (k, tk), scan = plain_gate_scan(ctx, P, args, Ntrue, S, q_plain)
            tk = plain_s if tk is None else tk
            at_gate = {"plain_iters_to_gate": k, "plain_at_gate_s": round(tk, 3),
                       "speedup_vs_plain_at_gate": round(tk / value, 3), "plain_gate_scan": scan}
    direct = direct_sums(ctx)
    cpu = precision = None
    if not args.no_cpu:
        cpu, sample = cpu_baseline(P, args.grid, args.iters, normals=plain_r.normals)
        precision = precision_check(ctx, sample)
    out = {"metric": METRIC, "value": round(value, 3), "unit": "s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(value * 1000, 1), "higher_is_better": False,
           "scaling": "strong", "vs_baseline": None, "dtype": "fp32+bf16", "data": "synthetic",
           "config": {"workload": f"union of 32 primitives, {args.n} points (seed {args.seed}), noise 0.001*diag, "
                                  f"{args.grid}^3 grids, octree-leaf sources, beta 2.3",
                      "method": f"learned+mixed: MLP start, {args.learned_iters} iterations, BF16 far field + Kahan",
                      "n_sources": int(learned_r.n_sources),
                      "phases_s": {"setup": round(learned_r.time_setup_s, 3),
                                   "iterations": round(learned_r.time_iter_s, 3),
                                   "final": round(learned_r.time_final_s, 3)},
                      "quality": q_learned,
                      "plain_fp32": {"seconds": round(plain_s, 3), "iterations": args.iters, "quality": q_plain,
                                     "phases_s": {"setup": round(plain_r.time_setup_s, 3),
                                                  "iterations": round(plain_r.time_iter_s, 3),
                                                  "final": round(plain_r.time_final_s, 3)}},
                      "speedup_vs_plain": round(plain_s / value, 3), "equal_error": passed, **at_gate},
           "roofline": roofline, "direct_sums": direct, "precision": precision, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
