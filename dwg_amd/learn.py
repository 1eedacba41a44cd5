"""Training and evaluation of the learned normal initialisation.

Training data: DWG's own converged normals.  Each training shape is the
union-of-32-primitives generator at 50,000 points (seeds 1000, 1001, …);
plain DWG (Gauss-map init, 40 iterations, points as sources) runs on the
GPU and its final point normals are the targets for the 16-neighbour
patches of the same points.  Training is Adam on the normal-line loss
mean(1 − (ŷ·t)²) with one random rotation applied to each batch (patches
and targets together): a local patch fixes the line of the normal but not
reliably its side, so the initialisation takes the side from the Gauss-map
hemisphere (away from the bounding-box centre) and the DWG iterations
resolve the rest.

Evaluation on held-out shapes (seeds 2000, …): the plain method's final
mean normal angle against the analytic normals (Gauss-map start, 40
iterations), and the angle reached after k DWG iterations from the learned
start and from the Gauss-map start, k = 0, 2, 4, …; the result is the
smallest k within 0.5° of the plain method and its wall time.

    python -m dwg_amd.learn --shapes 512 --steps 6000 --out weights.bin --eval 64
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import gpu, metrics, pipeline, shapes

TRAIN_SEED0 = 1000
EVAL_SEED0 = 2000


def log(*a):
    print(*a, flush=True)


def normalised(P, r):
    return ((P.astype(np.float64) - r.center) / r.scale).astype(np.float32)


def plain_options(n, iters=40):
    grid = 192 if n <= 100_000 else 384
    return dict(source_mode=1, grid_reso=grid, final_reso=grid, total_iter=iters, input_normal=3)


def training_set(ctx, net, seeds, n=50_000, iters=40):
    X, T = [], []
    t0 = time.time()
    for i, s in enumerate(seeds):
        P, _ = shapes.union(n, 32, s)
        r = pipeline.reconstruct(ctx, P, want_mesh=False, **dict(plain_options(n, iters), final_reso=32))
        idx = gpu.KnnIndex(ctx, normalised(P, r), k_hint=17)
        X.append(net.patches(idx))
        T.append(torch.from_numpy(r.normals).cuda())
        if i % 32 == 0:
            log(f"training shape {i}/{len(seeds)} (seed {s}): {time.time() - t0:.1f} s")
    return torch.cat(X), torch.cat(T)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]], np.float32)


def train(ctx, net, X, T, steps=6000, batch=8192, lr=2e-3, seed=0, augment=True, line=True):
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    xb = torch.empty((batch, 48), dtype=torch.int16, device=X.device)
    tb = torch.empty((batch, 3), dtype=torch.float32, device=X.device)
    hist = []
    t0 = time.time()
    for step in range(steps):
        sel = torch.from_numpy(rng.integers(0, n, batch)).to(X.device)
        R = random_rotation(rng) if augment else np.eye(3, dtype=np.float32)
        net.augment(X, T, sel, R, xb, tb)
        loss = net.loss_grad(xb, tb, line=line)
        # constant learning rate for 60% of the steps, then cosine decay
        f = max(0.0, (step - 0.6 * steps) / (0.4 * steps))
        net.adam(lr=lr * (0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * f))))
        hist.append(loss)
        if step % 500 == 0 or step == steps - 1:
            log(f"step {step}: loss {np.mean(hist[-100:]):.5f} ({time.time() - t0:.1f} s)")
    return hist


def evaluate(ctx, net, seeds, n=50_000, ks=(0, 2, 4, 6, 8, 10, 12, 16, 20, 24, 30), tol_deg=0.5):
    rows = []
    for s in seeds:
        P, N = shapes.union(n, 32, s)
        base = pipeline.reconstruct(ctx, P, want_mesh=False, **plain_options(n))
        plain = float(metrics.normal_angle_deg(base.normals, N).mean())
        row = {"seed": s, "plain_deg": plain, "plain_s": base.time_total_s, "learned": [], "plain_k": []}
        for k in ks:
            r = pipeline.reconstruct(ctx, P, want_mesh=False, mlp=net, **dict(plain_options(n, k), input_normal=4))
            row["learned"].append((k, float(metrics.normal_angle_deg(r.normals, N).mean()), r.time_total_s))
            r = pipeline.reconstruct(ctx, P, want_mesh=False, **plain_options(n, k))
            row["plain_k"].append((k, float(metrics.normal_angle_deg(r.normals, N).mean()), r.time_total_s))
        reach = [(k, a, t) for k, a, t in row["learned"] if a <= plain + tol_deg]
        row["k_reach"] = reach[0][0] if reach else None
        row["t_reach"] = reach[0][2] if reach else None
        reach = [(k, a, t) for k, a, t in row["plain_k"] if a <= plain + tol_deg]
        row["k_reach_plain"] = reach[0][0] if reach else None
        rows.append(row)
        log(f"eval seed {s}: plain {plain:.2f}° in {base.time_total_s:.3f} s; learned "
            + ", ".join(f"k={k}:{a:.2f}°" for k, a, _ in row["learned"]) + "; plain "
            + ", ".join(f"k={k}:{a:.2f}°" for k, a, _ in row["plain_k"]))
    return rows


def summary(rows):
    ks = [r["k_reach"] for r in rows if r["k_reach"] is not None]
    kp = [r["k_reach_plain"] for r in rows if r["k_reach_plain"] is not None]
    return {"shapes": len(rows), "reached": len(ks), "median_k": float(np.median(ks)) if ks else None,
            "reached_gauss": len(kp), "median_k_gauss": float(np.median(kp)) if kp else None,
            "plain_mean_deg": float(np.mean([r["plain_deg"] for r in rows])),
            "learned_k0_mean_deg": float(np.mean([r["learned"][0][1] for r in rows])),
            "median_plain_s": float(np.median([r["plain_s"] for r in rows])),
            "median_learned_s": float(np.median([r["t_reach"] for r in rows if r["t_reach"] is not None]))
            if ks else None}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", type=int, default=512)
    ap.add_argument("--points", type=int, default=50_000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=64)
    ap.add_argument("--out", default="dwg_mlp.bin")
    ap.add_argument("--report", default=None)
    ap.add_argument("--eval-only", default=None, metavar="WEIGHTS",
                    help="skip training: evaluate these weights on the held-out shapes")
    args = ap.parse_args(argv)
    ctx = gpu.Context(0)
    if args.eval_only:
        net = gpu.Mlp(ctx, path=args.eval_only)
        rows = evaluate(ctx, net, range(EVAL_SEED0, EVAL_SEED0 + args.eval), args.points)
        rep = {"weights": args.eval_only, "eval": summary(rows), "rows": rows}
        log(json.dumps(rep["eval"]))
        if args.report:
            with open(args.report, "w") as f:
                json.dump(rep, f, indent=1)
        return rep
    net = gpu.Mlp(ctx, 1)
    X, T = training_set(ctx, net, range(TRAIN_SEED0, TRAIN_SEED0 + args.shapes), args.points)
    log(f"training set: {X.shape[0]} patches")
    hist = train(ctx, net, X, T, args.steps, args.batch, args.lr)
    net.save(args.out)
    del X, T
    rows = evaluate(ctx, net, range(EVAL_SEED0, EVAL_SEED0 + args.eval), args.points) if args.eval else []
    rep = {"train": {"shapes": args.shapes, "points": args.points, "steps": args.steps, "batch": args.batch,
                     "final_loss": float(np.mean(hist[-100:]))}, "eval": summary(rows) if rows else None,
           "rows": rows}
    log(json.dumps(rep["eval"]))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
