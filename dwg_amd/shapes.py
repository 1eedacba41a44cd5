"""Seeded synthetic point clouds for the DWG workloads, with analytic normals.

Every generator returns ``(points float32 [N,3], normals float32 [N,3])``;
the normals are the exact outward unit normals of the clean surface at the
sampled point (before noise) and are used only to score results — the
reconstruction itself receives positions only, as the reference does with
unoriented input (--in_normal 3, DWGProcess.h:474-510).

Randomness: numpy's PCG64 generator seeded with the workload seed; all
sampling is area-uniform on the clean surface.  Rotations, dot products
and norms are written out as elementwise products and sums in a fixed
order (no BLAS matrix product, no einsum), so every inside / accept
decision, and with it the sample, is the same on every host.

Workloads:
  sphere(n=10_000, seed=1)            unit sphere, no noise
  torus(n=100_000, seed=2)            R=1, r=0.35, density 1..8 along the major
                                      angle, noise 0.002·bbox-diagonal
  union(n=1_000_000, m=32, seed=3)    union of m random spheres / ellipsoids /
                                      tori in the unit cube, noise 0.001·diag
  large(n=5_000_000, seed=5)          128 primitives + 0.01-thick sphere
                                      shells, 5% removed in 0.05-radius holes
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def _dot3(a, b):
    """Row-wise a·b of n×3 arrays, summed x, y, z in that order."""
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _mat3(x, m):
    """Rows of x times the 3×3 matrix m (x @ m), each entry a fixed-order sum."""
    return np.stack([x[:, 0] * m[0, j] + x[:, 1] * m[1, j] + x[:, 2] * m[2, j] for j in range(3)], -1)


def _unit(v):
    """v over its length along the last axis (summed in index order)."""
    ss = v[..., 0] * v[..., 0]
    for k in range(1, v.shape[-1]):
        ss = ss + v[..., k] * v[..., k]
    return v / np.sqrt(ss)[..., None]


def sphere(n=10_000, seed=1, radius=1.0, center=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    nrm = np.stack([s * np.cos(phi), s * np.sin(phi), z], -1)
    pts = np.asarray(center) + radius * nrm
    return pts.astype(np.float32), nrm.astype(np.float32)


def torus(n=100_000, seed=2, R=1.0, r=0.35, noise=0.002):
    """Torus with density ratio 1..8 along the major angle φ (rejection sampling)."""
    rng = np.random.default_rng(seed)
    out_p, out_n, have = [], [], 0
    while have < n:
        m = 4 * (n - have) + 1024
        th = rng.uniform(0, 2 * np.pi, m)
        ph = rng.uniform(0, 2 * np.pi, m)
        dens = 1.0 + 3.5 * (1.0 - np.cos(ph))  # in [1, 8]
        acc = rng.uniform(0, 1, m) < ((R + r * np.cos(th)) / (R + r)) * (dens / 8.0)
        th, ph = th[acc], ph[acc]
        rr = R + r * np.cos(th)
        out_p.append(np.stack([rr * np.cos(ph), rr * np.sin(ph), r * np.sin(th)], -1))
        out_n.append(np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), np.sin(th)], -1))
        have += int(acc.sum())
    p = np.concatenate(out_p)[:n]
    nn = np.concatenate(out_n)[:n]
    e = p.max(0) - p.min(0)
    diag = float(np.sqrt(_dot3(e, e)))
    p = p + rng.normal(0.0, noise * diag, p.shape)
    return p.astype(np.float32), nn.astype(np.float32)


class _Prim:
    """A sphere (0), ellipsoid (1), torus (2) or sphere shell (3) with a pose."""

    def __init__(self, kind, params, rot, ctr):
        self.kind, self.params, self.rot, self.ctr = kind, np.asarray(params, float), rot, np.asarray(ctr)

    def area(self):
        p = self.params
        if self.kind == 0:
            return 4 * np.pi * p[0] ** 2
        if self.kind == 1:
            a, b, c = p
            q = 1.6075  # Knud Thomsen's approximation, < 1.1% error
            return 4 * np.pi * (((a * b) ** q + (a * c) ** q + (b * c) ** q) / 3) ** (1 / q)
        if self.kind == 2:
            return 4 * np.pi ** 2 * p[0] * p[1]
        return 4 * np.pi * (p[0] ** 2 + (p[0] - p[1]) ** 2)

    def bound(self):
        """Radius of a ball about the centre that holds the solid."""
        p = self.params
        if self.kind == 1:
            return float(max(p))
        if self.kind == 2:
            return float(p[0] + p[1])
        return float(p[0])

    def to_local(self, x):
        return _mat3(x - self.ctr, self.rot)  # rot columns are the local axes

    def inside(self, x):
        """Strictly inside the solid (points on or near the boundary are not)."""
        p = self.params
        if self.kind in (0, 3):  # radial: no rotation needed
            d = x - self.ctr
            d2 = _dot3(d, d)
            if self.kind == 0:
                return d2 < (p[0] * (1 - 1e-6)) ** 2
            ro, t = p
            return (d2 < (ro * (1 - 1e-6)) ** 2) & (d2 > ((ro - t) * (1 + 1e-6)) ** 2)
        y = self.to_local(x)
        if self.kind == 1:
            z = y / p
            return _dot3(z, z) < 1 - 1e-6
        if self.kind == 2:
            R, r = p
            q = np.sqrt(y[:, 0] ** 2 + y[:, 1] ** 2) - R
            return q * q + y[:, 2] ** 2 < (r * (1 - 1e-6)) ** 2

    def sample(self, rng, m):
        """m area-uniform surface samples (points, outward normals) in world frame."""
        p = self.params
        if self.kind in (0, 3):
            u = _unit(rng.normal(size=(m, 3)))
            if self.kind == 0:
                y, n = p[0] * u, u
            else:  # outer sphere and inner sphere (normals pointing inwards), by area
                ro, ri = p[0], p[0] - p[1]
                outer = rng.uniform(0, 1, m) < ro ** 2 / (ro ** 2 + ri ** 2)
                y = np.where(outer[:, None], ro * u, ri * u)
                n = np.where(outer[:, None], u, -u)
        elif self.kind == 1:
            a, b, c = p
            ys, ns, have = [], [], 0
            gmax = max(b * c, a * c, a * b)
            while have < m:
                u = _unit(rng.normal(size=(2 * (m - have) + 64, 3)))
                g = np.sqrt((b * c * u[:, 0]) ** 2 + (a * c * u[:, 1]) ** 2 + (a * b * u[:, 2]) ** 2)
                keep = rng.uniform(0, 1, len(u)) < g / gmax
                u = u[keep]
                ys.append(u * p)
                ns.append(_unit(u / p))
                have += len(u)
            y, n = np.concatenate(ys)[:m], np.concatenate(ns)[:m]
        else:
            R, r = p
            ys, ns, have = [], [], 0
            while have < m:
                k = 2 * (m - have) + 64
                th = rng.uniform(0, 2 * np.pi, k)
                ph = rng.uniform(0, 2 * np.pi, k)
                keep = rng.uniform(0, 1, k) < (R + r * np.cos(th)) / (R + r)
                th, ph = th[keep], ph[keep]
                rr = R + r * np.cos(th)
                ys.append(np.stack([rr * np.cos(ph), rr * np.sin(ph), r * np.sin(th)], -1))
                ns.append(np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), np.sin(th)], -1))
                have += len(th)
            y, n = np.concatenate(ys)[:m], np.concatenate(ns)[:m]
        return _mat3(y, self.rot.T) + self.ctr, _mat3(n, self.rot.T)


def _random_rotation(rng):
    q = _unit(rng.normal(size=4))
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def random_primitives(rng, m, shells=0):
    prims = []
    for i in range(m + shells):
        kind = 3 if i >= m else int(rng.integers(0, 3))
        if kind == 0:
            params = [rng.uniform(0.1, 0.3)]
            b = params[0]
        elif kind == 1:
            params = list(rng.uniform(0.05, 0.3, 3))
            b = max(params)
        elif kind == 2:
            R = rng.uniform(0.15, 0.3)
            params = [R, R * rng.uniform(0.2, 0.4)]
            b = params[0] + params[1]
        else:
            params = [rng.uniform(0.1, 0.3), 0.01]
            b = params[0]
        b = min(b, 0.499)
        ctr = rng.uniform(b, 1.0 - b, 3)
        prims.append(_Prim(kind, params, _random_rotation(rng), ctr))
    return prims


def union_of(prims, n, rng, noise, adaptive=False, progress=None, clean=False):
    """Area-uniform samples of the union's boundary: each primitive is sampled
    in proportion to its area and samples inside any other primitive are
    rejected; then Gaussian noise of `noise`·bbox-diagonal.  Each pass draws
    1.5× the missing count, or (adaptive) the missing count over the
    acceptance rate seen so far, which needs few passes when the primitives
    overlap a lot.  progress(have, n) is called after every pass; clean=True
    also returns the points before the noise."""
    areas = np.array([p.area() for p in prims])
    probs = areas / areas.sum()
    # a sample of primitive i can only be inside primitive j when their
    # bounding balls overlap
    ctr = np.array([p.ctr for p in prims])
    rad = np.array([p.bound() for p in prims])
    dc = ctr[:, None] - ctr[None]
    gap = np.sqrt(_dot3(dc, dc)) - (rad[:, None] + rad[None])
    near = [[j for j in np.nonzero(gap[i] < 1e-9)[0] if j != i] for i in range(len(prims))]
    pts, nrms, have, drawn = [], [], 0, 0
    while have < n:
        if adaptive and have > 0:
            m = int(1.1 * (n - have) * drawn / have) + 4096
        else:
            m = int(1.5 * (n - have)) + 4096
        drawn += m
        counts = rng.multinomial(m, probs)
        drawn_i = [(i, *prims[i].sample(rng, int(c))) for i, c in enumerate(counts) if c > 0]

        def kept(item):
            i, x, _ = item
            keep = np.ones(len(x), bool)
            for j in near[i]:
                # only the samples inside j's bounding box can be inside j
                cand = np.nonzero(keep & (np.abs(x - ctr[j]) < rad[j]).all(1))[0]
                if cand.size:
                    keep[cand] = ~prims[j].inside(x[cand])
            return keep

        # the tests release the GIL inside numpy, so threads share the work;
        # the samples were all drawn first, in order, so the result does not
        # depend on the thread count
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
            keeps = list(pool.map(kept, drawn_i))
        for (i, x, nn), keep in zip(drawn_i, keeps):
            pts.append(x[keep])
            nrms.append(nn[keep])
            have += int(keep.sum())
        if progress:
            progress(have, n)
    p = np.concatenate(pts)
    nn = np.concatenate(nrms)
    perm = rng.permutation(len(p))[:n]
    p, nn = p[perm], nn[perm]
    e = p.max(0) - p.min(0)
    diag = float(np.sqrt(_dot3(e, e)))
    p0 = p
    p = p + rng.normal(0.0, noise * diag, p.shape)
    if clean:
        return p.astype(np.float32), nn.astype(np.float32), p0.astype(np.float32)
    return p.astype(np.float32), nn.astype(np.float32)


def union(n=1_000_000, m=32, seed=3, noise=0.001):
    rng = np.random.default_rng(seed)
    prims = random_primitives(rng, m)
    return union_of(prims, n, rng, noise)


def large(n=5_000_000, seed=5, m=128, shells=16, hole_frac=0.05, hole_radius=0.05, noise=0.001, progress=None,
          clean=False):
    """Large workload: primitives plus 0.01-thick sphere shells, with seeded
    hole patches of radius 0.05 removing about 5% of the points.  clean=True
    also returns the same points before the noise (for Chamfer distances)."""
    rng = np.random.default_rng(seed)
    prims = random_primitives(rng, m, shells)
    n0 = int(np.ceil(n / (1.0 - hole_frac)))
    p, nn, p0 = union_of(prims, n0, rng, noise, adaptive=True, progress=progress, clean=True)
    keep = np.ones(len(p), bool)
    while keep.sum() > n:
        c = p[rng.integers(0, len(p))]
        d = p - c
        keep &= _dot3(d, d) > hole_radius ** 2
    idx = np.nonzero(keep)[0][:n]
    if clean:
        return p[idx], nn[idx], p0[idx]
    return p[idx], nn[idx]


WORKLOADS = {
    "sphere10k": lambda: sphere(10_000, 1),
    "torus100k": lambda: torus(100_000, 2),
    "union1m": lambda: union(1_000_000, 32, 3),
    "large5m": lambda: large(5_000_000, 5),
}
