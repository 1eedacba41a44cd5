"""Reconstruction quality: normal angle error and Chamfer distance.

Both run on the GPU through the library's exact kNN.
"""
from __future__ import annotations

import numpy as np
import torch

from . import gpu


def normal_angle_deg(est, truth, oriented=True):
    """Per-point angle (degrees) between estimated and true normals."""
    e = np.asarray(est, np.float64)
    t = np.asarray(truth, np.float64)
    c = np.einsum("ij,ij->i", e, t) / (np.linalg.norm(e, axis=1) * np.linalg.norm(t, axis=1) + 1e-30)
    if not oriented:
        c = np.abs(c)
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def mesh_samples(tris, per_tri=1):
    """Area-weighted sample set of a triangle soup: triangle centroids with
    their areas as weights."""
    t = np.asarray(tris, np.float64)
    c = t.mean(1)
    a = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    return c.astype(np.float32), a


def chamfer(ctx: gpu.Context, mesh_tris, surface_pts):
    """Symmetric Chamfer distance between a triangle mesh (area-weighted
    centroids) and a dense sample of the true surface:
    mean_{m} min |m − s| (area-weighted) + mean_{s} min |s − m|."""
    c, a = mesh_samples(mesh_tris)
    s = np.ascontiguousarray(surface_pts, np.float32)
    if len(c) == 0:
        return float("inf")
    si = gpu.KnnIndex(ctx, s, k_hint=1)
    _, d_ms = si.query(c, 1, d2=True)
    mi = gpu.KnnIndex(ctx, c, k_hint=1)
    _, d_sm = mi.query(s, 1, d2=True)
    d1 = torch.sqrt(d_ms[:, 0].double()).cpu().numpy()
    d2 = torch.sqrt(d_sm[:, 0].double()).cpu().numpy()
    return float((d1 * a).sum() / a.sum() + d2.mean())
