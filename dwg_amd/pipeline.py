"""The whole DWG reconstruction (dwg_launch → the C++ TurboDWG driver).

``reconstruct(points, **options)`` mirrors the DWG_CUDA command line
(src/main.cpp:17-166): every option keeps the reference flag's name and
default; the MI355X-only options are ``source_mode``, ``leaf_size``,
``beta``, ``kahan``, ``bf16_far``, ``near_fp64``, ``residual_tol`` and
``seed``.
"""
from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import DWG_FAR_BF16, DWG_KAHAN, DWG_NEAR_FP64, DwgOptions, DwgResult


@dataclass
class Reconstruction:
    normals: np.ndarray          # final point normals (normalised frame; same as input frame up to scale)
    tris: np.ndarray | None      # final mesh, normalised frame (T×3×3)
    n_sources: int
    center: np.ndarray
    scale: float
    time_setup_s: float
    time_iter_s: float
    time_final_s: float
    time_total_s: float
    w_mean: np.ndarray = field(default_factory=lambda: np.zeros(0))
    iter_tris: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    iter_zeros: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    terms_src: int = 0
    terms_grid: int = 0
    grid_points: int = 0
    source_normals: np.ndarray | None = None  # the sources' normals after the iterations (on request)
    residual: np.ndarray = field(default_factory=lambda: np.zeros(0))  # convergence residual per iteration
    iterations: int = 0                      # iterations run

    def tris_input_frame(self):
        return None if self.tris is None else self.tris * self.scale + self.center


def options(**kw) -> DwgOptions:
    L = _lib.load()
    o = DwgOptions()
    L.dwg_options_default(ctypes.byref(o))
    kahan = kw.pop("kahan", None)
    bf16 = kw.pop("bf16_far", None)
    near64 = kw.pop("near_fp64", None)
    lam = kw.pop("lam", None)
    if lam is not None:
        kw["sigma"] = lam
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown DWG option {k}")
        setattr(o, k, v)
    if kahan is not None:
        o.precision = (o.precision | DWG_KAHAN) if kahan else (o.precision & ~DWG_KAHAN)
    if bf16 is not None:
        o.precision = (o.precision | DWG_FAR_BF16) if bf16 else (o.precision & ~DWG_FAR_BF16)
    if near64 is not None:
        o.precision = (o.precision | DWG_NEAR_FP64) if near64 else (o.precision & ~DWG_NEAR_FP64)
    return o


def _host_array(L, ptr, ntri):
    """The library's mesh buffer as a (T, 3, 3) array without a copy; the
    buffer goes back to dwg_free when the array is collected."""
    if not ptr or ntri == 0:
        if ptr:
            L.dwg_free(ptr)
        return np.zeros((0, 3, 3), np.float32)
    raw = (ctypes.c_float * (9 * ntri)).from_address(ptr)
    weakref.finalize(raw, L.dwg_free, ptr)
    return np.frombuffer(raw, np.float32).reshape(ntri, 3, 3)


def reconstruct(ctx, points, normals=None, want_mesh=True, out_base=None, mlp=None, want_source_normals=False,
                want_residual=False, **kw) -> Reconstruction:
    """Run DWG on host points (N×3); ``ctx`` is a dwg_amd.gpu.Context."""
    L = _lib.load()
    P = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
    n = len(P)
    o = options(**kw)
    if mlp is not None:
        o.mlp = mlp.h
    nin = None if normals is None else np.ascontiguousarray(np.asarray(normals, np.float32).reshape(-1, 3))
    r = DwgResult()
    out_n = np.zeros((n, 3), np.float32)
    r.normals = out_n.ctypes.data
    iters = max(int(o.total_iter), 1)
    wm = np.zeros(iters)
    it_t = np.zeros(iters, np.int64)
    it_z = np.zeros(iters, np.int64)
    r.w_mean, r.iter_tris, r.iter_zeros = wm.ctypes.data, it_t.ctypes.data, it_z.ctypes.data
    r.want_mesh = 1 if want_mesh else 0
    src_n = np.zeros((n, 3), np.float32) if want_source_normals else None
    r.source_normals = None if src_n is None else src_n.ctypes.data
    resid = np.full(iters, np.nan) if want_residual else None
    r.residual = None if resid is None else resid.ctypes.data
    ctx.check(L.dwg_launch(ctx.h, P.ctypes.data, None if nin is None else nin.ctypes.data, n, ctypes.byref(o),
                           None if out_base is None else out_base.encode(), ctypes.byref(r)))
    tris = None
    if want_mesh:
        tris = _host_array(L, r.mesh, int(r.tri_count))
    k = int(r.iterations)
    return Reconstruction(out_n, tris, int(r.n_sources), np.array(r.center[:]), float(r.scale), r.time_setup_s,
                          r.time_iter_s, r.time_final_s, r.time_total_s, wm[:k], it_t[:k], it_z[:k], int(r.terms_src),
                          int(r.terms_grid), int(r.grid_points),
                          None if src_n is None else src_n[: int(r.n_sources)],
                          np.zeros(0) if resid is None else resid[: int(r.iterations)], int(r.iterations))
