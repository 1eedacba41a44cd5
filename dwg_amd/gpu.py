"""Thin Python mirror of the C ABI over torch device tensors.

torch only provides device memory and the stream; every computation is a
call into libdwg_hip.so.  Names follow the reference's operators:
``Octree.update`` ↔ TurboLibiGWN::UpdateMassNormals, ``Octree.eval`` ↔
TurboLibiScreenedGWNLaunch, ``KnnIndex`` ↔ GenerateKDTreeForDWG +
CudaKNNWithKDTreeforDWG, ``marching_cubes`` ↔ TurboMC::TurboMarchingCubes.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import DWG_FAR_BF16, DWG_KAHAN, DWG_NEAR_FP64, DWG_SCREENED, DWG_USE_RADII, DwgGrid


class DwgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dwg error {code}: {msg}")
        self.code = code


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _f32(t, device):
    if isinstance(t, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(t, dtype=np.float32))
    return t.to(device=device, dtype=torch.float32).contiguous()


def cache_release() -> int:
    """Hand the library's cached device blocks back to HIP; returns the bytes freed."""
    L = _lib.load()
    b = ctypes.c_int64()
    if L.dwg_cache_release(ctypes.byref(b)) != 0:
        raise DwgError(2, "dwg_cache_release failed")
    return b.value


class Context:
    """dwg_ctx on one device, issuing work on torch's current stream."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("libdwg_hip needs a HIP device")
        self.L = _lib.load()
        self.device = torch.device("cuda", device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        h = ctypes.c_void_p()
        rc = self.L.dwg_ctx_create(int(device), ctypes.c_void_p(stream), ctypes.byref(h))
        if rc != 0:
            raise DwgError(rc, "dwg_ctx_create failed")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise DwgError(rc, self.L.dwg_ctx_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.dwg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.L.dwg_ctx_sync(self.h))

    def profile(self, enable=True):
        self.check(self.L.dwg_ctx_profile(self.h, 1 if enable else 0))

    def profile_get(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self.check(self.L.dwg_ctx_profile_get(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def profile_reset(self):
        self.check(self.L.dwg_ctx_profile_reset(self.h))

    def counter(self, name):
        v = ctypes.c_int64()
        self.check(self.L.dwg_ctx_counter_get(self.h, name.encode(), ctypes.byref(v)))
        return v.value

    def f32(self, x):
        return _f32(x, self.device)

    # ---- sums ---------------------------------------------------------------
    def gwn_direct(self, pts, nrm, area, q, sigma=10.0, screened=False, radii=None, kahan=True, grad=False,
                   near64=False):
        pts, nrm, area, q = self.f32(pts), self.f32(nrm), self.f32(area), self.f32(q)
        rad = None if radii is None else self.f32(radii)
        flags = (DWG_SCREENED if screened else 0) | (DWG_USE_RADII if rad is not None else 0) | \
                (DWG_KAHAN if kahan else 0) | (DWG_NEAR_FP64 if near64 else 0)
        nq = q.numel() // 3
        w = torch.empty(nq, device=self.device, dtype=torch.float32)
        g = torch.empty((nq, 3), device=self.device, dtype=torch.float32) if grad else None
        self.check(self.L.dwg_gwn_direct(self.h, _ptr(pts), _ptr(nrm), _ptr(area), _ptr(rad), pts.numel() // 3,
                                         _ptr(q), nq, float(sigma), flags, _ptr(w), _ptr(g)))
        return (w, g) if grad else w

    def reduce_sum(self, v):
        v = self.f32(v)
        out = ctypes.c_double()
        self.check(self.L.dwg_reduce_sum(self.h, _ptr(v), v.numel(), ctypes.byref(out)))
        return out.value

    def bounds(self, pts):
        pts = self.f32(pts)
        lo = (ctypes.c_float * 3)()
        hi = (ctypes.c_float * 3)()
        self.check(self.L.dwg_bounds(self.h, _ptr(pts), pts.numel() // 3, lo, hi))
        return np.array(lo[:], np.float32), np.array(hi[:], np.float32)


def make_grid(origin, h, res) -> DwgGrid:
    g = DwgGrid()
    for d in range(3):
        g.origin[d] = float(origin[d])
        g.res[d] = int(res[d])
    g.h = float(h)
    return g


class Octree:
    """dwg_octree: octreeDWG split rule + order-0 moments, resident on the GPU."""

    def __init__(self, ctx: Context, pts, area=None, max_depth=10, leaf_size=1):
        self.ctx = ctx
        P = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
        A = None if area is None else np.ascontiguousarray(np.asarray(area, dtype=np.float64).reshape(-1))
        self.n = len(P)
        h = ctypes.c_void_p()
        ctx.check(ctx.L.dwg_octree_build(ctx.h, P.ctypes.data, None if A is None else A.ctypes.data, self.n,
                                         int(max_depth), int(leaf_size), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.ctx.L.dwg_octree_destroy(self.h)
            self.h = None

    def info(self):
        m, l, d = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        self.ctx.L.dwg_octree_info(self.h, ctypes.byref(m), ctypes.byref(l), ctypes.byref(d))
        return {"nodes": m.value, "leaves": l.value, "depth": d.value}

    def point_order(self):
        """Caller index of each point in tree order (dwg_octree_point_order)."""
        c = self.ctx
        out = torch.empty(max(self.n, 1), device=c.device, dtype=torch.int32)
        c.check(c.L.dwg_octree_point_order(c.h, self.h, _ptr(out)))
        return out[:self.n].cpu().numpy()

    def leaves(self, min_num=0):
        """Centres and widths of the leaf cells holding more than min_num
        points, breadth-first (GetOctreeLeaves)."""
        c = self.ctx
        nl = self.info()["leaves"]
        cen = torch.empty((max(nl, 1), 3), device=c.device, dtype=torch.float64)
        wid = torch.empty(max(nl, 1), device=c.device, dtype=torch.float64)
        cnt = ctypes.c_int64()
        c.check(c.L.dwg_octree_leaves(c.h, self.h, int(min_num), _ptr(cen), _ptr(wid), ctypes.byref(cnt)))
        return cen[:cnt.value].cpu().numpy(), wid[:cnt.value].cpu().numpy()

    def leaf_normals(self, pts, nrm, min_num=0):
        """SmoothNormalstoOctreeLeaves + whole-matrix normalisation of the
        tree's points' normals onto its leaves (dwg_octree_leaf_normals)."""
        c = self.ctx
        P = torch.as_tensor(np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, 3))).to(c.device)
        N = torch.as_tensor(np.ascontiguousarray(np.asarray(nrm, np.float64).reshape(-1, 3))).to(c.device)
        nl = self.info()["leaves"]
        out = torch.empty((max(nl, 1), 3), device=c.device, dtype=torch.float32)
        cnt = ctypes.c_int64()
        c.check(c.L.dwg_octree_leaf_normals(c.h, self.h, int(min_num), _ptr(P), _ptr(N), _ptr(out), ctypes.byref(cnt)))
        return out[:cnt.value].cpu().numpy()

    def update(self, nrm, area, radii=None):
        c = self.ctx
        self._nrm, self._area = c.f32(nrm), c.f32(area)
        self._radii = None if radii is None else c.f32(radii)
        c.check(c.L.dwg_octree_update(c.h, self.h, _ptr(self._nrm), _ptr(self._area), _ptr(self._radii)))

    def set_point_areas(self, area):
        """Point terms with other areas; the far-field moments stay those of
        the last update (dwg_octree_set_point_areas)."""
        c = self.ctx
        self._point_area = c.f32(area)
        c.check(c.L.dwg_octree_set_point_areas(c.h, self.h, _ptr(self._nrm), _ptr(self._point_area)))

    def _flags(self, screened, use_radii, kahan, bf16, near64=False):
        if use_radii is None:  # per-source radii whenever update() was given them
            use_radii = getattr(self, "_radii", None) is not None
        return (DWG_SCREENED if screened else 0) | (DWG_USE_RADII if use_radii else 0) | \
               (DWG_KAHAN if kahan else 0) | (DWG_FAR_BF16 if bf16 else 0) | (DWG_NEAR_FP64 if near64 else 0)

    def eval(self, q, beta=2.3, sigma=10.0, screened=True, kahan=False, bf16=False, out=None, use_radii=None,
             near64=False):
        c = self.ctx
        q = c.f32(q)
        nq = q.numel() // 3
        out = torch.empty(nq, device=c.device, dtype=torch.float32) if out is None else out
        c.check(c.L.dwg_gwn_tree(c.h, self.h, _ptr(q), nq, float(beta), float(sigma),
                                 self._flags(screened, use_radii, kahan, bf16, near64), _ptr(out)))
        return out

    def eval_grad(self, q, beta=2.3, sigma=10.0, screened=True):
        """Value and gradient of the Barnes-Hut sums at q (dwg_gwn_tree_grad)."""
        c = self.ctx
        q = c.f32(q)
        nq = q.numel() // 3
        w = torch.empty(nq, device=c.device, dtype=torch.float32)
        g = torch.empty((nq, 3), device=c.device, dtype=torch.float32)
        c.check(c.L.dwg_gwn_tree_grad(c.h, self.h, _ptr(q), nq, float(beta), float(sigma),
                                      DWG_SCREENED if screened else 0, _ptr(w), _ptr(g)))
        return w, g

    def eval_grid(self, grid: DwgGrid, beta=2.3, sigma=10.0, screened=True, kahan=False, bf16=False, out=None,
                  use_radii=None):
        c = self.ctx
        G = grid.res[0] * grid.res[1] * grid.res[2]
        out = torch.empty(G, device=c.device, dtype=torch.float32) if out is None else out
        c.check(c.L.dwg_gwn_tree_grid(c.h, self.h, ctypes.byref(grid), float(beta), float(sigma),
                                      self._flags(screened, use_radii, kahan, bf16), _ptr(out)))
        return out

    def terms(self, q, beta=2.3):
        c = self.ctx
        q = c.f32(q)
        nq = q.numel() // 3
        out = torch.empty(nq, device=c.device, dtype=torch.int64)
        c.check(c.L.dwg_gwn_tree_terms(c.h, self.h, _ptr(q), nq, float(beta), _ptr(out)))
        return out

    def grid_terms(self, grid: DwgGrid, beta=2.3):
        c = self.ctx
        tot = ctypes.c_int64()
        c.check(c.L.dwg_gwn_tree_grid_terms(c.h, self.h, ctypes.byref(grid), float(beta), ctypes.byref(tot)))
        return tot.value


class KnnIndex:
    """dwg_knn_index: exact kNN, (d², index) order."""

    def __init__(self, ctx: Context, pts, k_hint=15):
        self.ctx = ctx
        self.pts = ctx.f32(pts)
        self.n = self.pts.numel() // 3
        h = ctypes.c_void_p()
        ctx.check(ctx.L.dwg_knn_build(ctx.h, _ptr(self.pts), self.n, int(k_hint), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.ctx.L.dwg_knn_destroy(self.h)
            self.h = None

    def query(self, q, k, skip=None, qmask=None, d2=False):
        c = self.ctx
        q = c.f32(q)
        nq = q.numel() // 3
        idx = torch.full((nq, k), -1, device=c.device, dtype=torch.int32)
        dd = torch.empty((nq, k), device=c.device, dtype=torch.float32) if d2 else None
        sk = None if skip is None else skip.to(device=c.device, dtype=torch.uint8).contiguous()
        qm = None if qmask is None else qmask.to(device=c.device, dtype=torch.uint8).contiguous()
        c.check(c.L.dwg_knn_query(c.h, self.h, _ptr(q), nq, int(k), _ptr(sk), _ptr(qm), _ptr(idx), _ptr(dd)))
        return (idx, dd) if d2 else idx


def marching_cubes(ctx: Context, field, grid: DwgGrid, iso, tris=False):
    """→ (tris or None, centroids, normals) as device tensors, cube order."""
    field = ctx.f32(field)
    cnt = ctypes.c_int64()
    rc = ctx.L.dwg_marching_cubes(ctx.h, _ptr(field), ctypes.byref(grid), float(iso), 0, None, None, None,
                                  ctypes.byref(cnt))
    if rc not in (0, _lib.DWG_ERR_CAPACITY):
        ctx.check(rc)
    T = cnt.value
    t = torch.empty((T, 3, 3), device=ctx.device, dtype=torch.float32) if tris else None
    c = torch.empty((T, 3), device=ctx.device, dtype=torch.float32)
    n = torch.empty((T, 3), device=ctx.device, dtype=torch.float32)
    ctx.check(ctx.L.dwg_marching_cubes(ctx.h, _ptr(field), ctypes.byref(grid), float(iso), T, _ptr(t), _ptr(c),
                                       _ptr(n), ctypes.byref(cnt)))
    return t, c, n


def marching_cubes_any_order(ctx: Context, field, grid: DwgGrid, iso, capacity=1 << 16):
    """→ (centroids, normals) as device tensors, one pass, in no fixed order.

    Starts with `capacity` triangles and retries once at the reported count."""
    field = ctx.f32(field)
    cnt = ctypes.c_int64()
    cap = int(capacity)
    for _ in range(2):
        c = torch.empty((max(cap, 1), 3), device=ctx.device, dtype=torch.float32)
        n = torch.empty((max(cap, 1), 3), device=ctx.device, dtype=torch.float32)
        rc = ctx.L.dwg_marching_cubes_any_order(ctx.h, _ptr(field), ctypes.byref(grid), float(iso), cap, _ptr(c),
                                                _ptr(n), ctypes.byref(cnt))
        if rc != _lib.DWG_ERR_CAPACITY:
            ctx.check(rc)
            return c[:cnt.value], n[:cnt.value]
        cap = cnt.value  # short: the call reported the count
    ctx.check(rc)


def diffuse_normals(ctx: Context, index: KnnIndex, cent, tri_nrm, nrm, k=15, zero_k=10, zero_threshold=10):
    """In-place update of nrm (n×3 float32 device tensor); returns the zero count."""
    cent, tri_nrm = ctx.f32(cent), ctx.f32(tri_nrm)
    nz = ctypes.c_int64()
    ctx.check(ctx.L.dwg_diffuse_normals(ctx.h, index.h, _ptr(index.pts), index.n, _ptr(cent), _ptr(tri_nrm),
                                        cent.numel() // 3, int(k), int(zero_k), int(zero_threshold), _ptr(nrm),
                                        ctypes.byref(nz)))
    return nz.value


def radii(ctx: Context, index: KnnIndex, k=10, rmin=0.0015, rmax=0.015):
    out = torch.empty(index.n, device=ctx.device, dtype=torch.float32)
    ctx.check(ctx.L.dwg_radii(ctx.h, index.h, _ptr(index.pts), index.n, int(k), float(rmin), float(rmax), _ptr(out)))
    return out


def point_areas(ctx: Context, index: KnnIndex, k=15, nrm=None):
    out = torch.empty(index.n, device=ctx.device, dtype=torch.float32)
    nr = None if nrm is None else ctx.f32(nrm)
    ctx.check(ctx.L.dwg_point_areas(ctx.h, index.h, _ptr(index.pts), index.n, int(k), _ptr(nr), _ptr(out)))
    return out


def pca_normals(ctx: Context, index: KnnIndex, k=10, viewpoint=(0.0, 0.0, 0.0)):
    out = torch.empty((index.n, 3), device=ctx.device, dtype=torch.float32)
    vp = (ctypes.c_float * 3)(*[float(v) for v in viewpoint])
    ctx.check(ctx.L.dwg_pca_normals(ctx.h, index.h, _ptr(index.pts), index.n, int(k), vp, _ptr(out)))
    return out


def gauss_map(ctx: Context, pts, center):
    pts = ctx.f32(pts)
    out = torch.empty_like(pts)
    c = (ctypes.c_float * 3)(*[float(v) for v in center])
    ctx.check(ctx.L.dwg_gauss_map(ctx.h, _ptr(pts), pts.numel() // 3, c, _ptr(out)))
    return out


def transfer_normals(ctx: Context, index: KnnIndex, src_nrm, q, k=10):
    src_nrm, q = ctx.f32(src_nrm), ctx.f32(q)
    out = torch.empty_like(q)
    ctx.check(ctx.L.dwg_transfer_normals(ctx.h, index.h, _ptr(src_nrm), _ptr(q), q.numel() // 3, int(k), _ptr(out)))
    return out


def outlier_filter(ctx: Context, area, threshold):
    ctx.check(ctx.L.dwg_outlier_filter(ctx.h, _ptr(area), area.numel(), float(threshold)))
    return area


class Mlp:
    """dwg_mlp: the learned-initialisation network (BF16 on the matrix cores)."""

    def __init__(self, ctx: Context, seed: int = 0, path: str | None = None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        if path is None:
            ctx.check(ctx.L.dwg_mlp_create(ctx.h, int(seed), ctypes.byref(h)))
        else:
            ctx.check(ctx.L.dwg_mlp_load(ctx.h, path.encode(), ctypes.byref(h)))
        self.h = h
        self.n_params = int(ctx.L.dwg_mlp_num_params())

    def __del__(self):
        if getattr(self, "h", None):
            self.ctx.L.dwg_mlp_destroy(self.h)
            self.h = None

    def save(self, path):
        self.ctx.check(self.ctx.L.dwg_mlp_save(self.ctx.h, self.h, path.encode()))

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx.L.dwg_mlp_get(self.ctx.h, self.h, out.ctypes.data, None))
        return out

    def grad(self):
        out = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx.L.dwg_mlp_get(self.ctx.h, self.h, None, out.ctypes.data))
        return out

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, np.float32)
        assert t.size == self.n_params
        self.ctx.check(self.ctx.L.dwg_mlp_set(self.ctx.h, self.h, t.ctypes.data))

    def patches(self, index: KnnIndex):
        out = torch.empty((index.n, 48), device=self.ctx.device, dtype=torch.int16)
        self.ctx.check(self.ctx.L.dwg_mlp_patches(self.ctx.h, index.h, _ptr(index.pts), index.n, _ptr(out)))
        return out

    def patches_from(self, index: KnnIndex, q):
        """Patches of the points q over the points of `index`."""
        q = self.ctx.f32(q)
        nq = q.numel() // 3
        out = torch.empty((nq, 48), device=self.ctx.device, dtype=torch.int16)
        self.ctx.check(self.ctx.L.dwg_mlp_patches_from(self.ctx.h, index.h, _ptr(index.pts), index.n, _ptr(q), nq,
                                                       _ptr(out)))
        return out

    def init_normals(self, pts, center):
        """Learned initialisation (dwg_mlp_init): normal line oriented away
        from `center`."""
        p = self.ctx.f32(pts)
        n = p.numel() // 3
        out = torch.empty((n, 3), device=self.ctx.device, dtype=torch.float32)
        c = (ctypes.c_float * 3)(*[float(v) for v in center])
        self.ctx.check(self.ctx.L.dwg_mlp_init(self.ctx.h, self.h, _ptr(p), n, c, _ptr(out)))
        return out

    def forward(self, patches, normalize=True):
        n = patches.shape[0]
        out = torch.empty((n, 3), device=self.ctx.device, dtype=torch.float32)
        self.ctx.check(self.ctx.L.dwg_mlp_forward(self.ctx.h, self.h, _ptr(patches), n, 1 if normalize else 0,
                                                  _ptr(out)))
        return out

    def loss_grad(self, patches, target, line=False):
        """Loss and gradient: 1 − ŷ·t (oriented) or, with line=True,
        1 − (ŷ·t)² (normal line, sign free)."""
        loss = ctypes.c_double()
        t = target.to(device=self.ctx.device, dtype=torch.float32).contiguous()
        self.ctx.check(self.ctx.L.dwg_mlp_loss_grad(self.ctx.h, self.h, _ptr(patches), _ptr(t), patches.shape[0],
                                                    1 if line else 0, ctypes.byref(loss)))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx.check(self.ctx.L.dwg_mlp_adam(self.ctx.h, self.h, float(lr), float(beta1), float(beta2), float(eps)))

    def augment(self, patches, target, index, rotation, out_patches, out_target):
        R = (ctypes.c_float * 9)(*[float(v) for v in np.asarray(rotation, np.float32).ravel()])
        self.ctx.check(self.ctx.L.dwg_mlp_augment(self.ctx.h, _ptr(patches), _ptr(target), _ptr(index),
                                                  index.shape[0], R, _ptr(out_patches), _ptr(out_target)))


def gemm_bf16(ctx: Context, a16, sa, b16, sb, m, n, k):
    """C = A·B (FP32) with BF16 operands given as int16 bit patterns and
    element strides sa = (s_ai, s_ak), sb = (s_bk, s_bj)."""
    c = torch.empty((m, n), device=ctx.device, dtype=torch.float32)
    ctx.check(ctx.L.dwg_gemm_bf16(ctx.h, _ptr(a16), int(sa[0]), int(sa[1]), _ptr(b16), int(sb[0]), int(sb[1]),
                                  _ptr(c), int(m), int(n), int(k)))
    return c
