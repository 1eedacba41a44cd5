"""CPU oracle for the DWG path — test infrastructure, never shipped.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  It wraps liboracle.so (dwg_oracle.cpp, a FP64 C++
restatement of the reference's algorithm; see that file's header for the
function-by-function reference map) plus a few numpy restatements.

Every function here is pinned by tests/test_oracle.py against closed forms,
brute force or invariants; the one exception is listed there as "parity
unpinned".
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double


def build() -> str:
    """Compile liboracle.so with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.orc_set_threads.argtypes = [_int]
    L.orc_get_threads.restype = _int
    L.orc_tree_build.argtypes = [_f64p, _i64, _int, _int]
    L.orc_tree_build.restype = _vp
    L.orc_tree_free.argtypes = [_vp]
    L.orc_tree_num_nodes.argtypes = [_vp]
    L.orc_tree_num_nodes.restype = _i64
    L.orc_tree_topology.argtypes = [_vp, _i32p, _i32p, _i32p]
    L.orc_tree_node_points.argtypes = [_vp, _i64, ctypes.c_void_p]
    L.orc_tree_node_points.restype = _i64
    L.orc_tree_precompute.argtypes = [_vp, _f64p, _f64p]
    L.orc_tree_refresh_ec.argtypes = [_vp, _f64p, _f64p]
    L.orc_tree_moments.argtypes = [_vp, _f64p, _f64p, _f64p]
    L.orc_tree_eval.argtypes = [_vp, _f64p, _f64p, _vp, _f64p, _i64, _dbl, _int, _dbl, _int, _int,
                                _f64p, _vp]
    L.orc_direct.argtypes = [_f64p, _f64p, _f64p, _vp, _i64, _f64p, _i64, _int, _dbl, _int, _f64p]
    L.orc_direct_grad.argtypes = [_f64p, _f64p, _f64p, _vp, _i64, _f64p, _i64, _int, _dbl, _int, _f64p, _f64p]
    L.orc_tree_eval_grad.argtypes = [_vp, _f64p, _f64p, _vp, _f64p, _i64, _dbl, _int, _dbl, _int, _f64p, _f64p]
    L.orc_voxel_grid.argtypes = [_f64p, _i64, _dbl, _int, _f64p, ctypes.POINTER(_dbl), _i32p]
    L.orc_mc.argtypes = [_f64p, _i32p, _f64p, _dbl, _dbl, _i64, _vp, _vp, _vp]
    L.orc_mc.restype = _i64
    L.orc_mc_table.argtypes = [_u64p]
    L.orc_knn_brute.argtypes = [_f32p, _i64, _f32p, _i64, _int, _vp, _i32p, _vp]
    L.orc_knn_grid.argtypes = [_f32p, _i64, _f32p, _i64, _int, _vp, _i32p]
    L.orc_radii.argtypes = [_f32p, _i64, _int, _dbl, _dbl, _f64p]
    L.orc_gauss_map.argtypes = [_f64p, _i64, _f64p]
    L.orc_sigma_schedule.argtypes = [_int, _dbl]
    L.orc_sigma_schedule.restype = _dbl
    L.orc_dwg_iterate.argtypes = [_f64p, _f64p, _f64p, _vp, _i64, _i32p, _f64p, _dbl, _int, _int, _dbl,
                                  _dbl, _int, _int, _f64p, _i64p, _i64p]
    _LIB = L
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def threads() -> int:
    return int(lib().orc_get_threads())


class Octree:
    """octreeDWG + order-0 precompute (octree.cpp:171-335, fast_winding_number.cpp:9-120)."""

    def __init__(self, P, max_depth=10, leaf_size=1):
        self.P = _c(P, np.float64).reshape(-1, 3)
        self.h = lib().orc_tree_build(self.P, len(self.P), int(max_depth), int(leaf_size))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    @property
    def num_nodes(self) -> int:
        return int(lib().orc_tree_num_nodes(self.h))

    def topology(self):
        m = self.num_nodes
        ch = np.zeros((m, 8), np.int32)
        cnt = np.zeros(m, np.int32)
        dep = np.zeros(m, np.int32)
        lib().orc_tree_topology(self.h, ch, cnt, dep)
        return ch, cnt, dep

    def node_points(self, node):
        n = lib().orc_tree_node_points(self.h, int(node), None)
        out = np.zeros(n, np.int32)
        lib().orc_tree_node_points(self.h, int(node), out.ctypes.data)
        return out

    def precompute(self, N, A):
        self.N = _c(N, np.float64).reshape(-1, 3)
        self.A = _c(A, np.float64).reshape(-1)
        lib().orc_tree_precompute(self.h, self.N, self.A)

    def refresh(self, N, A):
        self.N = _c(N, np.float64).reshape(-1, 3)
        self.A = _c(A, np.float64).reshape(-1)
        lib().orc_tree_refresh_ec(self.h, self.N, self.A)

    def moments(self):
        m = self.num_nodes
        cm = np.zeros((m, 3))
        r = np.zeros(m)
        ec = np.zeros((m, 3))
        lib().orc_tree_moments(self.h, cm, r, ec)
        return cm, r, ec

    def eval(self, Q, beta, kind="screened", sigma=10.0, radii=None, float_decisions=True,
             count_terms=False, point_areas=None):
        """Barnes-Hut sums at Q.  point_areas (optional) replaces the areas of
        the point terms only; far nodes keep the moments of the last
        precompute/refresh, as the reference's launch takes the areas apart
        from the node moments (TurboLibiGWN.cu:912-922, TurboDWG.cu:870-875)."""
        Q = _c(Q, np.float64).reshape(-1, 3)
        out = np.zeros(len(Q))
        terms = np.zeros(len(Q), np.int64) if count_terms else None
        rad = None if radii is None else _c(radii, np.float64)
        A = self.A if point_areas is None else _c(point_areas, np.float64).reshape(-1)
        lib().orc_tree_eval(self.h, self.N, A, _ptr(rad), Q, len(Q), float(beta),
                            0 if kind == "plain" else 1, float(sigma), 1 if rad is not None else 0,
                            1 if float_decisions else 0, out, _ptr(terms))
        return (out, terms) if count_terms else out

    def eval_grad(self, Q, beta, kind="screened", sigma=10.0, radii=None):
        """Barnes-Hut value and gradient: the term set of eval (float far
        decisions), each term with its exact gradient (no reference
        counterpart)."""
        Q = _c(Q, np.float64).reshape(-1, 3)
        w = np.zeros(len(Q))
        g = np.zeros((len(Q), 3))
        rad = None if radii is None else _c(radii, np.float64)
        lib().orc_tree_eval_grad(self.h, self.N, self.A, _ptr(rad), Q, len(Q), float(beta),
                                 0 if kind == "plain" else 1, float(sigma), 1 if rad is not None else 0, w, g)
        return w, g


def direct(P, N, A, Q, kind="plain", sigma=10.0, radii=None):
    """FP64 all-pairs sum (fast_winding_number.cpp:313-323; screened: TurboLibiGWN.cu:544-585)."""
    P = _c(P, np.float64).reshape(-1, 3)
    N = _c(N, np.float64).reshape(-1, 3)
    A = _c(A, np.float64).reshape(-1)
    Q = _c(Q, np.float64).reshape(-1, 3)
    rad = None if radii is None else _c(radii, np.float64)
    out = np.zeros(len(Q))
    lib().orc_direct(P, N, A, _ptr(rad), len(P), Q, len(Q), 0 if kind == "plain" else 1, float(sigma),
                     1 if rad is not None else 0, out)
    return out


def direct_grad(P, N, A, Q, kind="plain", sigma=10.0, radii=None):
    """FP64 all-pairs sum and its exact gradient with respect to the query
    point, plain or screened term (no reference counterpart)."""
    P = _c(P, np.float64).reshape(-1, 3)
    N = _c(N, np.float64).reshape(-1, 3)
    A = _c(A, np.float64).reshape(-1)
    Q = _c(Q, np.float64).reshape(-1, 3)
    rad = None if radii is None else _c(radii, np.float64)
    w = np.zeros(len(Q))
    g = np.zeros((len(Q), 3))
    lib().orc_direct_grad(P, N, A, _ptr(rad), len(P), Q, len(Q), 0 if kind == "plain" else 1, float(sigma),
                          1 if rad is not None else 0, w, g)
    return w, g


def voxel_grid(V, offset, s):
    """igl::voxel_grid(V, offset, s, pad=1) → (origin[3], h, res[3])."""
    V = _c(V, np.float64).reshape(-1, 3)
    origin = np.zeros(3)
    h = _dbl(0)
    res = np.zeros(3, np.int32)
    lib().orc_voxel_grid(V, len(V), float(offset), int(s), origin, ctypes.byref(h), res)
    return origin, float(h.value), res


def grid_points(origin, h, res):
    i, j, k = np.meshgrid(np.arange(res[0]), np.arange(res[1]), np.arange(res[2]), indexing="ij")
    pts = np.stack([origin[0] + i * h, origin[1] + j * h, origin[2] + k * h], -1)
    return np.ascontiguousarray(pts.transpose(2, 1, 0, 3).reshape(-1, 3))


def marching_cubes(field, res, origin, h, iso):
    """→ (tris T×3×3, centroids T×3, normals T×3) in cube order."""
    field = _c(field, np.float64).reshape(-1)
    res = _c(res, np.int32)
    origin = _c(origin, np.float64)
    T = lib().orc_mc(field, res, origin, float(h), float(iso), 0, None, None, None)
    tris = np.zeros((T, 3, 3))
    c = np.zeros((T, 3))
    n = np.zeros((T, 3))
    lib().orc_mc(field, res, origin, float(h), float(iso), T, tris.ctypes.data, c.ctypes.data, n.ctypes.data)
    return tris, c, n


def mc_table():
    rows = np.zeros(256, np.uint64)
    lib().orc_mc_table(rows)
    return rows


def knn_brute(P, Q, k, skip=None, return_d2=False):
    P = _c(P, np.float32).reshape(-1, 3)
    Q = _c(Q, np.float32).reshape(-1, 3)
    idx = np.zeros((len(Q), k), np.int32)
    d2 = np.zeros((len(Q), k), np.float32) if return_d2 else None
    sk = None if skip is None else _c(skip, np.uint8)
    lib().orc_knn_brute(P, len(P), Q, len(Q), int(k), _ptr(sk), idx, _ptr(d2))
    return (idx, d2) if return_d2 else idx


def knn_grid(P, Q, k, skip=None):
    P = _c(P, np.float32).reshape(-1, 3)
    Q = _c(Q, np.float32).reshape(-1, 3)
    idx = np.zeros((len(Q), k), np.int32)
    sk = None if skip is None else _c(skip, np.uint8)
    lib().orc_knn_grid(P, len(P), Q, len(Q), int(k), _ptr(sk), idx)
    return idx


def radii(P, k=10, rmin=0.0015, rmax=0.015):
    P = _c(P, np.float32).reshape(-1, 3)
    out = np.zeros(len(P))
    lib().orc_radii(P, len(P), int(k), float(rmin), float(rmax), out)
    return out


def gauss_map(P):
    P = _c(P, np.float64).reshape(-1, 3)
    N = np.zeros_like(P)
    lib().orc_gauss_map(P, len(P), N)
    return N


def pca_normals(P, k=10, viewpoint=None):
    """PCA normals (PCL NormalEstimation with setKSearch(k) and
    setViewPoint, DWGProcess.h:152-200): the smallest-eigenvalue eigenvector
    of the covariance of each point's k nearest neighbours (the point itself
    included, float32 (d², index) ranking), FP64 LAPACK eigen-solver, turned
    so that n·(viewpoint − p) ≥ 0.  viewpoint defaults to the bounding-box
    centre of the float32 points."""
    P32 = _c(P, np.float32).reshape(-1, 3)
    if viewpoint is None:
        viewpoint = (P32.min(0) + P32.max(0)) / np.float32(2)
    idx = knn_brute(P32, P32, k)
    Q = P32.astype(np.float64)[idx]                      # n × k × 3
    D = Q - Q.mean(1, keepdims=True)
    C = np.einsum("nki,nkj->nij", D, D)
    _, U = np.linalg.eigh(C)
    N = U[:, :, 0]
    t = np.asarray(viewpoint, np.float64) - P32.astype(np.float64)
    N[(N * t).sum(1) < 0] *= -1
    return N


def point_areas(P, I, N=None):
    """Voronoi point areas, igl::copyleft::cgal::point_areas
    (deps/libigl/include/igl/copyleft/cgal/point_areas.cpp:41-205), FP64:
    neighbours I[i] (the first is the point itself), kept where
    n_first·n_j > 0; ≤ 2 kept → 0; PCA plane of the kept neighbours (SVD
    scores, first two columns); 2-D Delaunay triangulation (here Qhull via
    scipy instead of CGAL); every triangle at the first point adds its
    Voronoi share (half the triangle if obtuse there, a quarter if obtuse
    elsewhere, else the circumcentre split); non-finite → 0."""
    from scipy.spatial import Delaunay
    from scipy.spatial import QhullError
    P = np.asarray(P, np.float64).reshape(-1, 3)
    I = np.asarray(I)
    A = np.zeros(len(I))
    for i in range(len(I)):
        nb = I[i][I[i] >= 0]
        X = P[nb]
        if N is not None and len(nb) > 1:
            Nn = np.asarray(N, np.float64)[nb]
            X = X[(Nn @ Nn[0]) > 0]
        if len(X) <= 2:
            continue
        C = X - X.mean(0)
        U, S, _ = np.linalg.svd(C, full_matrices=False)
        plane = (U * S)[:, :2]
        try:
            tri = Delaunay(plane)
        except QhullError:
            continue
        acc = 0.0
        for f in tri.simplices:
            if 0 not in f:
                continue
            f = list(f)
            X0 = f.index(0)
            Y, Z = f[(X0 + 1) % 3], f[(X0 + 2) % 3]
            x = np.linalg.norm(plane[Y] - plane[Z])
            y = np.linalg.norm(plane[0] - plane[Z])
            z = np.linalg.norm(plane[Y] - plane[0])
            cx = (z * z + y * y - x * x) / (2 * y * z)
            cy = (z * z + x * x - y * y) / (2 * x * z)
            cz = (x * x + y * y - z * z) / (2 * y * x)
            bary = np.array([x * cx, y * cy, z * cz])
            bary /= bary.sum()
            full = 0.25 * np.sqrt((x + (y + z)) * (z - (x - y)) * (z + (x - y)) * (x + (y - z)))
            if cx < 0:
                acc += 0.5 * full
            elif cy < 0 or cz < 0:
                acc += 0.25 * full
            else:
                acc += (bary[1] + bary[2]) * full / 2
        A[i] = acc if np.isfinite(acc) else 0.0
    return A


def outlier_filter(A, threshold):
    """OutlierFilter (TurboDWG.cu:991-1009): sort the areas in descending order
    (thrust::sort_by_key with thrust::greater, a stable radix sort: equal
    areas keep ascending index order) and divide the first
    ceil(float(n)·threshold) by 1000 in float (AsFiltertoZero, :442-454)."""
    A = np.array(A, dtype=np.float32, copy=True).reshape(-1)
    m = int(np.ceil(np.float32(len(A)) * np.float32(threshold)))
    if m <= 0:
        return A
    order = np.argsort(-A, kind="stable")
    A[order[:m]] = A[order[:m]] / np.float32(1000.0)
    return A


def normal_change(old, new):
    """Convergence residual of a diffusion step (new; no reference
    counterpart): mean over the sources of 1 − |cos ∠(old_i, new_i)|, a zero
    normal counting 1, in FP64."""
    a = np.asarray(old, np.float64).reshape(-1, 3)
    b = np.asarray(new, np.float64).reshape(-1, 3)
    la, lb = (a * a).sum(1), (b * b).sum(1)
    ok = (la > 0) & (lb > 0)
    c = np.ones(len(a))
    c[ok] = 1.0 - np.abs((a[ok] * b[ok]).sum(1)) / np.sqrt(la[ok] * lb[ok])
    return float(c.mean()) if len(a) else 0.0


def sigma_schedule(it, sigma_in):
    return float(lib().orc_sigma_schedule(int(it), float(sigma_in)))


def dwg_iterate(P, N0, A, res, origin, h, iters, sigma=10.0, beta=2.3, max_depth=10, leaf_size=1,
                radii=None):
    """The DWG loop over a fixed source set (TurboDWG.cu:516-652) in FP64.

    Returns (normals, w_mean[iters], ntri[iters], nzero[iters])."""
    P = _c(P, np.float64).reshape(-1, 3)
    N = _c(N0, np.float64).reshape(-1, 3).copy()
    A = _c(A, np.float64).reshape(-1)
    rad = None if radii is None else _c(radii, np.float64)
    wm = np.zeros(iters)
    nt = np.zeros(iters, np.int64)
    nz = np.zeros(iters, np.int64)
    lib().orc_dwg_iterate(P, N, A, _ptr(rad), len(P), _c(res, np.int32), _c(origin, np.float64), float(h),
                          int(max_depth), int(leaf_size), float(beta), float(sigma), int(iters),
                          1 if rad is not None else 0, wm, nt, nz)
    return N, wm, nt, nz


def octree_leaves(P, max_depth=10, min_num=0, widths=False):
    """GetOctreeLeaves (DWGProcess.h:383-428): breadth-first over octreeDWG,
    the centres of leaf cells holding more than min_num points.
    Returns (centres [m,3], point index lists[, cell widths [m]])."""
    from collections import deque
    P = _c(P, np.float64).reshape(-1, 3)
    t = Octree(P, max_depth, 1)
    ch, cnt, _ = t.topology()
    lo, hi = P.min(0), P.max(0)
    centres = {0: ((lo + hi) / 2, (hi - lo).max())}
    q = deque([0])
    out, pts, wids = [], [], []
    while q:
        i = q.popleft()
        c, w = centres[i]
        if ch[i, 0] == -1:
            if cnt[i] > min_num:
                out.append(c)
                pts.append(t.node_points(i))
                wids.append(w)
            continue
        for k in range(8):
            kid = int(ch[i, k])
            sign = np.array([1.0 if k & 1 else -1.0, 1.0 if k & 2 else -1.0, 1.0 if k & 4 else -1.0])
            centres[kid] = (c + sign * w / 4, w / 2)
            if cnt[kid] > 0:
                q.append(kid)
    if widths:
        return np.array(out), pts, np.array(wids)
    return np.array(out), pts


def leaf_normals(P, N, max_depth=10, min_num=0):
    """SmoothNormalstoOctreeLeaves (DWGProcess.h:430-472): component k of a
    leaf's normal is Σ_j exp(−((P_jk − C_k)/W)²/4)·N_jk over the leaf's
    points in octree order (C the leaf centre, W its width, octree.cpp
    CN/W); then .normalized() (DWGProcess.h:537) divides the whole matrix by
    its Frobenius norm."""
    P = _c(P, np.float64).reshape(-1, 3)
    N = _c(N, np.float64).reshape(-1, 3)
    C, pts, W = octree_leaves(P, max_depth, min_num, widths=True)
    out = np.zeros((len(C), 3))
    for i, (c, idx, w) in enumerate(zip(C, pts, W)):
        x = (P[idx] - c) / w
        out[i] = (np.exp(-x * x / 4.0) * N[idx]).sum(0)
    f = np.sqrt((out * out).sum())
    return out / f if f > 0 else out
