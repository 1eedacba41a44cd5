"""dwg_gwn_direct_noalloc against dwg_gwn_direct, both called through the C ABI.

The allocation-free entry point runs the same kernels on sources packed into
the library's fixed device array, so its results must equal dwg_gwn_direct's
bit for bit.  Two calls in a row, the second with fewer sources and queries
(the array keeps the first call's contents), for the plain, screened (with
per-source radii) and gradient forms; sizes past the limit are refused.
"""
import ctypes

import numpy as np
import pytest
import torch

from dwg_amd import _lib, shapes
from dwg_amd._lib import DWG_KAHAN, DWG_SCREENED, DWG_USE_RADII

pytestmark = pytest.mark.gpu


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _inputs(n, nq, seed):
    P, N = shapes.sphere(n, seed)
    rng = np.random.default_rng(seed)
    A = rng.uniform(1e-4, 3e-3, n).astype(np.float32)
    R = rng.uniform(0.0015, 0.015, n).astype(np.float32)
    Q = np.concatenate([rng.uniform(-1.4, 1.4, (nq - nq // 4, 3)), P[: nq // 4]]).astype(np.float32)
    dev = torch.device("cuda", 0)
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (P, N, A, R, Q)]


def _call(ctx, fn, P, N, A, R, Q, flags, sigma, grad):
    nq = Q.shape[0]
    w = torch.full((nq,), float("nan"), device=Q.device)
    g = torch.full((nq, 3), float("nan"), device=Q.device) if grad else None
    rc = fn(ctx.h, _p(P), _p(N), _p(A), _p(R if flags & DWG_USE_RADII else None), P.shape[0], _p(Q), nq,
            ctypes.c_float(sigma), flags, _p(w), _p(g))
    ctx.check(rc)
    torch.cuda.synchronize()
    return w.cpu().numpy(), None if g is None else g.cpu().numpy()


@pytest.mark.parametrize("flags,sigma,grad", [
    (0, 0.0, False),
    (DWG_KAHAN, 0.0, True),
    (DWG_SCREENED | DWG_USE_RADII | DWG_KAHAN, 10.0, False),
    (DWG_SCREENED, 20.0, False),
])
def test_noalloc_matches_direct_twice(ctx, flags, sigma, grad):
    L = ctx.L
    for n, nq, seed in [(5000, 3000, 1), (1234, 777, 2)]:  # the second call is smaller
        P, N, A, R, Q = _inputs(n, nq, seed)
        w_ref, g_ref = _call(ctx, L.dwg_gwn_direct, P, N, A, R, Q, flags, sigma, grad)
        w, g = _call(ctx, L.dwg_gwn_direct_noalloc, P, N, A, R, Q, flags, sigma, grad)
        print(f"noalloc flags={flags} sigma={sigma} grad={grad} n={n} nq={nq}: "
              f"max |w| {np.abs(w_ref).max():.3g}", flush=True)
        assert np.isfinite(w_ref).all()
        np.testing.assert_array_equal(w, w_ref)
        if grad:
            np.testing.assert_array_equal(g, g_ref)


def test_noalloc_full_size_and_limit(ctx):
    L = ctx.L
    n = _lib.DWG_NOALLOC_MAX_SOURCES
    P, N, A, R, Q = _inputs(n, 512, 3)
    w_ref, _ = _call(ctx, L.dwg_gwn_direct, P, N, A, R, Q, DWG_KAHAN, 0.0, False)
    w, _ = _call(ctx, L.dwg_gwn_direct_noalloc, P, N, A, R, Q, DWG_KAHAN, 0.0, False)
    np.testing.assert_array_equal(w, w_ref)
    # one source more than the array holds: refused on the host, nothing launched
    w = torch.zeros(Q.shape[0], device=Q.device)
    rc = L.dwg_gwn_direct_noalloc(ctx.h, _p(P), _p(N), _p(A), None, n + 1, _p(Q), Q.shape[0], ctypes.c_float(0.0),
                                  0, _p(w), None)
    assert rc == _lib.DWG_ERR_ARG
    assert b"exceed" in L.dwg_ctx_error(ctx.h)
