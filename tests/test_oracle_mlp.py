"""Pins the FP64 MLP oracle: analytic gradient vs central finite differences,
GELU vs its erf form, and the patch definition's invariances."""
import math

import numpy as np
import pytest

from oracle import mlp as M


def _theta(seed=0):
    rng = np.random.default_rng(seed)
    th = []
    for s in M.SIZES:
        fan_in = s[1] if len(s) == 2 else 64
        th.append(rng.uniform(-1, 1, s).ravel() / math.sqrt(fan_in))
    return np.concatenate(th)


def test_param_count():
    assert M.N_PARAMS == 23171


def test_gelu_close_to_erf_form():
    x = np.linspace(-6, 6, 241)
    erf = np.array([0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in x])
    assert np.abs(M.gelu(x) - erf).max() < 1e-3
    h = 1e-6
    assert np.allclose(M.gelu_grad(x), (M.gelu(x + h) - M.gelu(x - h)) / (2 * h), atol=1e-8)


@pytest.mark.parametrize("line", [False, True])
def test_gradient_matches_finite_differences(line):
    rng = np.random.default_rng(1)
    th = _theta(2)
    X = rng.normal(size=(24, 48))
    T = rng.normal(size=(24, 3))
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    loss, g = M.loss_grad(th, X, T, line=line)
    assert loss == M.loss_only(th, X, T, line=line)
    h = 1e-6
    for i in rng.choice(M.N_PARAMS, 60, replace=False).tolist() + [M.N_PARAMS - 1, 0, 6144, 6144 + 128]:
        e = np.zeros_like(th)
        e[i] = h
        fd = (M.loss_only(th + e, X, T, line=line) - M.loss_only(th - e, X, T, line=line)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_line_loss_is_sign_free():
    th = _theta(3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(16, 48))
    T = rng.normal(size=(16, 3))
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    flip = np.where(rng.uniform(size=(16, 1)) < 0.5, -1.0, 1.0)
    a, ga = M.loss_grad(th, X, T, line=True)
    b, gb = M.loss_grad(th, X, T * flip, line=True)
    assert a == pytest.approx(b, rel=1e-12) and np.allclose(ga, gb, rtol=1e-10, atol=1e-15)
    # a perfect predictor of the line (either sign) has zero loss
    y, _ = M.forward(th, X)
    yh = y / np.linalg.norm(y, axis=1, keepdims=True)
    assert M.loss_only(th, X, yh * flip, line=True) == pytest.approx(0.0, abs=1e-12)


def test_patch_invariances():
    rng = np.random.default_rng(3)
    P = rng.uniform(0, 1, (300, 3)).astype(np.float32)
    a = M.patches(P)
    b = M.patches((P * np.float32(4.0) + np.float32(2.0)).astype(np.float32))
    assert np.allclose(a, b, atol=1e-5)
    nrm = np.linalg.norm(a.reshape(-1, 16, 3), axis=2)
    assert np.all(np.diff(nrm, axis=1) >= -1e-6)  # distance order
    assert np.allclose(nrm[:, -1], 1.0, atol=1e-6)  # scaled by the 16th distance


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.3, 65504.0, 1e-30], np.float32)
    r = M.to_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_patches_from_matches_self_patches():
    # over its own points (with exact duplicates) the general form equals the
    # self-dropping form, and a point off the set keeps its 16 nearest
    rng = np.random.default_rng(9)
    P = rng.normal(size=(300, 3)).astype(np.float32)
    P = np.concatenate([P, P[:20]])
    assert np.array_equal(M.patches(P), M.patches_from(P, P))
    q = (P[:5] + 0.01).astype(np.float32)
    X = M.patches_from(P, q).reshape(5, 16, 3)
    assert np.allclose(np.linalg.norm(X, axis=2).max(1), 1.0, atol=1e-6)
