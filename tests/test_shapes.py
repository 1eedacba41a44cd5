"""The synthetic workloads: sizes, determinism, points on the stated
surfaces, unit normals, noise of the stated size."""
import numpy as np

from dwg_amd import shapes


def test_sphere_and_torus():
    P, N = shapes.sphere(3000, 1)
    assert P.shape == N.shape == (3000, 3)
    assert np.allclose(np.linalg.norm(P, axis=1), 1.0, atol=1e-6)
    assert np.allclose(P, N, atol=1e-6)
    P, N = shapes.torus(4000, 2, noise=0.0)
    q = np.sqrt(P[:, 0] ** 2 + P[:, 1] ** 2) - 1.0
    assert np.allclose(q ** 2 + P[:, 2] ** 2, 0.35 ** 2, atol=1e-5)
    assert np.allclose(np.linalg.norm(N, axis=1), 1.0, atol=1e-6)
    # density ratio 1..8 along the major angle: the far side (φ = π) holds more
    phi = np.arctan2(P[:, 1], P[:, 0])
    assert np.mean(np.abs(phi) > 2.5) > 3 * np.mean(np.abs(phi) < 0.64)
    P2, _ = shapes.torus(4000, 2, noise=0.0)
    assert np.array_equal(P, P2)


def test_rotations_orthonormal():
    rng = np.random.default_rng(3)
    for _ in range(20):
        R = shapes._random_rotation(rng)
        assert np.abs(R @ R.T - np.eye(3)).max() < 1e-14 and abs(np.linalg.det(R) - 1) < 1e-14


def test_union_and_large():
    P, N = shapes.union(5000, 8, 11)
    P2, _ = shapes.union(5000, 8, 11)
    assert P.shape == (5000, 3) and np.array_equal(P, P2)
    assert np.allclose(np.linalg.norm(N, axis=1), 1.0, atol=1e-5)
    P, N, C = shapes.large(20_000, 5, m=16, shells=4, clean=True)
    assert len(P) <= 20_000 and len(P) > 19_000
    d = np.linalg.norm(P - C, axis=1)
    diag = np.linalg.norm(C.max(0) - C.min(0))
    assert 0.5 * 0.001 * diag < np.sqrt((d ** 2).mean() / 3) < 1.5 * 0.001 * diag


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def test_headline_workload_checksum():
    # the 1M-point union (seed 3): the same bits on every host (fixed-order
    # elementwise arithmetic, no BLAS in the sampler)
    P, N = shapes.union(1_000_000, 32, 3)
    assert len(P) == 1_000_000
    assert (_sha(P), _sha(N)) == ("447f76d592ec930c", "14c40acdce53e93d")


def test_large_workload_checksum():
    # the 5M-point large configuration (seed 5): point count and bits pinned
    P, N = shapes.large(5_000_000, 5)
    assert len(P) == 4_996_750
    assert _sha(P) == "1663876fdafd7ced"
