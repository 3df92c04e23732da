"""Pins of the oracle's affine functions (DLT fit, FP64 resample) against
exact recovery, order invariance, scipy.ndimage and exact integer cases."""
import numpy as np
import pytest
from scipy import ndimage


def rand_affine(rng, scale=0.15):
    M = np.eye(4)
    M[:3, :3] += rng.uniform(-scale, scale, (3, 3))
    M[:3, 3] = rng.uniform(-10, 10, 3)
    return M


def apply(M, P):
    return P.astype(np.float64) @ M[:3, :3].T + M[:3, 3]


def test_affine_fit_recovers_matrix(oracle, rng):
    for n in (4, 6, 30):
        # Points exactly representable in float32 on both sides.
        A = rng.integers(0, 200, (n, 3)).astype(np.float32)
        M = rand_affine(rng)
        M[:3, :3] = np.round(M[:3, :3] * 8) / 8
        M[:3, 3] = np.round(M[:3, 3])
        B = apply(M, A).astype(np.float32)
        assert np.array_equal(B.astype(np.float64), apply(M, A))
        Mf = oracle.affine_fit(A, B)
        assert np.allclose(Mf, M, atol=1e-9), (n, Mf - M)


def test_affine_fit_order_invariant(oracle, rng):
    A = rng.uniform(0, 100, (25, 3)).astype(np.float32)
    B = (apply(rand_affine(rng), A) + rng.normal(0, 0.3, (25, 3))).astype(np.float32)
    M1 = oracle.affine_fit(A, B)
    perm = rng.permutation(25)
    M2 = oracle.affine_fit(A[perm], B[perm])
    assert np.allclose(M1, M2, atol=1e-10)
    # Least squares: residual no larger than for perturbed matrices.
    r1 = np.abs(apply(M1, A) - B).sum()
    for _ in range(5):
        Mp = M1.copy()
        Mp[:3] += rng.normal(0, 1e-3, (3, 4))
        assert np.abs(apply(Mp, A) - B).sum() >= r1 - 1e-6


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_affine_identity_and_integer_shift(oracle, rng, dtype):
    hi = np.iinfo(dtype).max
    v = rng.integers(0, hi + 1, (9, 11, 13)).astype(dtype)
    assert np.array_equal(oracle.affine_resample(v, np.eye(4)), v)
    M = np.eye(4)
    M[:3, 3] = (2, -1, 3)  # subject = target + t
    out = oracle.affine_resample(v, M)
    expect = np.zeros_like(v)
    expect[:-3, 1:, :-2] = v[3:, :-1, 2:]
    assert np.array_equal(out, expect)


def test_affine_matches_scipy_with_truncation(oracle, rng):
    v = rng.integers(0, 256, (14, 16, 18)).astype(np.uint8)
    M = rand_affine(rng, 0.1)
    M[:3, 3] = rng.uniform(-1.5, 1.5, 3)
    out = oracle.affine_resample(v, M)
    # scipy works in (z, y, x) index order.
    P = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], float)
    Ms = P @ M[:3, :3] @ P
    off = P @ M[:3, 3]
    ref = ndimage.affine_transform(v.astype(np.float64), Ms, offset=off, order=1, mode="constant", cval=0.0)
    z, y, x = np.meshgrid(*[np.arange(s) for s in v.shape], indexing="ij")
    pos = np.einsum("ab,bzyx->azyx", M[:3, :3], np.stack([x, y, z]).astype(float)) + M[:3, 3, None, None, None]
    inside = ((pos[0] >= 0) & (pos[0] <= v.shape[2] - 1) & (pos[1] >= 0) & (pos[1] <= v.shape[1] - 1)
              & (pos[2] >= 0) & (pos[2] <= v.shape[0] - 1))
    near_int = np.abs(ref - np.round(ref)) < 1e-7
    assert np.all(out[~inside] == 0)
    ok = inside & ~near_int
    assert np.array_equal(out[ok], np.floor(ref[ok]).astype(np.uint8))


def test_affine_multichannel(oracle, rng):
    v = rng.integers(0, 256, (3, 6, 7, 8)).astype(np.uint8)
    M = rand_affine(rng, 0.05)
    out = oracle.affine_resample(v, M)
    for c in range(3):
        assert np.array_equal(out[c], oracle.affine_resample(v[c], M))


def test_tps_with_affine_landmarks_matches_affine_within_one_lsb(oracle, rng):
    """Whole-chain cross-check: TPS rounds (+0.5), affine truncates."""
    size = (40, 36, 24)
    v = rng.integers(0, 256, size[::-1]).astype(np.uint8)
    v = ndimage.gaussian_filter(v.astype(float), 1.5).astype(np.uint8)
    M = np.eye(4)
    M[:3, :3] += rng.uniform(-0.03, 0.03, (3, 3))
    M[:3, 3] = rng.uniform(-2, 2, 3)
    tar = (rng.uniform(0, 1, (30, 3)) * (np.array(size) - 1)).astype(np.float32)
    sub = apply(M, tar).astype(np.float32)
    warped, _, _ = oracle.warp_tps(v, tar, sub, g=4)
    aff = oracle.affine_resample(v, M)
    both = (warped > 0) & (aff > 0)
    diff = warped.astype(int)[both] - aff.astype(int)[both]
    assert diff.size > 0.8 * v.size
    assert diff.min() >= 0 and diff.max() <= 1
