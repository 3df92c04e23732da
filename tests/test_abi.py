"""The C-ABI library loads, exports every entry point include/terawarp.h
declares, and its host-side functions agree with the oracle (no GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from terawarp_amd import _lib as L
from terawarp_amd import warp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("terawarp.h", "terawarp_host.h")]


def declared_symbols(path=None):
    names = set()
    for h in ([path] if path else HEADERS):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(tw_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_and_binding_agree():
    assert declared_symbols(HEADERS[0]) == sorted(L.EXPORTS)
    assert declared_symbols(HEADERS[1]) == sorted(L.HOST_EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts():
    # Field offsets of the C structs (x86-64 / gfx950 ABI, 8-byte aligned).
    assert ctypes.sizeof(L.TwVolume) == 8 + 10 * 8 + 2 * 4
    assert L.TwGrid.qtab.offset == 8 + 3 * 8 + 2 * 4 + 3 * 8 + 2 * 4
    assert ctypes.sizeof(L.TwTimings) == 7 * 8 + 2 * 4 + 8
    assert ctypes.sizeof(L.TwStreamStats) == 8 + 9 * 8 and L.TwStreamStats.issue_s.offset == 8 + 8 * 8


def test_version_and_grid_dim():
    lib = L.load()
    assert b"gfx950" in lib.tw_version()
    assert lib.tw_grid_dim(256, 4) == 67 and lib.tw_grid_dim(1024, 8) == 131
    assert lib.tw_grid_dim(2048, 8) == 259 and lib.tw_grid_dim(10, 3) == 7
    assert lib.tw_grid_dim(0, 4) == 0


def test_null_context_is_rejected():
    lib = L.load()
    assert lib.tw_sync(None) == L.TW_E_ARG
    assert lib.tw_resample_affine(None, None, None, None, 0) == L.TW_E_ARG
    assert lib.tw_tps_fit(None, None, None, 0, 0.2, 0, None, None) == L.TW_E_ARG


def test_affine_fit_matches_oracle(oracle, rng):
    A = rng.uniform(0, 500, (40, 3)).astype(np.float32)
    M = np.eye(4)
    M[:3, :3] += rng.uniform(-0.2, 0.2, (3, 3))
    M[:3, 3] = rng.uniform(-20, 20, 3)
    B = (A.astype(np.float64) @ M[:3, :3].T + M[:3, 3] + rng.normal(0, 0.5, (40, 3))).astype(np.float32)
    Mp = warp.affine_fit(A, B)
    Mo = oracle.affine_fit(A, B)
    assert np.allclose(Mp, Mo, rtol=1e-10, atol=1e-10)


def test_affine_fit_rejects_bad_input():
    with pytest.raises(warp.TeraWarpError):
        warp.affine_fit(np.zeros((3, 3), np.float32), np.zeros((3, 3), np.float32))


def test_mode_matrices():
    size = (11, 9, 7)
    c = (np.array(size) - 1) / 2
    M = warp.mode_matrix(warp.TRANSLATION, (3, -2, 1), size)
    assert np.array_equal(M[:3, 3], [-3, 2, -1]) and np.array_equal(M[:3, :3], np.eye(3))
    M = warp.mode_matrix(warp.SCALING, (2, 4, 0.5), size)
    assert np.array_equal(np.diag(M)[:3], [0.5, 0.25, 2.0])
    # 90 degrees about z: exact entries, the centre maps to itself.
    M = warp.mode_matrix(warp.ROTATION, (0, 0, 90), size)
    assert set(np.unique(M[:3, :3])) <= {-1.0, 0.0, 1.0}
    assert np.allclose(M[:3, :3] @ c + M[:3, 3], c)
    # Rotation matrices are orthonormal and compose x, then y, then z.
    M = warp.mode_matrix(warp.ROTATION, (10, -25, 40), size)
    R = M[:3, :3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15)

    def rot(axis, deg):
        t = np.deg2rad(deg)
        cs, sn = np.cos(t), np.sin(t)
        return {0: np.array([[1, 0, 0], [0, cs, -sn], [0, sn, cs]]),
                1: np.array([[cs, 0, sn], [0, 1, 0], [-sn, 0, cs]]),
                2: np.array([[cs, -sn, 0], [sn, cs, 0], [0, 0, 1]])}[axis]
    Rf = rot(2, 40) @ rot(1, -25) @ rot(0, 10)
    assert np.allclose(R, Rf.T, atol=1e-14)
    with pytest.raises(warp.TeraWarpError):
        warp.mode_matrix(7, (0, 0, 0), size)
    with pytest.raises(warp.TeraWarpError):
        warp.mode_matrix(warp.SCALING, (0, 1, 1), size)


def test_rotation_90_matches_rot90(oracle, rng):
    """The mode matrix + the oracle resample reproduce numpy's rot90 (pins the
    rotation direction and centre convention on the CPU)."""
    v = rng.integers(0, 256, (5, 9, 9)).astype(np.uint8)
    M = warp.mode_matrix(warp.ROTATION, (0, 0, 90), (9, 9, 5))
    out = oracle.affine_resample(v, M)
    # output(q) = subject(R^T (q - c) + c) with R = +90 deg about z:
    # out[z, y, x] = v[z, cx + (x - cx) * 0 + ..]; equals rot90 in the (y, x) plane.
    assert np.array_equal(out, np.rot90(v, k=1, axes=(2, 1)))
