"""ctypes view of the FP64 CPU oracle (oracle/tw_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference the GPU results are compared
with.  The terawarp_amd product never imports this module and has no CPU
fallback.

Array conventions (numpy):
  points          (K, 3) float32, xyz
  volumes         (c, z, y, x) uint8 or uint16 (a 3-D (z, y, x) array is one channel)
  displacement    (3, gz, gy, gx) float32 planes sx, sy, sz
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libtw_oracle.so")
_lib = None

_i64 = ctypes.c_int64
_P = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (g++ -O2 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.or_stps_fit.argtypes = [_P, _P, ctypes.c_int, ctypes.c_double, _P, _P, ctypes.c_int]
        L.or_stps_fit_bordered.argtypes = L.or_stps_fit.argtypes
        L.or_grid_eval.argtypes = [_P, ctypes.c_int, _P, _P, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, _i64, _i64, _i64, _P, ctypes.c_int]
        L.or_bspline_basis.argtypes = [ctypes.c_int, _P, _P]
        L.or_df_basis.argtypes = [ctypes.c_int, ctypes.c_int, _P, _P]
        L.or_tps_resample.argtypes = [_P, _P, ctypes.c_int, _P, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int,
                                      _P, _i64, _i64, ctypes.c_int]
        L.or_tps_resample_window.argtypes = [_P, _P, _P, _P, ctypes.c_int, _P, _i64, _i64, _i64, _P, ctypes.c_int,
                                             ctypes.c_int, _P, _P, _P, ctypes.c_int]
        L.or_tps_warp_tera.argtypes = [_P, _P, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, _i64,
                                       ctypes.c_double, ctypes.c_int, _P, ctypes.c_int]
        L.or_dense_df.argtypes = [_P, _P, _i64, _i64, _i64, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]
        L.or_affine_fit.argtypes = [_P, _P, ctypes.c_int, _P]
        L.or_affine_resample.argtypes = [_P, _P, ctypes.c_int, _P, _P, _P, _i64, _i64, ctypes.c_int]
        for f in ("or_stps_fit", "or_stps_fit_bordered", "or_grid_eval", "or_bspline_basis", "or_df_basis",
                  "or_tps_resample",
                  "or_tps_resample_window", "or_tps_warp_tera", "or_dense_df", "or_affine_fit", "or_affine_resample", "or_max_threads"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def _pts(p):
    p = np.ascontiguousarray(p, dtype=np.float32)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError("points must be (K, 3)")
    return p


def _vol4(v):
    v = np.ascontiguousarray(v)
    if v.ndim == 3:
        v = v[None]
    if v.dtype == np.uint8:
        dt = 1
    elif v.dtype == np.uint16:
        dt = 2
    else:
        raise ValueError("volume dtype must be uint8 or uint16")
    return v, dt


def max_threads():
    return lib().or_max_threads()


def grid_shape(size_xyz, g):
    """Grid node counts per axis: ceil(sz/g) + 3 (q_imgwarp_tps_quicksmallmemory.cpp:592)."""
    gs = g if hasattr(g, "__len__") else (g, g, g)
    return tuple(int(math.ceil(s / f)) + 3 for s, f in zip(size_xyz, gs))


def stps_fit(centres, targets, lam=0.2, nthreads=0, bordered=False):
    centres, targets = _pts(centres), _pts(targets)
    K = centres.shape[0]
    c = np.zeros((K, 4), np.float64)
    d = np.zeros((4, 4), np.float64)
    f = lib().or_stps_fit_bordered if bordered else lib().or_stps_fit
    _check(f(_ptr(centres), _ptr(targets), K, lam, _ptr(c), _ptr(d), nthreads), "stps_fit")
    return c, d


def grid_eval(centres, c, d, gshape_xyz, g, origin=(0, 0, 0), nthreads=0):
    centres = _pts(centres)
    c = np.ascontiguousarray(c, np.float64)
    d = np.ascontiguousarray(d, np.float64)
    gx, gy, gz = gshape_xyz
    fx, fy, fz = g if hasattr(g, "__len__") else (g, g, g)
    df = np.zeros((3, gz, gy, gx), np.float32)
    _check(lib().or_grid_eval(_ptr(centres), centres.shape[0], _ptr(c), _ptr(d), gx, gy, gz, fx, fy, fz,
                              origin[0], origin[1], origin[2], _ptr(df), nthreads), "grid_eval")
    return df


# OR into a resample's interp: the linear DF upsample (df_basis kind 1).
DF_LINEAR = 0x100


def bspline_basis(n, full=True):
    TB = np.zeros((n, 4), np.float64)
    basis = np.zeros((n ** 3, 64), np.float64) if full else None
    _check(lib().or_bspline_basis(n, _ptr(TB), _ptr(basis) if full else None), "bspline_basis")
    return TB, basis


def df_basis(n, kind, full=True):
    """Upsampling weights of kind 0 (cubic B-spline) or 1 (linear)."""
    TB = np.zeros((n, 4), np.float64)
    basis = np.zeros((n ** 3, 64), np.float64) if full else None
    _check(lib().or_df_basis(n, kind, _ptr(TB), _ptr(basis) if full else None), "df_basis")
    return TB, basis


def tps_resample(src, df, n, interp=0, zblocks=(0, -1), dst=None, nthreads=0):
    """Warp src by the grid DF (block size n); returns the output volume."""
    v, dt = _vol4(src)
    df = np.ascontiguousarray(df, np.float32)
    sz = np.array([v.shape[3], v.shape[2], v.shape[1], v.shape[0]], np.int64)
    out = np.zeros_like(v) if dst is None else dst
    _check(lib().or_tps_resample(_ptr(v), _ptr(sz), dt, _ptr(df), df.shape[3], df.shape[2], df.shape[1], n, interp,
                                 _ptr(out), zblocks[0], zblocks[1], nthreads), "tps_resample")
    return out if src.ndim == 4 else out[0]


def tps_resample_window(src_win, src_origin, full_xyz, df, n, grid_origin, out_size_xyz, out_origin, interp=0,
                        nthreads=0):
    """Warp the output window [out_origin, out_origin + out_size) of a volume of
    extent full_xyz from the source window src_win (z, y, x) at src_origin,
    with a grid whose node (i, j, k) sits at (i-1)n + grid_origin."""
    v = np.ascontiguousarray(src_win)
    dt = 1 if v.dtype == np.uint8 else 2
    df = np.ascontiguousarray(df, np.float32)
    ssz, s0, fsz, g0, dsz, d0 = (np.array(list(a), np.int64) for a in (
        (v.shape[2], v.shape[1], v.shape[0]), src_origin, full_xyz, grid_origin, out_size_xyz, out_origin))
    out = np.zeros((out_size_xyz[2], out_size_xyz[1], out_size_xyz[0]), v.dtype)
    _check(lib().or_tps_resample_window(_ptr(v), _ptr(ssz), _ptr(s0), _ptr(fsz), dt, _ptr(df), df.shape[3],
                                        df.shape[2], df.shape[1], _ptr(g0), n, interp, _ptr(out), _ptr(dsz),
                                        _ptr(d0), nthreads), "tps_resample_window")
    return out


def warp_tps_tera(src, tar_markers, sub_markers, g=4, block=1000, lam=0.2, interp=0, nthreads=0):
    """The tera-block TPS driver of the reference's CPU path on a volume in
    memory: markers scaled by 4, output blocks of `block` voxels, each with
    its own grid and the block-local float order of the sample position."""
    v = np.ascontiguousarray(src)
    assert v.ndim == 3, "one channel"
    dt = 1 if v.dtype == np.uint8 else 2
    tar, sub = _pts(tar_markers), _pts(sub_markers)
    sz = np.array([v.shape[2], v.shape[1], v.shape[0]], np.int64)
    out = np.zeros_like(v)
    _check(lib().or_tps_warp_tera(_ptr(v), _ptr(sz), dt, _ptr(tar), _ptr(sub), tar.shape[0], g, block, lam, interp,
                                  _ptr(out), nthreads), "tps_warp_tera")
    return out


def dense_df(size_xyz, df, n, nthreads=0, kind=0):
    df = np.ascontiguousarray(df, np.float32)
    sz = np.array(list(size_xyz) + [1], np.int64)
    out = np.zeros((3, size_xyz[2], size_xyz[1], size_xyz[0]), np.float32)
    _check(lib().or_dense_df(_ptr(sz), _ptr(df), df.shape[3], df.shape[2], df.shape[1], n, _ptr(out), nthreads,
                             kind), "dense_df")
    return out


def affine_fit(pts_a, pts_b):
    """4x4 matrix M with pts_b ~= M pts_a (target -> subject at the plugin call site)."""
    a, b = _pts(pts_a), _pts(pts_b)
    M = np.zeros((4, 4), np.float64)
    _check(lib().or_affine_fit(_ptr(a), _ptr(b), a.shape[0], _ptr(M)), "affine_fit")
    return M


def affine_resample(src, M, out_size_xyz=None, zrange=(0, -1), nthreads=0):
    v, dt = _vol4(src)
    M = np.ascontiguousarray(M, np.float64).reshape(4, 4) if np.size(M) == 16 else \
        np.vstack([np.asarray(M, np.float64).reshape(3, 4), [0, 0, 0, 1]])
    ssz = np.array([v.shape[3], v.shape[2], v.shape[1], v.shape[0]], np.int64)
    if out_size_xyz is None:
        out_size_xyz = tuple(ssz[:3])
    dsz = np.array(list(out_size_xyz) + [v.shape[0]], np.int64)
    out = np.zeros((v.shape[0], dsz[2], dsz[1], dsz[0]), v.dtype)
    _check(lib().or_affine_resample(_ptr(v), _ptr(ssz), dt, _ptr(M), _ptr(out), _ptr(dsz), zrange[0], zrange[1],
                                    nthreads), "affine_resample")
    return out if src.ndim == 4 else out[0]


def warp_tps(src, tar_markers, sub_markers, g=4, lam=0.2, interp=0, nthreads=0):
    """Whole in-memory TPS chain of the reference CPU path: centres are the
    target-space markers, mapped onto the subject markers."""
    v, _ = _vol4(src)
    c, d = stps_fit(tar_markers, sub_markers, lam, nthreads)
    size = (v.shape[3], v.shape[2], v.shape[1])
    gsh = grid_shape(size, g)
    df = grid_eval(tar_markers, c, d, gsh, g, nthreads=nthreads)
    return tps_resample(src, df, g, interp, nthreads=nthreads), df, (c, d)
