"""Python view of the host-side surface (include/terawarp_host.h): Vaa3D
.v3draw / .marker I/O and the plugin's dofunc / bigimagewarp entry."""
import ctypes

import numpy as np

from . import _lib as L

_DT = {1: np.uint8, 2: np.uint16, 4: np.float32}


def _err():
    return L.load().tw_io_last_error().decode(errors="replace")


class HostError(RuntimeError):
    pass


def read_v3draw(path):
    """(c, z, y, x) array of a .v3draw stack."""
    lib = L.load()
    sz = (ctypes.c_int64 * 4)()
    dt, sw, hl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if lib.tw_v3draw_header(path.encode(), sz, ctypes.byref(dt), ctypes.byref(sw), ctypes.byref(hl)) != 0:
        raise HostError(_err())
    arr = np.empty((sz[3], sz[2], sz[1], sz[0]), _DT[dt.value])
    if lib.tw_v3draw_read(path.encode(), arr.ctypes.data, arr.nbytes) != 0:
        raise HostError(_err())
    return arr


def write_v3draw(path, arr):
    arr = np.ascontiguousarray(arr)
    if arr.ndim == 3:
        arr = arr[None]
    code = {np.dtype(np.uint8): 1, np.dtype(np.uint16): 2, np.dtype(np.float32): 4}[arr.dtype]
    sz = (ctypes.c_int64 * 4)(arr.shape[3], arr.shape[2], arr.shape[1], arr.shape[0])
    if L.load().tw_v3draw_write(path.encode(), arr.ctypes.data, sz, code) != 0:
        raise HostError(_err())


def read_markers(path):
    lib = L.load()
    n = lib.tw_markers_read(path.encode(), None, 0)
    if n < 0:
        raise HostError(_err())
    xyz = np.zeros((max(n, 1), 3), np.float32)
    lib.tw_markers_read(path.encode(), xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n)
    return xyz[:n]


def write_markers(path, xyz):
    xyz = np.ascontiguousarray(xyz, np.float32)
    if L.load().tw_markers_write(path.encode(), xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 xyz.shape[0]) != 0:
        raise HostError(_err())


def _cstrs(items):
    arr = (ctypes.c_char_p * max(len(items), 1))(*[str(s).encode() for s in items])
    return arr, len(items)


def parse_params(paras):
    p = L.TwWarpParams()
    arr, n = _cstrs(paras)
    rc = L.load().tw_plugin_parse_params(arr, n, ctypes.byref(p))
    if rc != 0:
        raise HostError(_err())
    return p


def dofunc(func_name, inputs, paras, outputs):
    """The plugin's dofunc: True on success (bool, like the reference)."""
    i, ni = _cstrs(inputs)
    p, np_ = _cstrs(paras)
    o, no = _cstrs(outputs)
    return bool(L.load().tw_plugin_dofunc(func_name.encode(), i, ni, p, np_, o, no))


def last_error():
    return _err()
