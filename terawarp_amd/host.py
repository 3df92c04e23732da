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


# ---- slice directories (the tera driver's input and output) -----------------

def write_slices(directory, vol, name="{:07d}.raw"):
    """Write a (z, y, x) volume as one single-slice Vaa3D raw stack per z."""
    import os
    os.makedirs(directory, exist_ok=True)
    for z in range(vol.shape[0]):
        write_v3draw(os.path.join(directory, name.format(z * 10)), vol[z:z + 1])


def _terr():
    return L.load().tw_tera_last_error().decode(errors="replace")


class SliceDir:
    """An open slice directory (tw_slice_dir): size, windows, and a read
    callback for the streaming driver."""

    def __init__(self, directory):
        self.lib = L.load()
        self.h = ctypes.c_void_p()
        if self.lib.tw_slice_dir_open(str(directory).encode(), ctypes.byref(self.h)) != 0:
            raise HostError(_terr())
        nx, ny, nz, dt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        self.lib.tw_slice_dir_size(self.h, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz), ctypes.byref(dt))
        self.size = (nx.value, ny.value, nz.value)
        self.dtype = np.dtype(np.uint8 if dt.value == L.TW_U8 else np.uint16)

    def read(self, x0, y0, z0, nx, ny, nz):
        out = np.empty((nz, ny, nx), self.dtype)
        if self.lib.tw_slice_dir_read(self.h, x0, y0, z0, nx, ny, nz, out.ctypes.data) != 0:
            raise HostError(_terr())
        return out

    def close(self):
        if self.h:
            self.lib.tw_slice_dir_close(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_slices(directory):
    d = SliceDir(directory)
    try:
        nx, ny, nz = d.size
        return d.read(0, 0, 0, nx, ny, nz)
    finally:
        d.close()


def downsample_factor(nx, ny, nz):
    return int(L.load().tw_downsample_factor(nx, ny, nz))
