"""tw_warp_tps_noalloc through the C ABI against tw_warp_tps.

The allocation-free entry point runs the same kernels with its workspaces in
the library's static scratch, so its output must equal tw_warp_tps's bit for
bit.  Two calls in a row, the second on a smaller volume (the scratch keeps
the first call's contents), for each precision form; oversized requests are
rejected before any work.
"""
import ctypes

import numpy as np
import pytest

from terawarp_amd import _lib as L
from terawarp_amd import synth
from terawarp_amd.warp import _vol

pytestmark = pytest.mark.gpu


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available()
    lib = L.load()
    c = ctypes.c_void_p()
    assert lib.tw_create(0, ctypes.byref(c)) == L.TW_OK
    yield lib, c
    lib.tw_destroy(c)


def _case(size, K, dmax, seed):
    import torch
    vol = synth.blob_volume(size[::-1], np.uint16, n_blobs=30, amp=(100, 4000), seed=seed)
    sub, tar = synth.landmarks(size, K, dmax, seed=seed)
    src = torch.from_numpy(vol.view(np.int16)).cuda()
    return src, np.ascontiguousarray(tar, np.float32), np.ascontiguousarray(sub, np.float32)


def _warp(lib, c, fn, src, tar, sub, f, precision):
    import torch
    dst = torch.full_like(src, 7)
    vs, vd = _vol(src), _vol(dst)
    rc = fn(c, ctypes.byref(vs), ctypes.byref(vd), _fp(tar), _fp(sub), tar.shape[0], f, 0.2, L.TW_INTERP_LINEAR,
            precision)
    assert rc == L.TW_OK, lib.tw_last_error(c)
    torch.cuda.synchronize()
    return dst.cpu().numpy()


@pytest.mark.parametrize("precision", [L.TW_EXACT, L.TW_PLAIN, L.TW_MIXED])
def test_noalloc_equals_warp_tps_twice(ctx, precision):
    lib, c = ctx
    # First a larger call (K above the one-work-group fit: the blocked
    # factor, the pipelined solve and its flags), then a smaller one.
    for size, K, f, seed in (((200, 160, 72), 300, 8, 1), ((96, 64, 40), 40, 4, 2)):
        src, tar, sub = _case(size, K, 6.0, seed)
        got = _warp(lib, c, lib.tw_warp_tps_noalloc, src, tar, sub, f, precision)
        ref = _warp(lib, c, lib.tw_warp_tps, src, tar, sub, f, precision)
        assert np.array_equal(got, ref), (size, precision, np.count_nonzero(got != ref))


def test_noalloc_rejects_what_it_cannot_serve(ctx):
    import torch
    lib, c = ctx
    src, tar, sub = _case((64, 48, 24), 20, 3.0, 3)
    dst = torch.empty_like(src)
    vs, vd = _vol(src), _vol(dst)
    big_t = np.ascontiguousarray(np.tile(tar, (60, 1))[:1100], np.float32)
    big_s = np.ascontiguousarray(np.tile(sub, (60, 1))[:1100], np.float32)
    rc = lib.tw_warp_tps_noalloc(c, ctypes.byref(vs), ctypes.byref(vd), _fp(big_t), _fp(big_s), 1100, 4, 0.2,
                                 L.TW_INTERP_LINEAR, L.TW_MIXED)
    assert rc == L.TW_E_UNSUPPORTED and b"1024" in lib.tw_last_error(c)
    host = np.zeros((24, 48, 64), np.uint16)
    vh = _vol(host)
    rc = lib.tw_warp_tps_noalloc(c, ctypes.byref(vs), ctypes.byref(vh), _fp(tar), _fp(sub), 20, 4, 0.2,
                                 L.TW_INTERP_LINEAR, L.TW_MIXED)
    assert rc == L.TW_E_ARG and b"device buffers" in lib.tw_last_error(c)
    # A grid beyond the scratch: 4096 x 4096 x 64 at f = 2 (about 140 M nodes).
    vbig = _vol(src)
    vbig.nx, vbig.ny, vbig.nz = 4096, 4096, 64
    rc = lib.tw_warp_tps_noalloc(c, ctypes.byref(vs), ctypes.byref(vbig), _fp(tar), _fp(sub), 20, 2, 0.2,
                                 L.TW_INTERP_LINEAR, L.TW_MIXED)
    assert rc == L.TW_E_UNSUPPORTED and b"MiB of scratch" in lib.tw_last_error(c)
