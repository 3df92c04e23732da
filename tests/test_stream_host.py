"""Host side of the streaming driver (no GPU): the procedural volume is a
function of the global coordinate (every window equals the same part of the
whole volume), and tw_checksum64 is the stated position-weighted sum."""
import numpy as np
import pytest

from terawarp_amd import Synth, checksum64


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_synth_windows_match_the_whole_volume(dtype):
    s = Synth((70, 45, 38), dtype, seed=9, cell=16, blobs_per_cell=3, sigma=(1.5, 4.0))
    whole = s.fill(0, 0, 0, 70, 45, 38)
    assert whole.std() > 1.0  # blobs and noise, not a constant
    for (x0, y0, z0, nx, ny, nz) in [(0, 0, 0, 1, 1, 1), (13, 7, 5, 40, 30, 20), (69, 44, 37, 1, 1, 1),
                                     (30, 0, 17, 40, 45, 21)]:
        assert np.array_equal(s.fill(x0, y0, z0, nx, ny, nz), whole[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx])
    other = Synth((70, 45, 38), dtype, seed=10, cell=16, blobs_per_cell=3, sigma=(1.5, 4.0))
    assert not np.array_equal(other.fill(0, 0, 0, 70, 45, 38), whole)


def test_synth_rejects_boxes_outside():
    from terawarp_amd import TeraWarpError
    s = Synth((10, 10, 10), np.uint8)
    with pytest.raises(TeraWarpError):
        s.fill(5, 0, 0, 6, 1, 1)


def _checksum_np(a):
    b = np.ascontiguousarray(a).view(np.uint8).ravel()
    n = b.size
    pad = np.zeros((-n) % 8, np.uint8)
    w = np.concatenate([b, pad]).view("<u8")
    i = np.arange(w.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(w * (2 * i + 1), dtype=np.uint64)) ^ n


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 1000, 4096 + 3])
def test_checksum64_is_the_weighted_word_sum(n):
    a = np.random.default_rng(n).integers(0, 256, n).astype(np.uint8)
    assert checksum64(a) == _checksum_np(a)
    if n >= 16:  # position counts: swapping two words changes it
        b = a.copy()
        b[:8], b[8:16] = a[8:16], a[:8]
        if not np.array_equal(a[:8], a[8:16]):
            assert checksum64(b) != checksum64(a)


def test_tera_block_files_and_slices(tmp_path):
    """Block files named as the reference's tera driver names them, written
    through the stream's write callback, stitched back into per-z slices."""
    import ctypes
    from terawarp_amd import _lib as L
    from terawarp_amd.warp import stitch_slices, tera_paths
    b, s = tera_paths("/out/vol_stps", 512, 1024, 0, z=37)
    assert b == "/out/vol_stps_block/000000_010240_005120.v3draw"
    assert s == "/out/vol_stps_slice/0000370.raw"
    lib = L.load()
    size, brick = (45, 30, 21), 16                     # uneven bricks on every axis
    vol = np.random.default_rng(3).integers(0, 65536, size[::-1]).astype(np.uint16)
    base = str(tmp_path / "w")
    files = L.TwBlockFiles(base.encode(), L.TW_U16, 0)
    nbx, nby, nbz = (-(-n // brick) for n in size)
    for bid in range(nbx * nby * nbz):
        bx, by, bz = bid % nbx, (bid // nbx) % nby, bid // (nbx * nby)
        box = vol[bz * brick:(bz + 1) * brick, by * brick:(by + 1) * brick, bx * brick:(bx + 1) * brick]
        box = np.ascontiguousarray(box)
        rc = lib.tw_block_file_write(ctypes.byref(files), bid, bx * brick, by * brick, bz * brick, box.shape[2],
                                     box.shape[1], box.shape[0], box.ctypes.data)
        assert rc == 0
    stitch_slices(base, size, brick, np.uint16)
    for z in range(size[2]):
        path = tera_paths(base, z=z)[1]
        sz = (ctypes.c_int64 * 4)()
        dt, sw, hl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        assert lib.tw_v3draw_header(path.encode(), sz, ctypes.byref(dt), ctypes.byref(sw), ctypes.byref(hl)) == 0
        assert list(sz) == [size[0], size[1], 1, 1] and dt.value == 2
        out = np.zeros((size[1], size[0]), np.uint16)
        assert lib.tw_v3draw_read(path.encode(), out.ctypes.data, out.nbytes) == 0
        assert np.array_equal(out, vol[z]), z
