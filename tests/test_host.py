"""Host-side surface on the CPU: .v3draw / .marker I/O, -p parsing, plugin
lists and dofunc's argument errors (all return before any GPU call)."""
import struct

import numpy as np
import pytest

from terawarp_amd import _lib as L
from terawarp_amd import host


def test_v3draw_round_trip(tmp_path, rng):
    for dt in (np.uint8, np.uint16):
        a = rng.integers(0, np.iinfo(dt).max + 1, (2, 5, 6, 7)).astype(dt)
        p = str(tmp_path / f"a{dt.__name__}.v3draw")
        host.write_v3draw(p, a)
        raw = open(p, "rb").read()
        assert raw[:24] == b"raw_image_stack_by_hpeng" and raw[24:25] == b"L"
        assert struct.unpack("<h4i", raw[25:43]) == (np.dtype(dt).itemsize, 7, 6, 5, 2)
        assert len(raw) == 43 + a.nbytes
        assert np.array_equal(host.read_v3draw(p), a)


def test_v3draw_big_endian_is_swapped(tmp_path, rng):
    a = rng.integers(0, 65536, (1, 3, 4, 5)).astype(np.uint16)
    p = str(tmp_path / "be.v3draw")
    with open(p, "wb") as f:
        f.write(b"raw_image_stack_by_hpeng" + b"B" + struct.pack(">h4i", 2, 5, 4, 3, 1))
        f.write(a.astype(">u2").tobytes())
    assert np.array_equal(host.read_v3draw(p), a)


@pytest.mark.parametrize("payload,msg", [
    (b"not_a_stack_file_at_all_xxxxxxxxxxxxxxxxxxxxxxx", "Unrecognized"),
    (b"raw_image_stack_by_hpengL" + struct.pack("<h4i", 3, 1, 1, 1, 1) + b"\0" * 8, "data type"),
    (b"raw_image_stack_by_hpengL" + struct.pack("<h4i", 1, 4, 4, 4, 1) + b"\0" * 8, "shorter"),
])
def test_v3draw_errors(tmp_path, payload, msg):
    p = str(tmp_path / "bad.v3draw")
    open(p, "wb").write(payload)
    with pytest.raises(host.HostError, match=msg):
        host.read_v3draw(p)


def test_markers_round_trip_and_comments(tmp_path):
    p = str(tmp_path / "m.marker")
    open(p, "w").write("##x,y,z,radius,shape,name,comment, color_r,color_g,color_b\n"
                       "10.5,20,30,0,1,,,255,0,0\n\n# a comment\n1,2,3\n")
    xyz = host.read_markers(p)
    assert np.allclose(xyz, [[10.5, 20, 30], [1, 2, 3]])
    q = str(tmp_path / "n.marker")
    pts = np.array([[1.25, 2.5, 3.75], [100.125, 0, 7]], np.float32)
    host.write_markers(q, pts)
    assert np.array_equal(host.read_markers(q), pts)


def test_param_parsing():
    p = host.parse_params(["0", "1", "0", "1", "0", "0", "0", "0", "0", "0"])
    assert (p.kind, p.device_mode, p.resize, p.legacy) == (1, 1, 0, 0)
    p = host.parse_params(["1", "2", "1", "0", "10", "-20", "90", "64", "32", "16"])
    assert p.kind == 3 and p.device_mode == 2 and list(p.para) == [10, -20, 90] and list(p.newsize) == [64, 32, 16]
    p = host.parse_params(["0", "0", "0", "2", "2", "0.5", "1", "0", "0", "0"])
    assert p.kind == 2
    p = host.parse_params(["1", "0", "0", "0", "0", "0", "0"])  # reference code form: program_mode 1 = affine
    assert p.legacy == 1 and p.kind == 0 and p.device_mode == 0
    p = host.parse_params(["0", "1", "1", "0", "0", "0", "0"])
    assert p.kind == 1 and p.downsample == 1
    for bad in (["1", "2"], ["0", "1", "0", "5", "0", "0", "0", "0", "0", "0"],
                ["0", "1", "0", "2", "20", "1", "1", "0", "0", "0"], ["0", "9", "0", "1", "0", "0", "0", "0", "0", "0"],
                ["1", "0", "0", "0", "200", "0", "0", "0", "0", "0"], ["0", "1", "1", "1", "0", "0", "0", "0", "4", "4"],
                ["x", "1", "0", "1", "0", "0", "0", "0", "0", "0"]):
        with pytest.raises(host.HostError):
            host.parse_params(bad)


def test_plugin_lists():
    lib = L.load()
    assert [lib.tw_plugin_menu_item(i) for i in range(lib.tw_plugin_menu_count())] == [b"bigimagewarp", b"about"]
    assert [lib.tw_plugin_func_item(i) for i in range(lib.tw_plugin_func_count())] == [b"bigimagewarp", b"help"]
    assert abs(lib.tw_plugin_version() - 1.1) < 1e-6


def test_dofunc_argument_errors(tmp_path):
    ten = ["0", "1", "0", "1", "0", "0", "0", "0", "0", "0"]
    img = str(tmp_path / "s.v3draw")
    host.write_v3draw(img, np.zeros((4, 4, 4), np.uint8))
    m = str(tmp_path / "a.marker")
    host.write_markers(m, np.zeros((5, 3), np.float32))
    m2 = str(tmp_path / "b.marker")
    host.write_markers(m2, np.zeros((6, 3), np.float32))
    assert not host.dofunc("bigimagewarp", [img, m], ten, [str(tmp_path)])
    assert "no enough para" in host.last_error()
    assert not host.dofunc("bigimagewarp", [img, m, m], ["1", "2", "3"], [str(tmp_path)])
    assert not host.dofunc("bigimagewarp", [img, m, m + ".txt"], ten, [str(tmp_path)])
    assert "marker" in host.last_error()
    assert not host.dofunc("bigimagewarp", [img, m, m2], ten, [str(tmp_path)])
    assert "control points" in host.last_error()
    # A directory is a slice series (the tera branch): these files are not slices.
    assert not host.dofunc("bigimagewarp", [str(tmp_path), m, m], ten, [str(tmp_path)])
    assert "not a Vaa3D raw stack" in host.last_error()
    assert host.dofunc("help", [], [], [])


def test_slice_dir_reads_windows(tmp_path, rng):
    """A directory of single-slice stacks (sorted by name; mdata.bin and
    .iim.format skipped) reads back any window; big-endian slices are swapped."""
    vol = rng.integers(0, 65536, (7, 13, 17)).astype(np.uint16)
    d = tmp_path / "slices"
    host.write_slices(str(d), vol)
    (d / "mdata.bin").write_bytes(b"x")
    (d / ".iim.format").write_bytes(b"y")
    # slice 3 rewritten big-endian
    with open(d / "0000030.raw", "wb") as f:
        f.write(b"raw_image_stack_by_hpeng" + b"B" + struct.pack(">h4i", 2, 17, 13, 1, 1))
        f.write(vol[3].astype(">u2").tobytes())
    sd = host.SliceDir(str(d))
    assert sd.size == (17, 13, 7) and sd.dtype == np.uint16
    assert np.array_equal(sd.read(0, 0, 0, 17, 13, 7), vol)
    for (x0, y0, z0, nx, ny, nz) in [(3, 2, 1, 9, 5, 4), (16, 12, 6, 1, 1, 1), (0, 5, 2, 17, 3, 5)]:
        assert np.array_equal(sd.read(x0, y0, z0, nx, ny, nz), vol[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx])
    with pytest.raises(host.HostError):
        sd.read(10, 0, 0, 8, 1, 1)  # past the row
    sd.close()


def test_slice_dir_rejects_mixed_slices(tmp_path):
    d = tmp_path / "s"
    host.write_slices(str(d), np.zeros((3, 4, 5), np.uint8))
    host.write_v3draw(str(d / "0000030.raw"), np.zeros((1, 4, 6), np.uint8))
    with pytest.raises(host.HostError):
        host.SliceDir(str(d))


def test_downsample_factor():
    # m / 600 rounded half up, m the largest extent (q_warp_affine_tps.cpp:1477-1479)
    assert host.downsample_factor(299, 10, 10) == 0
    assert host.downsample_factor(300, 10, 10) == 1
    assert host.downsample_factor(10, 899, 10) == 1
    assert host.downsample_factor(10, 10, 900) == 2
    assert host.downsample_factor(8192, 8192, 2048) == 14
