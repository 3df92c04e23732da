"""End to end through the plugin surface on the GPU: the CLI (dofunc)
reads .v3draw + .marker files, warps, and writes <out>/<name>_<mode>.v3draw,
checked against the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from terawarp_amd import host, synth, warp

from .parity import check_fast_form

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "terawarp_amd", "terawarp")


def run_cli(img, msub, mtar, out, params):
    r = subprocess.run([CLI, "-i", img, msub, mtar, "-o", out, "-p"] + [str(p) for p in params],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.fixture
def files(tmp_path):
    size = (96, 72, 40)
    vol = synth.blob_volume(size[::-1], np.uint16, n_blobs=40, amp=(100, 4000), seed=8)
    sub, tar = synth.landmarks(size, 32, 4.0, seed=8)
    img = str(tmp_path / "subject.v3draw")
    host.write_v3draw(img, vol)
    ms, mt = str(tmp_path / "sub.marker"), str(tmp_path / "tar.marker")
    host.write_markers(ms, sub)
    host.write_markers(mt, tar)
    return dict(size=size, vol=vol, sub=host.read_markers(ms), tar=host.read_markers(mt), img=img, ms=ms, mt=mt,
                out=str(tmp_path))


def test_cli_tps_reference_exact(files, oracle):
    f = files
    run_cli(f["img"], f["ms"], f["mt"], f["out"], [0, 0, 0, 1, 0, 0, 0, 0, 0, 0])
    got = host.read_v3draw(os.path.join(f["out"], "subject_stps.v3draw"))[0]
    ref, _, _ = oracle.warp_tps(f["vol"], f["tar"], f["sub"], g=4)
    diff = np.abs(got.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and np.mean(diff == 0) >= 0.9999


def test_cli_legacy_affine(files, oracle):
    f = files
    run_cli(f["img"], f["ms"], f["mt"], f["out"], [1, 1, 0, 0, 0, 0, 0])
    got = host.read_v3draw(os.path.join(f["out"], "subject_affine.v3draw"))[0]
    M = oracle.affine_fit(f["tar"], f["sub"])
    ref = oracle.affine_resample(f["vol"], M)
    diff = np.abs(got.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and np.mean(diff == 0) >= 0.9999


@pytest.mark.parametrize("params,suffix,mode,para", [
    ([1, 1, 0, 0, 90, 0, -90, 0, 0, 0], "_rotation", warp.ROTATION, (90, 0, -90)),
    ([1, 1, 0, 1, 3, -2, 5, 0, 0, 0], "_translation", warp.TRANSLATION, (3, -2, 5)),
    ([0, 2, 1, 2, 2, 2, 1, 192, 144, 40], "_scaling", warp.SCALING, (2, 2, 1)),
])
def test_cli_rigid_and_scaling_bit_exact(files, oracle, params, suffix, mode, para):
    f = files
    run_cli(f["img"], f["ms"], f["mt"], f["out"], params)
    got = host.read_v3draw(os.path.join(f["out"], "subject" + suffix + ".v3draw"))[0]
    M = warp.mode_matrix(mode, para, f["size"])
    out_size = tuple(params[7:10]) if params[2] else f["size"]
    ref = oracle.affine_resample(f["vol"], M, out_size_xyz=out_size)
    assert np.array_equal(got, ref)


# ---- tera branch: a directory of slices --------------------------------------
# The reference's tera drivers scale the marker files by 4 and warp 1000^3
# output blocks (imgwarp_smallmemory_CYF, q_imagewarp_affine_cyf).  The
# volume is 1100 voxels long in x, so the branch writes two blocks per
# z-layer, the second one clipped.

@pytest.fixture
def tera(tmp_path):
    size = (1100, 40, 24)
    vol = synth.blob_volume(size[::-1], np.uint8, n_blobs=60, amp=(20, 200), seed=9)
    sub, tar = synth.landmarks(size, 24, 4.0, seed=9)
    d = str(tmp_path / "stack")
    host.write_slices(d, vol)
    ms, mt = str(tmp_path / "sub.marker"), str(tmp_path / "tar.marker")
    host.write_markers(ms, sub / 4)  # marker files on the 4x-downsampled view
    host.write_markers(mt, tar / 4)
    out = str(tmp_path / "out")
    os.makedirs(out)
    return dict(size=size, vol=vol, dir=d, ms=ms, mt=mt, out=out, sub=host.read_markers(ms),
                tar=host.read_markers(mt))


def test_cli_tera_tps_reference_exact(tera, oracle):
    """device_mode 0: the slices equal the oracle's restatement of the tera
    TPS driver (x4 markers, 1000-voxel blocks, block-local float positions)
    bit for bit; block files on the 1000 lattice."""
    t = tera
    run_cli(t["dir"], t["ms"], t["mt"], t["out"], [0, 0, 0, 1, 0, 0, 0, 0, 0, 0])
    got = host.read_slices(os.path.join(t["out"], "stack_stps_slice"))
    ref = oracle.warp_tps_tera(t["vol"], t["tar"], t["sub"], g=4, block=1000)
    assert np.array_equal(got, ref), np.count_nonzero(got != ref)
    blocks = sorted(os.listdir(os.path.join(t["out"], "stack_stps_block")))
    assert blocks == ["000000_000000_000000.v3draw", "000000_000000_010000.v3draw"]
    b1 = host.read_v3draw(os.path.join(t["out"], "stack_stps_block", blocks[1]))[0]
    assert b1.shape == (24, 40, 100) and np.array_equal(b1, ref[:, :, 1000:])
    slices = sorted(os.listdir(os.path.join(t["out"], "stack_stps_slice")))
    assert slices[:2] == ["0000000.raw", "0000010.raw"] and len(slices) == 24


@pytest.mark.parametrize("mode", [1, 2])
def test_cli_tera_tps_fast_forms(tera, oracle, mode):
    """device_mode 1 (plain) and 2 (mixed): within one level of the tera
    oracle except where the oracle's sample lies within the displacement
    error bound of a face (bounds-test flips), and on at least 99.99 % of
    all voxels."""
    t = tera
    run_cli(t["dir"], t["ms"], t["mt"], t["out"], [0, mode, 0, 1, 0, 0, 0, 0, 0, 0])
    got = host.read_slices(os.path.join(t["out"], "stack_stps_slice"))
    ref = oracle.warp_tps_tera(t["vol"], t["tar"], t["sub"], g=4, block=1000)
    _, df, _ = oracle.warp_tps(t["vol"], 4 * t["tar"], 4 * t["sub"], g=4)
    check_fast_form(got, ref, oracle.dense_df(t["size"], df, 4), t["size"])


def test_cli_tera_affine_uses_4x_markers(tera, oracle):
    """Affine warp of a slice directory: the DLT fit of the x4 markers,
    resampled block by block with the affine truncation."""
    t = tera
    run_cli(t["dir"], t["ms"], t["mt"], t["out"], [1, 1, 0, 0, 0, 0, 0])
    got = host.read_slices(os.path.join(t["out"], "stack_affine_slice"))
    M = oracle.affine_fit(4 * t["tar"], 4 * t["sub"])
    ref = oracle.affine_resample(t["vol"], M)
    diff = np.abs(got.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and np.mean(diff == 0) >= 0.9999
    assert sorted(os.listdir(os.path.join(t["out"], "stack_affine_block"))) == [
        "000000_000000_000000.v3draw", "000000_000000_010000.v3draw"]


def test_cli_tera_rotation_bit_exact(tera, oracle):
    """A rigid rotation of a slice directory (no markers) equals the oracle's
    affine resample of the mode matrix bit for bit."""
    t = tera
    run_cli(t["dir"], t["ms"], t["mt"], t["out"], [1, 1, 0, 0, 2, -3, 5, 0, 0, 0])
    got = host.read_slices(os.path.join(t["out"], "stack_rotation_slice"))
    M = warp.mode_matrix(warp.ROTATION, (2, -3, 5), t["size"])
    assert np.array_equal(got, oracle.affine_resample(t["vol"], M))


def test_cli_tera_downsample(tmp_path):
    """downsample_mode 1: the warped slices downsampled by the factor of the
    largest extent (900 -> 2); each output voxel is the warped voxel at twice
    its coordinates; block files on the 1000 lattice (one block here) and
    slices under _stps_downsample."""
    size = (900, 40, 24)
    vol = synth.blob_volume(size[::-1], np.uint8, n_blobs=30, amp=(20, 200), seed=3)
    sub, tar = synth.landmarks(size, 16, 3.0, seed=3)
    d = str(tmp_path / "stack")
    host.write_slices(d, vol)
    ms, mt = str(tmp_path / "s.marker"), str(tmp_path / "t.marker")
    host.write_markers(ms, sub / 4)
    host.write_markers(mt, tar / 4)
    out = str(tmp_path / "out")
    os.makedirs(out)
    run_cli(d, ms, mt, out, [0, 1, 1, 0, 0, 0, 0])  # legacy form: TPS, plain, downsample
    assert host.downsample_factor(*size) == 2
    warped = host.read_slices(os.path.join(out, "stack_stps_slice"))
    ds = host.read_slices(os.path.join(out, "stack_stps_downsample_slice"))
    assert ds.shape == (12, 20, 450)
    assert np.array_equal(ds, warped[::2, ::2, ::2][:12, :20, :450])
    blocks = sorted(os.listdir(os.path.join(out, "stack_stps_downsample_block")))
    assert blocks == ["000000_000000_000000.v3draw"]
