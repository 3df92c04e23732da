"""End to end through the plugin surface on the GPU: the CLI (dofunc)
reads .v3draw + .marker files, warps, and writes <out>/<name>_<mode>.v3draw,
checked against the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from terawarp_amd import host, synth, warp

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
# References come from the CLI's in-memory branch on the same volume written
# as a .v3draw stack (every GPU call of this file runs in a CLI process).

@pytest.fixture
def tera(tmp_path):
    size = (230, 150, 60)  # 3 x 2 x 1 output blocks of 100, uneven
    vol = synth.blob_volume(size[::-1], np.uint8, n_blobs=40, amp=(20, 200), seed=9)
    sub, tar = synth.landmarks(size, 24, 4.0, seed=9)
    d = str(tmp_path / "stack")
    host.write_slices(d, vol)
    img = str(tmp_path / "stack.v3draw")
    host.write_v3draw(img, vol)
    ms, mt = str(tmp_path / "sub.marker"), str(tmp_path / "tar.marker")
    host.write_markers(ms, sub)
    host.write_markers(mt, tar)
    out = str(tmp_path / "out")
    os.makedirs(out)
    return dict(size=size, vol=vol, dir=d, img=img, ms=ms, mt=mt, out=out)


def test_cli_tera_tps_equals_in_memory(tera):
    """Slice-directory TPS (streaming driver, 100^3 blocks) writes the block
    files and slices of the reference's layout; the slices equal the
    in-memory warp of the same volume."""
    t = tera
    params = [0, 1, 0, 1, 0, 0, 0, 0, 0, 0]
    run_cli(t["dir"], t["ms"], t["mt"], t["out"], params)
    run_cli(t["img"], t["ms"], t["mt"], t["out"], params)
    got = host.read_slices(os.path.join(t["out"], "stack_stps_slice"))
    ref = host.read_v3draw(os.path.join(t["out"], "stack_stps.v3draw"))[0]
    assert np.array_equal(got, ref), np.count_nonzero(got != ref)
    blocks = sorted(os.listdir(os.path.join(t["out"], "stack_stps_block")))
    assert blocks == sorted(f"{0:06d}_{y * 10:06d}_{x * 10:06d}.v3draw" for y in (0, 100) for x in (0, 100, 200))


@pytest.mark.parametrize("params,suffix", [([1, 1, 0, 0, 0, 0, 0], "_affine"),
                                           ([1, 1, 0, 0, 30, -20, 45, 0, 0, 0], "_rotation")])
def test_cli_tera_affine_equals_in_memory(tera, params, suffix):
    """Affine and rigid warps of a slice directory, block by block, equal the
    in-memory warp."""
    t = tera
    run_cli(t["dir"], t["ms"], t["mt"], t["out"], params)
    run_cli(t["img"], t["ms"], t["mt"], t["out"], params)
    got = host.read_slices(os.path.join(t["out"], "stack" + suffix + "_slice"))
    ref = host.read_v3draw(os.path.join(t["out"], "stack" + suffix + ".v3draw"))[0]
    assert np.array_equal(got, ref), np.count_nonzero(got != ref)


def test_cli_tera_downsample(tmp_path):
    """downsample_mode 1: the warped slices downsampled by the factor of the
    largest extent (900 -> 2); each output voxel is the warped voxel at twice
    its coordinates; 500^3 block files and slices under _downsample."""
    size = (900, 40, 24)
    vol = synth.blob_volume(size[::-1], np.uint8, n_blobs=30, amp=(20, 200), seed=3)
    sub, tar = synth.landmarks(size, 16, 3.0, seed=3)
    d = str(tmp_path / "stack")
    host.write_slices(d, vol)
    ms, mt = str(tmp_path / "s.marker"), str(tmp_path / "t.marker")
    host.write_markers(ms, sub)
    host.write_markers(mt, tar)
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
