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
