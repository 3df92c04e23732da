"""GPU affine/rigid resample vs the FP64 oracle: bit-exact (same matrix,
same FP64 operation order)."""
import numpy as np
import pytest

from terawarp_amd import synth, warp

pytestmark = pytest.mark.gpu


def _vol(rng, shape, dtype):
    hi = np.iinfo(dtype).max
    return rng.integers(0, hi + 1, shape).astype(dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("kind", ["random", "translate", "rotate90", "rotate", "scale2", "scale3"])
def test_affine_bit_exact(warper, oracle, rng, dtype, kind):
    size = (97, 66, 41)
    v = synth.blob_volume(size[::-1], dtype, n_blobs=30, amp=(20, 200) if dtype == np.uint8 else (100, 4000),
                          seed=5)
    if kind == "random":
        M = synth.random_affine(size, seed=7)
    elif kind == "translate":
        M = warp.mode_matrix(warp.TRANSLATION, (3, -5, 2), size)
    elif kind == "rotate90":
        M = warp.mode_matrix(warp.ROTATION, (90, 0, -90), size)
    elif kind == "rotate":
        M = warp.mode_matrix(warp.ROTATION, (12.5, -31, 170), size)
    elif kind == "scale2":
        M = warp.mode_matrix(warp.SCALING, (2, 2, 2), size)
    else:
        M = warp.mode_matrix(warp.SCALING, (3, 0.5, 1), size)
    got = warper.resample_affine(v, M)
    ref = oracle.affine_resample(v, M)
    assert np.array_equal(got, ref), (kind, np.count_nonzero(got != ref))


def test_affine_fixture(warper):
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "affine_u16.npz"))
    assert np.array_equal(warper.resample_affine(f["vol"], f["M"]), f["out"])


def test_affine_multichannel_resize_and_nn(warper, oracle, rng):
    v = _vol(rng, (2, 20, 30, 40), np.uint8)
    M = synth.random_affine((40, 30, 20), seed=3)
    got = warper.resample_affine(v, M, out_size_xyz=(50, 25, 22))
    ref = oracle.affine_resample(v, M, out_size_xyz=(50, 25, 22))
    assert np.array_equal(got, ref)
    # Nearest neighbour on an integer translation is the same shifted copy.
    T = warp.mode_matrix(warp.TRANSLATION, (2, -3, 1), (40, 30, 20))
    nn = warper.resample_affine(v[0], T, interp=warp.NN)
    assert np.array_equal(nn, oracle.affine_resample(v[0], T))


def test_affine_device_buffers_and_windows(warper, oracle, rng):
    import torch
    size = (64, 48, 32)
    v = synth.blob_volume(size[::-1], np.uint16, n_blobs=20, amp=(100, 4000), seed=9)
    M = synth.random_affine(size, seed=9)
    ref = oracle.affine_resample(v, M)
    src = torch.from_numpy(v.view(np.int16)).cuda()
    dst = torch.zeros_like(src)
    warper.resample_affine(src, M, dst=dst)
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy().view(np.uint16), ref)
    # Output brick at an offset, source window = whole volume.
    sub = np.zeros((10, 12, 16), np.uint16)
    warper.resample_affine(v, M, dst=sub, dst_origin=(20, 30, 7))
    assert np.array_equal(sub, ref[7:17, 30:42, 20:36])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_affine_plain_gpu_config_size(warper, oracle, dtype):
    """Config-sized 1024x1024x256 volume, checked on sampled slabs."""
    import torch
    size = (1024, 1024, 256)
    v = synth.blob_volume(size[::-1], dtype, n_blobs=200, amp=(100, 4000) if dtype == np.uint16 else (20, 200),
                          seed=21)
    M = synth.random_affine(size, seed=21)
    src = torch.from_numpy(v.view(tdt_np(dtype))).cuda()
    dst = torch.empty_like(src)
    warper.resample_affine(src, M, dst=dst)
    torch.cuda.synchronize()
    out = dst.cpu().numpy().view(dtype)
    for z0 in (0, 101, 255):
        ref = oracle.affine_resample(v, M, zrange=(z0, z0 + 1))
        assert np.array_equal(out[z0], ref[z0]), z0


def tdt_np(dtype):
    return np.uint8 if dtype == np.uint8 else np.int16


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("interp", [warp.LINEAR, warp.NN])
def test_affine_identity_reproduces_every_voxel(warper, rng, dtype, interp):
    """Identity matrix: every sample sits exactly on a voxel, the last one on
    the channel's final voxel (x_b = x_s everywhere, pairs at the last column
    shifted): the output is the input, channel by channel."""
    v = _vol(rng, (2, 13, 7, 33), dtype)
    I = np.eye(4)
    assert np.array_equal(warper.resample_affine(v, I, interp=interp), v)
