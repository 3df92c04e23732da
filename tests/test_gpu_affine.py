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


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("case", ["rot_2ch", "rot_nn", "rot_scaled", "rot_window", "rot_edges", "rot_srcwin"])
def test_affine_large_rotation_tiles(warper, oracle, rng, dtype, case):
    """Maps whose x step crosses source rows every lane or two run the
    LDS-staged tile (source box staged per 16 x 16 x 8 output tile).  Tiles
    whose samples all leave the volume store zeros; tiles whose samples
    partly leave a whole-volume source stage the box clipped to the volume;
    a box over the LDS budget, or samples leaving a partial source window,
    gather.  Every case is bit-identical to the oracle."""
    size = (83, 71, 45)
    amp = (20, 200) if dtype == np.uint8 else (100, 4000)
    v = synth.blob_volume(size[::-1], dtype, n_blobs=40, amp=amp, seed=13)
    R = warp.mode_matrix(warp.ROTATION, (37.0, -71.0, 113.0), size)
    # Outputs start as 0xA5 bytes: a voxel left unwritten (an all-outside
    # tile's zeros included) differs from the oracle.
    def filled(a):
        return np.full_like(a, 0xA5A5 if a.dtype == np.uint16 else 0xA5)
    if case == "rot_2ch":
        v2 = np.stack([v, v[::-1].copy()])
        got = warper.resample_affine(v2, R, dst=filled(v2))
        assert np.array_equal(got, oracle.affine_resample(v2, R))
        return
    if case == "rot_nn":
        got = warper.resample_affine(v, R, interp=warp.NN, dst=filled(v))
        assert np.array_equal(got, _affine_nn(v, R))
        return
    if case == "rot_scaled":
        # Rotation then 5x minification: a tile's source box is far over the
        # LDS budget, so those tiles gather.
        M = R @ warp.mode_matrix(warp.SCALING, (0.2, 0.2, 0.2), size)
        got = warper.resample_affine(v, M, dst=filled(v))
        assert np.array_equal(got, oracle.affine_resample(v, M))
        return
    if case == "rot_window":
        # Output brick at an offset, output larger than the source.
        ref = oracle.affine_resample(v, R, out_size_xyz=(96, 80, 50))
        sub = filled(np.zeros((20, 33, 40), dtype))
        warper.resample_affine(v, R, dst=sub, dst_origin=(31, 17, 9))
        assert np.array_equal(sub, ref[9:29, 17:50, 31:71])
        return
    if case == "rot_srcwin":
        # A partial source window of the volume: a sample whose taps leave
        # the window is 0 (q_warp_affine_tps.cpp's block form reads only the
        # block it loaded), the others equal the whole-volume result.
        full = oracle.affine_resample(v, R)
        w0, wn = (10, 5, 3), (60, 55, 37)
        win = v[w0[2]:w0[2] + wn[2], w0[1]:w0[1] + wn[1], w0[0]:w0[0] + wn[0]].copy()
        got = warper.resample_affine(win, R, dst=filled(v), src_origin=w0, full=size)
        z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in v.shape), indexing="ij")
        keep = np.ones(v.shape, bool)
        for c, coord in enumerate((x, y, z)):
            p = R[c, 0] * x + R[c, 1] * y + R[c, 2] * z + R[c, 3] * 1.0
            keep &= (np.floor(p) >= w0[c]) & (np.ceil(p) <= w0[c] + wn[c] - 1)
        exp = np.where(keep, full, 0).astype(v.dtype)
        assert keep.any() and (~keep & (full != 0)).any()
        assert np.array_equal(got, exp)
        return
    # rot_edges: a 90-degree turn about z with a shift, so that tiles' boxes
    # reach the volume's last rows and planes exactly (the channel's last
    # 16-byte piece is read byte by byte).
    M = warp.mode_matrix(warp.ROTATION, (0.0, 0.0, 90.0), size) @ warp.mode_matrix(warp.TRANSLATION, (1, 2, 0), size)
    got = warper.resample_affine(v, M, dst=filled(v))
    assert np.array_equal(got, oracle.affine_resample(v, M))


def _affine_nn(v, M):
    """Nearest-neighbour affine resample with the linear path's position and
    bounds (q_warp_affine_tps.cpp:1383-1400: p = M [x y z 1] in FP64, summed
    left to right, strict bounds against the extent) and the TPS resampler's
    nearest-neighbour rule ((long)(p + 0.5), clamped to the extent;
    q_imgwarp_tps_quicksmallmemory.cpp:119-129)."""
    nz, ny, nx = v.shape
    gz, gy, gx = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                             np.arange(nx, dtype=np.float64), indexing="ij")
    p = [M[c, 0] * gx + M[c, 1] * gy + M[c, 2] * gz + M[c, 3] * 1.0 for c in range(3)]
    ext = (nx, ny, nz)
    inside = np.ones(v.shape, bool)
    q = []
    for c in range(3):
        inside &= (p[c] >= 0) & (p[c] <= ext[c] - 1)
        qc = np.floor(np.where(inside, p[c], 0.0) + 0.5).astype(np.int64)
        q.append(np.minimum(qc, ext[c] - 1))
    out = v[q[2], q[1], q[0]]
    return np.where(inside, out, 0).astype(v.dtype)
