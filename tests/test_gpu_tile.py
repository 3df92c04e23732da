"""The fused TPS resample's two tiles give identical outputs.

The LDS tile streams each 16x16 output column's source box through a ring
of planes in LDS; the gather tile reads every tap from global memory.  Both
compute the same per-voxel displacement, position, bounds tests and blend, so
the bar is
bit-identity, on: whole volumes and source windows, uint8 and uint16, PLAIN
and MIXED, FP32 and FP16 grids, grid spacings 1..64, tiles whose box exceeds
the LDS ring (the gather fallback inside the LDS kernel), samples past the
volume's faces and the window's edges, volumes smaller than one tile, two
channels, and the bricks launch.  One case is also held against the FP64
oracle (tests/parity.py bars).
"""
import numpy as np
import pytest

from terawarp_amd import synth, warp

from .parity import check_fast_form

pytestmark = pytest.mark.gpu


def smooth_df(gshape_xyz, amp, seed, shift=(0.0, 0.0, 0.0)):
    """Random grid nodes uniform in [-amp, amp] per component, plus shift."""
    gx, gy, gz = gshape_xyz
    r = np.random.default_rng(seed)
    df = r.uniform(-amp, amp, (3, gz, gy, gx)).astype(np.float32)
    for c in range(3):
        df[c] += np.float32(shift[c])
    return df


def volume(size_xyz, dtype, seed, nc=1):
    amp = (20, 200) if dtype == np.uint8 else (100, 4000)
    vols = [synth.blob_volume(size_xyz[::-1], dtype, n_blobs=40, amp=amp, seed=seed + i) for i in range(nc)]
    return vols[0] if nc == 1 else np.ascontiguousarray(np.stack(vols))


def both_tiles(warper, fn):
    """fn() under the gather tile and under the LDS column tile."""
    try:
        warper.set_tile("gather")
        a = fn()
        warper.set_tile("lds")
        b = fn()
    finally:
        warper.set_tile("auto")
    return a, b


def assert_same(a, b, what):
    assert a.shape == b.shape
    n = int(np.count_nonzero(a != b))
    assert n == 0, f"{what}: {n} of {a.size} voxels differ"


CASES = [
    # size (x, y, z), dtype, f, amplitude, shift, precision
    ((200, 150, 70), np.uint16, 8, 1.5, (0, 0, 0), warp.PLAIN),
    ((200, 150, 70), np.uint16, 8, 1.5, (0, 0, 0), warp.MIXED),
    ((131, 77, 41), np.uint8, 4, 1.0, (0, 0, 0), warp.MIXED),
    ((131, 77, 41), np.uint8, 5, 1.0, (2.5, -1.25, 0.5), warp.PLAIN),
    ((96, 80, 48), np.uint16, 16, 12.0, (0, 0, 0), warp.MIXED),     # most boxes past the pitches
    ((96, 80, 48), np.uint16, 8, 0.5, (30.5, -20.25, 7.75), warp.MIXED),  # samples past the faces
    ((96, 80, 48), np.uint8, 8, 0.5, (-6.5, 9.25, -3.75), warp.PLAIN),
    ((33, 17, 9), np.uint16, 3, 1.0, (0, 0, 0), warp.MIXED),
    ((5, 3, 2), np.uint16, 1, 0.75, (0, 0, 0), warp.MIXED),       # below one tile
    ((70, 40, 30), np.uint16, 1, 0.4, (0, 0, 0), warp.PLAIN),
    ((70, 40, 30), np.uint16, 2, 0.8, (0, 0, 0), warp.MIXED),
    ((130, 70, 40), np.uint16, 32, 2.0, (0, 0, 0), warp.MIXED),
    ((130, 70, 40), np.uint8, 64, 2.0, (1.0, 1.0, 1.0), warp.PLAIN),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_tiles_identical_whole_volume(warper, i):
    size, dtype, f, amp, shift, prec = CASES[i]
    vol = volume(size, dtype, seed=i)
    df = smooth_df(tuple(warp.grid_dim(n, f) for n in size), amp, seed=100 + i, shift=shift)
    grid = warper.grid_from_array(df, f)
    a, b = both_tiles(warper, lambda: warper.resample_tps(vol, grid, precision=prec))
    assert_same(a, b, CASES[i])


def test_tiles_identical_two_channels(warper):
    size, f = (90, 60, 36), 8
    vol = volume(size, np.uint16, seed=7, nc=2)
    df = smooth_df(tuple(warp.grid_dim(n, f) for n in size), 1.0, seed=8, shift=(4.0, -3.0, 2.0))
    grid = warper.grid_from_array(df, f)
    a, b = both_tiles(warper, lambda: warper.resample_tps(vol, grid, precision=warp.MIXED))
    assert_same(a, b, "two channels")


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8])
def test_tiles_identical_windows(warper, dtype):
    """Source windows of a larger volume (corner, edge, interior) and output
    boxes at their global origins, against grids anchored at the box."""
    full = (300, 260, 120)
    vol = volume(full, dtype, seed=21)
    f = 8
    for k, (o, box, margin) in enumerate([((0, 0, 0), (64, 48, 40), 12), ((236, 212, 80), (64, 48, 40), 12),
                                          ((100, 96, 40), (80, 64, 48), 6), ((150, 0, 72), (150, 100, 48), 3)]):
        lo = tuple(max(0, o[a] - margin) for a in range(3))
        hi = tuple(min(full[a], o[a] + box[a] + margin) for a in range(3))
        win = np.ascontiguousarray(vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])
        df = smooth_df(tuple(warp.grid_dim(n, f) for n in box), 1.5, seed=30 + k, shift=(margin - 1.5, -2.0, 1.0))
        grid = warper.grid_from_array(df, f, origin=o)
        for prec in (warp.PLAIN, warp.MIXED):
            def run():
                out = np.zeros(box[::-1], dtype)
                return warper.resample_tps(win, grid, dst=out, precision=prec, src_origin=lo, dst_origin=o, full=full)
            a, b = both_tiles(warper, run)
            assert_same(a, b, (o, box, prec))


def test_tiles_identical_fp16_grid_and_oracle(warper, oracle):
    """A TPS warp with an FP16 grid (TW_MIXED) and an FP32 grid (TW_PLAIN):
    tiles identical; the LDS tile within the oracle bars."""
    import torch
    size = (160, 128, 64)
    vol = volume(size, np.uint16, seed=41)
    sub, tar = synth.landmarks(size, 64, 6.0, seed=41)
    src = torch.from_numpy(vol.view(np.int16)).cuda()
    for prec in (warp.MIXED, warp.PLAIN):
        def run():
            dst = torch.zeros_like(src)
            warper.warp_tps(src, tar, sub, f=8, lam=0.2, precision=prec, dst=dst)
            return dst.cpu().numpy().view(np.uint16)
        a, b = both_tiles(warper, run)
        assert_same(a, b, prec)
    ref, df, _ = oracle.warp_tps(vol, tar, sub, g=8, lam=0.2)
    check_fast_form(b, ref, oracle.dense_df(size, df, 8), size)


def test_tiles_identical_bricks(warper):
    """tw_resample_tps_bricks with the LDS tile equals the gather tile and
    the per-brick single calls (default tile)."""
    import torch
    full = (256, 192, 96)
    vol = volume(full, np.uint16, seed=51)
    f, brick, margin = 8, 64, 10
    gshape = tuple(warp.grid_dim(n, f) for n in full)
    df = torch.from_numpy(smooth_df(gshape, 1.5, seed=52, shift=(3.0, -2.0, 1.0))).cuda()
    grid = warper.grid_from_array(df, f)
    srcs, dsts, so, do = [], [], [], []
    for bz in range(0, full[2], brick // 2):
        for by in range(0, full[1], brick):
            for bx in range(0, full[0], brick):
                o = (bx, by, bz)
                box = (min(brick, full[0] - bx), min(brick, full[1] - by), min(brick // 2, full[2] - bz))
                lo = tuple(max(0, o[a] - margin) for a in range(3))
                hi = tuple(min(full[a], o[a] + box[a] + margin) for a in range(3))
                w = np.ascontiguousarray(vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])
                srcs.append(torch.from_numpy(w.view(np.int16)).cuda())
                dsts.append(torch.zeros(box[::-1], dtype=torch.int16, device="cuda"))
                so.append(lo)
                do.append(o)

    def run():
        for d in dsts:
            d.zero_()
        warper.resample_bricks(srcs, dsts, grid, so, do, full, precision=warp.MIXED)
        torch.cuda.synchronize()
        return [d.cpu().numpy().view(np.uint16).copy() for d in dsts]

    a, b = both_tiles(warper, run)
    for i, (x, y) in enumerate(zip(a, b)):
        assert_same(x, y, ("brick", i))
    for i in (0, len(srcs) // 2, len(srcs) - 1):
        one = torch.zeros_like(dsts[i])
        warper.resample_tps(srcs[i], grid, dst=one, precision=warp.MIXED, src_origin=so[i], dst_origin=do[i],
                            full=full)
        torch.cuda.synchronize()
        assert_same(one.cpu().numpy().view(np.uint16), b[i], ("single", i))
