"""Bricked streaming (tw_warp_stream) vs the in-memory warp and the oracle.

Bars (stated per test):
  TW_PLAIN / TW_MIXED   the assembled bricks are bit-identical to tw_warp_tps on
                        the whole volume (same grid, same per-voxel arithmetic;
                        exact source windows cover every tap)
  TW_EXACT              every brick's checksum equals that of the same box of
                        the FP64 oracle's in-memory warp
  failures              an error in the read callback ends the call with an
                        error (no hang); bad descriptors are rejected
"""
import numpy as np
import pytest

from terawarp_amd import Synth, checksum64, synth, warp

pytestmark = pytest.mark.gpu


def boxes(size, brick):
    nbx, nby, nbz = (-(-n // brick) for n in size)
    for b in range(nbx * nby * nbz):
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
        yield b, (slice(bz * brick, min((bz + 1) * brick, size[2])), slice(by * brick, min((by + 1) * brick, size[1])),
                  slice(bx * brick, min((bx + 1) * brick, size[0])))


def in_memory(warper, vol, tar, sub, f, precision, interp):
    import torch
    src = torch.from_numpy(vol.view(np.int16) if vol.dtype == np.uint16 else vol).cuda()
    dst = torch.zeros_like(src)
    warper.warp_tps(src, tar, sub, f=f, lam=0.2, interp=interp, precision=precision, dst=dst)
    torch.cuda.synchronize()
    out = dst.cpu().numpy()
    return out.view(np.uint16) if vol.dtype == np.uint16 else out


@pytest.mark.parametrize("precision,interp", [(warp.PLAIN, warp.LINEAR), (warp.MIXED, warp.LINEAR),
                                              (warp.MIXED, warp.NN), (warp.MIXED, warp.LINEAR | warp.DF_LINEAR)])
def test_stream_equals_in_memory_warp(warper, precision, interp):
    size, brick, f = (200, 150, 90), 64, 8          # 4 x 3 x 2 bricks, uneven on every axis
    s = Synth(size, np.uint16, seed=4, cell=64, blobs_per_cell=4)
    vol = s.fill(0, 0, 0, *size)
    sub, tar = synth.landmarks(size, 48, 6.0, seed=4)
    ref = in_memory(warper, vol, tar, sub, f, precision, interp)
    nb = 4 * 3 * 2
    sums, kept, st = warper.warp_stream(size, np.uint16, tar, sub, s, f=f, brick=brick, precision=precision,
                                        interp=interp, threads=4, keep=range(nb))
    assert st["bricks"] == nb
    for b, box in boxes(size, brick):
        assert kept[b].shape == ref[box].shape
        assert np.array_equal(kept[b], ref[box]), (b, np.count_nonzero(kept[b] != ref[box]))
        assert int(sums[b]) == checksum64(ref[box])
    # Exact windows: no wider than the field's reach plus the rounding margin.
    assert 0 < st["max_margin"] <= 6 + 4
    # The issuing thread's own enqueue time lies inside the run's wall time.
    assert 0 < st["issue_s"] < st["wall_s"]


def test_stream_exact_checksums_match_oracle(warper, oracle):
    size, brick, f = (130, 97, 70), 32, 4
    s = Synth(size, np.uint8, seed=8, cell=32, blobs_per_cell=4)
    vol = s.fill(0, 0, 0, *size)
    sub, tar = synth.landmarks(size, 30, 4.0, seed=8)
    ref = oracle.warp_tps(vol, tar, sub, g=f, lam=0.2)[0]
    sums, _, st = warper.warp_stream(size, np.uint8, tar, sub, s, f=f, brick=brick, precision=warp.EXACT,
                                     threads=3)
    for b, box in boxes(size, brick):
        assert int(sums[b]) == checksum64(ref[box]), b
    assert st["h2d_bytes"] > 0 and st["d2h_bytes"] == vol.nbytes


def test_stream_block_origin_exact_equals_tera_oracle(warper, oracle):
    """TW_STREAM_BLOCK_ORIGIN with TW_EXACT: bricks are the tera driver's
    blocks and positions take its block-local float order; the assembled
    output equals the oracle's restatement of the tera TPS driver (whose
    marker files carry a quarter of the coordinates) bit for bit, uint16,
    bricks uneven on every axis."""
    size, brick, f = (150, 97, 70), 40, 4
    s = Synth(size, np.uint16, seed=11, cell=32, blobs_per_cell=4)
    vol = s.fill(0, 0, 0, *size)
    sub, tar = synth.landmarks(size, 30, 6.0, seed=11)
    sub4, tar4 = sub / 4, tar / 4
    ref = oracle.warp_tps_tera(vol, tar4, sub4, g=f, block=brick)
    nb = 4 * 3 * 2
    _, kept, st = warper.warp_stream(size, np.uint16, 4 * tar4, 4 * sub4, s, f=f, brick=brick,
                                     precision=warp.EXACT, threads=3, keep=range(nb), block_origin=True)
    assert st["bricks"] == nb
    for b, box in boxes(size, brick):
        assert np.array_equal(kept[b], ref[box]), (b, np.count_nonzero(kept[b] != ref[box]))


def test_stream_python_source_and_read_failure(warper):
    size, brick, f = (96, 80, 48), 32, 8
    s = Synth(size, np.uint8, seed=2, cell=32, blobs_per_cell=2)
    vol = s.fill(0, 0, 0, *size)
    sub, tar = synth.landmarks(size, 20, 3.0, seed=2)

    def source(x0, y0, z0, nx, ny, nz):
        return vol[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx]

    a, _, _ = warper.warp_stream(size, np.uint8, tar, sub, s, f=f, brick=brick, precision=warp.MIXED, threads=2)
    b, _, _ = warper.warp_stream(size, np.uint8, tar, sub, source, f=f, brick=brick, precision=warp.MIXED, threads=2)
    assert np.array_equal(a, b)

    def failing(x0, y0, z0, nx, ny, nz):
        if z0 >= 32:
            raise IOError("read failed")
        return source(x0, y0, z0, nx, ny, nz)

    with pytest.raises(warp.TeraWarpError) as e:
        warper.warp_stream(size, np.uint8, tar, sub, failing, f=f, brick=brick, precision=warp.MIXED, threads=2)
    assert "read callback failed" in str(e.value)
    # The context still works afterwards.
    c, _, _ = warper.warp_stream(size, np.uint8, tar, sub, s, f=f, brick=brick, precision=warp.MIXED, threads=2)
    assert np.array_equal(a, c)


def test_stream_rejects_bad_descriptors(warper):
    size = (64, 64, 32)
    s = Synth(size, np.uint8)
    sub, tar = synth.landmarks(size, 10, 2.0, seed=1)
    with pytest.raises(warp.TeraWarpError):  # brick not a multiple of the spacing
        warper.warp_stream(size, np.uint8, tar, sub, s, f=8, brick=20)
    with pytest.raises(warp.TeraWarpError):  # spacing not below the image size
        warper.warp_stream(size, np.uint8, tar, sub, s, f=32, brick=32)


def test_stream_to_block_files_and_slices(warper, tmp_path):
    """The tera driver's output layout: block files from the stream, stitched
    into slices that equal the in-memory warp."""
    from terawarp_amd import _lib as L
    from terawarp_amd.warp import stitch_slices, tera_paths
    size, brick, f = (100, 70, 40), 32, 8
    s = Synth(size, np.uint8, seed=6, cell=32, blobs_per_cell=3)
    vol = s.fill(0, 0, 0, *size)
    sub, tar = synth.landmarks(size, 24, 4.0, seed=6)
    ref = in_memory(warper, vol, tar, sub, f, warp.MIXED, warp.LINEAR)
    base = str(tmp_path / "tw")
    warper.warp_stream(size, np.uint8, tar, sub, s, f=f, brick=brick, precision=warp.MIXED, threads=2,
                       block_base=base)
    stitch_slices(base, size, brick, np.uint8)
    lib = L.load()
    for z in range(size[2]):
        out = np.zeros((size[1], size[0]), np.uint8)
        assert lib.tw_v3draw_read(tera_paths(base, z=z)[1].encode(), out.ctypes.data, out.nbytes) == 0
        assert np.array_equal(out, ref[z]), z


@pytest.mark.parametrize("precision,f16", [(warp.PLAIN, False), (warp.MIXED, False), (warp.MIXED, True)])
def test_resample_bricks_equals_single_calls(warper, precision, f16):
    """tw_resample_tps_bricks: every brick of one launch equals
    tw_resample_tps on the same source window, output box and grid, and the
    bricks assemble the whole-volume resample (uneven bricks, windows from
    the field's bound)."""
    import torch
    size, brick, g = (200, 150, 90), 64, 8
    s = Synth(size, np.uint16, seed=6, cell=64, blobs_per_cell=4)
    vol = s.fill(0, 0, 0, *size)
    sub, tar = synth.landmarks(size, 40, 6.0, seed=6)
    c, d = warper.tps_fit(tar, sub, precision=warp.PLAIN)
    gsh = tuple(warp.grid_dim(n, g) for n in size)
    grid = warper.new_grid(gsh, g, storage=warp.F16 if f16 else warp.F32, device=True)
    warper.grid_eval(tar, c, d, grid, precision=precision)
    m = 6 + 4  # field bound plus the rounding margin
    srcs, dsts, so, do, boxes_ = [], [], [], [], []
    for b, box in boxes(size, brick):
        o = (box[2].start, box[1].start, box[0].start)
        n = (box[2].stop - o[0], box[1].stop - o[1], box[0].stop - o[2])
        lo = [max(0, o[a] - m) for a in range(3)]
        hi = [min(size[a], o[a] + n[a] + m) for a in range(3)]
        win = np.ascontiguousarray(vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])
        srcs.append(torch.from_numpy(win.view(np.int16)).cuda())
        dsts.append(torch.zeros((n[2], n[1], n[0]), dtype=torch.int16, device="cuda"))
        so.append(tuple(lo))
        do.append(o)
        boxes_.append(box)
    warper.resample_bricks(srcs, dsts, grid, so, do, size, precision=precision)
    # No explicit synchronisation: the wrapper orders the library's stream
    # after torch's (the zero fills) and torch's after the library's.
    whole = torch.zeros((size[2], size[1], size[0]), dtype=torch.int16, device="cuda")
    warper.resample_tps(torch.from_numpy(vol.view(np.int16)).cuda(), grid, dst=whole, precision=precision)
    whole = whole.cpu().numpy().view(np.uint16)
    for i in range(len(srcs)):
        one = torch.zeros_like(dsts[i])
        warper.resample_tps(srcs[i], grid, dst=one, precision=precision, src_origin=so[i], dst_origin=do[i],
                            full=size)
        got = dsts[i].cpu().numpy().view(np.uint16)
        assert np.array_equal(got, one.cpu().numpy().view(np.uint16)), i
        assert np.array_equal(got, whole[boxes_[i]]), i
    grid.release()
