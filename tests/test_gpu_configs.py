"""The precision, streaming and batched-instance configurations at their own
workloads, on the GPU, against the FP64 oracle.

Bars (stated per test):
  precision (2048x2048x512 uint16, K = 4096 lattice + 10 % outliers,
    |d| <= 40, grid 8): fit coefficients within 1e-8 (relative) of the
    oracle's bordered solve; voxel displacement within 1e-2 voxel; on two
    sampled block rows the plain and mixed chains within one level off the
    volume's faces and on >= 99.99 % of voxels (tests/parity.py), plain
    bit-identical on >= 99.9 %
  streaming (8192x8192x2048 uint16, 256 GiB, K = 2048, |d| <= 40, 512^3
    bricks): sampled corner, edge and interior bricks (16 slices each) against
    the oracle's window form with the same bars; two bricks through the
    reference-order form (TW_EXACT) on the oracle's grid bit-identical to the
    oracle, checksums included
  batch (256 uneven instances): every item bit-identical to its single call;
    sampled TPS items within the bars above, affine and rigid items
    bit-identical to the oracle
"""
import numpy as np
import pytest

from terawarp_amd import Synth, checksum64, synth, warp

from .parity import DF_TOL, check_fast_form

pytestmark = pytest.mark.gpu
THREADS = 16


def _host(t, dtype=np.uint16):
    a = t.cpu().numpy()
    return a.view(np.uint16) if dtype == np.uint16 else a


def test_precision_configuration(warper, oracle):
    import torch
    size, K, dmax, g, seed = (2048, 2048, 512), 4096, 40.0, 8, 3
    vol = np.empty(size[::-1], np.uint16)
    Synth(size, np.uint16, seed=seed, cell=256, blobs_per_cell=16).fill_into(vol, 0, 0, 0, *size, THREADS)
    sub, tar = synth.landmarks(size, K, dmax, seed=seed, layout="lattice", outliers=0.1)
    c_ref, d_ref = oracle.stps_fit(tar, sub, nthreads=THREADS, bordered=True)
    zb0, zb1 = 31, 33
    gsh = oracle.grid_shape(size, g)
    df_sl = oracle.grid_eval(tar, c_ref, d_ref, (gsh[0], gsh[1], zb1 - zb0 + 3), g, origin=(0, 0, zb0 * g),
                             nthreads=THREADS)
    box = (size[0], size[1], (zb1 - zb0) * g)
    dense_ref = oracle.dense_df(box, df_sl, g, nthreads=THREADS)
    full = np.zeros((3, gsh[2], gsh[1], gsh[0]), np.float32)
    full[:, zb0:zb0 + df_sl.shape[1]] = df_sl
    sl = slice(zb0 * g, zb1 * g)
    ref = oracle.tps_resample(vol, full, g, zblocks=(zb0, zb1), nthreads=THREADS)[sl]
    del full
    src = torch.from_numpy(vol.view(np.int16)).cuda()
    dst = torch.empty_like(src)
    for prec in (warp.PLAIN, warp.MIXED):
        c, d = warper.tps_fit(tar, sub, precision=prec)
        assert np.abs(c - c_ref).max() <= 1e-8 * np.abs(c_ref).max(), prec
        assert np.abs(d - d_ref).max() <= 1e-8 * np.abs(d_ref).max(), prec
        grid = warper.new_grid(gsh, g, storage=warp.F16 if prec == warp.MIXED else warp.F32, device=True)
        warper.grid_eval(tar, c, d, grid, precision=prec)
        dense = warper.upsample_df(grid, box, origin=(0, 0, zb0 * g), precision=prec)
        assert np.abs(dense - dense_ref).max() <= DF_TOL, prec
        grid.release()
        warper.warp_tps(src, tar, sub, f=g, lam=0.2, precision=prec, dst=dst)
        got = _host(dst[sl])
        st = check_fast_form(got, ref, dense_ref, size, origin=(0, 0, zb0 * g))
        if prec == warp.PLAIN:
            assert st["exact"] >= 0.999, st


def test_streaming_configuration(warper, oracle):
    size, K, dmax, g, brick, seed = (8192, 8192, 2048), 2048, 40.0, 8, 512, 4
    src = Synth(size, np.uint16, seed=seed, cell=256, blobs_per_cell=16)
    sub, tar = synth.landmarks(size, K, dmax, seed=seed)
    nbx, nby, nbz = (n // brick for n in size)
    picks = [(0, 0, 0), (nbx - 1, nby - 1, nbz - 1), (0, nby - 1, 0), (nbx // 2, nby // 2, nbz // 2),
             (nbx - 1, nby // 2, 1)]
    sample = sorted({bx + nbx * (by + nby * bz) for bx, by, bz in picks})
    sums, kept, st = warper.warp_stream(size, np.uint16, tar, sub, src, f=g, brick=brick, lam=0.2,
                                        precision=warp.MIXED, threads=THREADS, keep=sample)
    assert st["bricks"] == nbx * nby * nbz
    c, d = oracle.stps_fit(tar, sub, nthreads=THREADS, bordered=True)
    nz = 16
    margin = int(dmax) + 8
    for i, b in enumerate(sample):
        bx, by, bz = b % nbx, (b // nbx) % nby, b // (nbx * nby)
        o = (bx * brick, by * brick, bz * brick)
        box = (brick, brick, nz)
        lo = tuple(max(0, o[a] - margin) for a in range(3))
        hi = tuple(min(size[a], o[a] + box[a] + margin) for a in range(3))
        win = np.empty((hi[2] - lo[2], hi[1] - lo[1], hi[0] - lo[0]), np.uint16)
        src.fill_into(win, lo[0], lo[1], lo[2], hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], THREADS)
        df = oracle.grid_eval(tar, c, d, tuple(warp.grid_dim(v, g) for v in box), g, origin=o, nthreads=THREADS)
        ref = oracle.tps_resample_window(win, lo, size, df, g, o, box, o, nthreads=THREADS)
        got = kept[b][:nz]
        assert int(sums[b]) == checksum64(kept[b]), b
        check_fast_form(got, ref, oracle.dense_df(box, df, g, nthreads=THREADS), size, origin=o)
        if i < 2:
            # The reference-order form on the oracle's grid: bit-identical.
            grid = warper.grid_from_array(df, g, origin=o)
            ex = np.zeros(box[::-1], np.uint16)
            warper.resample_tps(win, grid, dst=ex, precision=warp.EXACT, src_origin=lo, dst_origin=o, full=size)
            assert np.array_equal(ex, ref), (b, np.count_nonzero(ex != ref))
            assert checksum64(ex) == checksum64(ref)


def test_batch_configuration(warper, oracle):
    import torch
    specs = synth.batch_instances(seed=5, n=256, dmax=8.0)
    srcs = [synth.blob_volume_device(sp["size"], sp["dtype"], sp["seed"]) for sp in specs]
    # Outputs start as different byte patterns for the batch and the single
    # calls: a voxel either leaves unwritten differs between the two.
    def filled(s, byte):
        return torch.full_like(s, byte if s.dtype == torch.uint8 else (byte << 8 | byte) - (1 << 16 if byte >= 128 else 0))
    dsts = [filled(s, 0xA5) for s in srcs]
    items = []
    for sp, s, d in zip(specs, srcs, dsts):
        it = {"src": s, "dst": d}
        if sp["kind"] == "tps":
            it.update(centres=sp["centres"], targets=sp["targets"])
        else:
            it["M"] = sp["M"]
        items.append(it)
    status, _ = warper.warp_batch(items, f=8, lam=0.2, precision=warp.MIXED)
    assert all(s == 0 for s in status)
    one = [filled(s, 0x5A) for s in srcs]
    for sp, s, o in zip(specs, srcs, one):
        if sp["kind"] == "tps":
            warper.warp_tps(s, sp["centres"], sp["targets"], f=8, lam=0.2, precision=warp.MIXED, dst=o)
        else:
            warper.resample_affine(s, sp["M"], dst=o)
    torch.cuda.synchronize()
    same = [bool(torch.equal(a, b)) for a, b in zip(dsts, one)]
    assert all(same), [i for i, x in enumerate(same) if not x][:8]
    vox = np.array([np.prod(sp["size"]) for sp in specs])
    order = np.argsort(vox)
    picked = []
    for kind, cnt in (("tps", 3), ("affine", 2), ("rigid", 2)):
        picked += [int(i) for i in order if specs[i]["kind"] == kind][:cnt]
    for i in picked:
        sp = specs[i]
        v = _host(srcs[i], sp["dtype"])
        out = _host(dsts[i], sp["dtype"])
        if sp["kind"] == "tps":
            ref, df, _ = oracle.warp_tps(v, sp["centres"], sp["targets"], g=8, lam=0.2, nthreads=THREADS)
            check_fast_form(out, ref, oracle.dense_df(sp["size"], df, 8, nthreads=THREADS), sp["size"])
        else:
            assert np.array_equal(out, oracle.affine_resample(v, sp["M"], nthreads=THREADS)), i
