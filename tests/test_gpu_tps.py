"""GPU TPS path vs the FP64 oracle.

Bars (stated per test):
  TW_EXACT grid / upsample / resample  bit-identical to the oracle
  TW_PLAIN resample                   |diff| <= 1 LSB on >= 99.99 % of voxels, exact on >= 99.9 %
  TW_PLAIN / TW_MIXED grid field      max |DF - DF_oracle| <= 1e-2 voxel (typically ~1e-4)
  fit (FP64 LU / FP32 LU + FP64 IR)   coefficients within 1e-8 relative of the oracle
"""
import os

import numpy as np
import pytest

from terawarp_amd import synth, warp

from .parity import check_fast_form

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DF_TOL = 1e-2  # voxels


def case(size, dtype, K, dmax, seed, n_blobs=60):
    amp = (20, 200) if dtype == np.uint8 else (100, 4000)
    vol = synth.blob_volume(size[::-1], dtype, n_blobs=n_blobs, amp=amp, seed=seed)
    sub, tar = synth.landmarks(size, K, dmax, seed=seed)
    return vol, sub, tar


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_fit_large_system_matches_oracle(warper, oracle):
    """K = 1200: the projected system (1196 unknowns) is solved by per-block
    launches over many work-groups; coefficients match the oracle's bordered
    FP64 solve."""
    sub, tar = synth.landmarks((1024, 1024, 256), 1200, 20.0, seed=11)
    c_ref, d_ref = oracle.stps_fit(tar, sub, bordered=True)
    for precision in (warp.PLAIN, warp.MIXED):
        c, d = warper.tps_fit(tar, sub, precision=precision)
        assert rel_err(c, c_ref) < 1e-8, (precision, rel_err(c, c_ref))
        assert np.abs(d - d_ref).max() < 1e-8
        assert warper.timings()["fit_residual"] < 1e-10


@pytest.mark.parametrize("K", [257, 1024, 1028, 1029])
def test_fit_path_boundaries_match_oracle(warper, oracle, K):
    """Fits at the boundaries of the large-system paths: K = 257 (fewest
    landmarks on the kernel sequence; a partial last 64-block), 1024 (the
    one-wave QR at its largest N; 16 blocks, the last partial, in the fused
    factor steps that also run the forward substitution), 1028 (16 full
    blocks; four-wave QR) and 1029 (17 blocks: per-block-column factor and
    the two-pass pipelined solve).  Coefficients match the oracle's bordered
    FP64 solve."""
    sub, tar = synth.landmarks((512, 512, 256), K, 10.0, seed=K)
    c_ref, d_ref = oracle.stps_fit(tar, sub, bordered=True)
    c, d = warper.tps_fit(tar, sub, precision=warp.PLAIN)
    assert rel_err(c, c_ref) < 1e-8, rel_err(c, c_ref)
    assert np.abs(d - d_ref).max() < 1e-8
    assert warper.timings()["fit_residual"] < 1e-10


def test_fit_four_wave_qr_at_its_largest(warper):
    """K = 4096: the four-wave QR kernels at their largest N (16 rows per
    thread).  The device's FP64 residual of the bordered system (formed from
    the landmarks, independent of the QR) certifies the solution, and the
    chain's affine part reproduces an exact affine map (pure translation
    targets give d = identity + shift, c = 0)."""
    sub, tar = synth.landmarks((2048, 2048, 512), 4096, 20.0, seed=5)
    c, d = warper.tps_fit(tar, sub, precision=warp.PLAIN)
    assert warper.timings()["fit_residual"] < 1e-10
    # Integer centres and shift: the shifted centres are exact in FP32.
    p = np.unique(np.round(sub), axis=0)[:4096].astype(np.float32)
    assert len(p) == 4096
    shift = np.array([3.0, -2.0, 5.0], np.float32)
    c2, d2 = warper.tps_fit(p + shift, p, precision=warp.PLAIN)
    assert warper.timings()["fit_residual"] < 1e-10
    assert np.abs(c2).max() < 1e-9, np.abs(c2).max()
    assert np.abs(d2[0, 1:] + shift).max() < 1e-6, d2[0]
    assert np.abs(d2[1:, 1:] - np.eye(3)).max() < 1e-9, d2[1:, 1:]


@pytest.mark.parametrize("K", [1029, 2048, 4200])
def test_fit_fp32_factor_with_refinement(warper, oracle, K):
    """TW_MIXED's FP32 factor (MFMA trailing updates) with FP64 refinement,
    forced on below its default threshold: it converges without the FP64
    refactor (refine_iters > 0) and the coefficients match the oracle's
    bordered FP64 solve and the FP64 factor's."""
    sub, tar = synth.landmarks((2048, 2048, 512), K, 20.0, seed=7)
    c64, d64 = warper.tps_fit(tar, sub, precision=warp.PLAIN)
    default = warper.get_tune("fp32_factor_min")
    try:
        warper.set_tune("fp32_factor_min", 1)
        c32, d32 = warper.tps_fit(tar, sub, precision=warp.MIXED)
        t = warper.timings()
    finally:
        warper.set_tune("fp32_factor_min", default)
    assert t["refine_iters"] > 0, t["refine_iters"]
    assert rel_err(c32, c64) < 1e-8, rel_err(c32, c64)
    assert rel_err(d32, d64) < 1e-8, rel_err(d32, d64)
    assert t["fit_residual"] < 1e-10
    if K <= 2048:
        c_ref, d_ref = oracle.stps_fit(tar, sub, bordered=True)
        assert rel_err(c32, c_ref) < 1e-8, rel_err(c32, c_ref)
        assert np.abs(d32 - d_ref).max() < 1e-8


def test_tunables_round_trip(warper):
    """tw_set_tune / tw_get_tune: values stick, bad keys and values are
    rejected with TW_E_ARG, and the default threshold keeps config-size
    mixed fits on the FP64 factor (refine_iters == 0)."""
    for key, val in (("tile", 1), ("tz", 32), ("fp32_factor_min", 5000)):
        old = warper.get_tune(key)
        warper.set_tune(key, val)
        assert warper.get_tune(key) == val
        warper.set_tune(key, old)
    with pytest.raises(warp.TeraWarpError):
        warper.set_tune("tz", 999)
    with pytest.raises(warp.TeraWarpError):
        warper.set_tune("tile", 7)
    sub, tar = synth.landmarks((1024, 1024, 256), 2048, 20.0, seed=9)
    warper.tps_fit(tar, sub, precision=warp.MIXED)
    assert warper.timings()["refine_iters"] == 0


@pytest.mark.parametrize("K", [600, 1500])
def test_solve_wait_bound_falls_back_to_block_launches(warper, oracle, K):
    """The pipelined solve's waits between work-groups are bounded.  With the
    bound at 0 every wait runs out at once (each work-group stops waiting and
    the kernel ends normally), the fit state reports it, and the library
    turns the pipelined solve off for the context and redoes the fit with one
    launch per block.  K = 600: fused factor steps, whose forward substitution
    the redo leaves to the per-block solve; K = 1500: per-column factor.  Each
    entry point that fits (tw_tps_fit, the warp chain, the batch) is forced
    through the fallback on its own (setting the bound turns the pipelined
    solve back on) and must match the oracle, or the default context's
    output."""
    import torch
    sub, tar = synth.landmarks((512, 512, 128), K, 10.0, seed=K + 3)
    c_ref, d_ref = oracle.stps_fit(tar, sub, bordered=True)
    size = (96, 80, 40)
    vol = synth.blob_volume(size[::-1], np.uint16, n_blobs=40, amp=(100, 4000), seed=K)
    lm_t, lm_s = tar * 0.15, sub * 0.15  # landmarks inside the small volume
    out_ref = warper.warp_tps(vol, lm_t, lm_s, f=4, precision=warp.PLAIN)
    src = torch.from_numpy(vol.view(np.int16)).cuda()
    dst_ref = torch.empty_like(src)
    warper.warp_batch([{"src": src, "dst": dst_ref, "centres": lm_t, "targets": lm_s}], f=4,
                      precision=warp.PLAIN)
    torch.cuda.synchronize()
    w = warp.Warper(0)
    try:
        assert w.get_tune("solve_wait_us") == 20000
        assert w.get_tune("solve_pipelined") == 1
        with pytest.raises(warp.TeraWarpError):
            w.set_tune("solve_wait_us", -1)
        with pytest.raises(warp.TeraWarpError):
            w.set_tune("solve_pipelined", 1)
        w.set_tune("solve_wait_us", 0)
        c, d = w.tps_fit(tar, sub, precision=warp.PLAIN)
        assert w.get_tune("solve_pipelined") == 0  # the wait ran out; the fit was redone
        assert rel_err(c, c_ref) < 1e-8, rel_err(c, c_ref)
        assert np.abs(d - d_ref).max() < 1e-8
        w.set_tune("solve_wait_us", 0)
        out = w.warp_tps(vol, lm_t, lm_s, f=4, precision=warp.PLAIN)
        assert w.get_tune("solve_pipelined") == 0
        diff = np.abs(out.astype(np.int64) - out_ref.astype(np.int64))
        assert diff.max() <= 1 and np.mean(diff == 0) >= 0.9999, (diff.max(), np.mean(diff == 0))
        w.set_tune("solve_wait_us", 0)
        dst = torch.empty_like(src)
        status, _ = w.warp_batch([{"src": src, "dst": dst, "centres": lm_t, "targets": lm_s}], f=4,
                                 precision=warp.PLAIN)
        assert status == [0] and w.get_tune("solve_pipelined") == 0
        torch.cuda.synchronize()
        bd = (dst.cpu().numpy().view(np.uint16).astype(np.int64) - dst_ref.cpu().numpy().view(np.uint16))
        assert np.abs(bd).max() <= 1 and np.mean(bd == 0) >= 0.9999
        w.set_tune("solve_wait_us", 20000)
        c2, _ = w.tps_fit(tar, sub, precision=warp.PLAIN)  # pipelined again
        assert rel_err(c2, c_ref) < 1e-8 and w.get_tune("solve_pipelined") == 1
    finally:
        w.close()


@pytest.mark.parametrize("K", [16, 17, 64, 100, 256, 512])
@pytest.mark.parametrize("precision", [warp.PLAIN, warp.MIXED])
def test_fit_matches_oracle(warper, oracle, K, precision):
    size = (256, 256, 64)
    _, sub, tar = case(size, np.uint8, K, 5.0, seed=K, n_blobs=1)
    c_ref, d_ref = oracle.stps_fit(tar, sub)
    c, d = warper.tps_fit(tar, sub, precision=precision)
    t = warper.timings()
    assert rel_err(c, c_ref) < 1e-8, rel_err(c, c_ref)
    assert np.abs(d - d_ref).max() < 1e-8
    assert t["fit_residual"] < 1e-10


def test_fit_rejects_degenerate(warper):
    P = np.zeros((10, 3), np.float32)
    P[:, 0] = np.arange(10)
    with pytest.raises(warp.TeraWarpError):
        warper.tps_fit(P, P)
    with pytest.raises(warp.TeraWarpError):
        warper.tps_fit(P[:4], P[:4])


def _quantize_f16(v):
    """FP16 offset-scaled storage restated in numpy (float32 arithmetic):
    per 8 x 8 x 8 node block and component, offset = (lo + hi) / 2 and scale =
    max(hi - offset, offset - lo) (1 where that is 0) over the block's nodes,
    stored h = (v - offset) / scale clamped to [-1, 1], rounded to nearest-even
    FP16.  v: (3, gz, gy, gx) float32."""
    out = np.empty(v.shape, np.float16)
    _, gz, gy, gx = v.shape
    half = np.float32(0.5)
    for z0 in range(0, gz, 8):
        for y0 in range(0, gy, 8):
            for x0 in range(0, gx, 8):
                b = v[:, z0:z0 + 8, y0:y0 + 8, x0:x0 + 8]
                lo = b.min(axis=(1, 2, 3))
                hi = b.max(axis=(1, 2, 3))
                off = half * (lo + hi)
                sc = np.maximum(hi - off, off - lo)
                sc = np.where(sc > 0, sc, np.float32(1)).astype(np.float32)
                h = (b - off[:, None, None, None]) / sc[:, None, None, None]
                out[:, z0:z0 + 8, y0:y0 + 8, x0:x0 + 8] = np.clip(h, np.float32(-1), np.float32(1)).astype(np.float16)
    return out


def test_fp16_grid_storage_bitwise(warper):
    """The mixed form's FP16 offset-scaled grid equals the numpy restatement
    of the quantisation applied to the same fit's FP32 grid, bit for bit;
    the grid's extents leave partial 8-node blocks on every axis and a
    partial run of 8 blocks along x."""
    size = (1190, 550, 150)
    sub, tar = synth.landmarks(size, 200, 20.0, seed=41)
    c, d = warper.tps_fit(tar, sub, precision=warp.MIXED)
    gsh = (150, 70, 20)
    g32 = warper.new_grid(gsh, 8, device=False)
    warper.grid_eval(tar, c, d, g32, precision=warp.PLAIN)
    g16 = warper.new_grid(gsh, 8, storage=warp.F16, device=True)
    warper.grid_eval(tar, c, d, g16, precision=warp.MIXED)
    got = g16.data.cpu().numpy()
    exp = _quantize_f16(g32.data)
    assert np.array_equal(got.view(np.uint16), exp.view(np.uint16)), np.count_nonzero(got.view(np.uint16) != exp.view(np.uint16))
    g16.release()


@pytest.mark.parametrize("g", [4, 8])
def test_grid_eval_exact_bitwise(warper, oracle, g):
    size = (200, 150, 60)
    _, sub, tar = case(size, np.uint8, 64, 8.0, seed=g, n_blobs=1)
    c, d = oracle.stps_fit(tar, sub)
    gsh = tuple(warp.grid_dim(s, g) for s in size)
    ref = oracle.grid_eval(tar, c, d, gsh, g, origin=(16, 0, 8))
    grid = warper.new_grid(gsh, g, origin=(16, 0, 8), device=False)
    warper.grid_eval(tar, c, d, grid, precision=warp.EXACT)
    mism = np.count_nonzero(grid.data != ref)
    assert mism == 0, (mism, np.abs(grid.data - ref).max())


@pytest.mark.parametrize("K", [64, 512])
def test_grid_eval_fp32_accuracy(warper, oracle, K):
    size = (1024, 1024, 256)
    g = 8
    _, sub, tar = case(size, np.uint8, K, 20.0, seed=3, n_blobs=1)
    c, d = oracle.stps_fit(tar, sub)
    gsh = (warp.grid_dim(1024, g), warp.grid_dim(1024, g), 12)  # a z-slab of nodes
    ref = oracle.grid_eval(tar, c, d, gsh, g)
    grid = warper.new_grid(gsh, g, device=False)
    warper.grid_eval(tar, c, d, grid, precision=warp.PLAIN)
    err = np.abs(grid.data - ref).max()
    assert err < 1e-3, err


@pytest.mark.parametrize("g", [4, 8])
def test_upsample_exact_and_plain(warper, oracle, g):
    size = (70, 45, 33)
    _, sub, tar = case(size, np.uint8, 32, 6.0, seed=g, n_blobs=1)
    c, d = oracle.stps_fit(tar, sub)
    gsh = tuple(warp.grid_dim(s, g) for s in size)
    df = oracle.grid_eval(tar, c, d, gsh, g)
    ref = oracle.dense_df(size, df, g)
    grid = warper.grid_from_array(df, g)
    ex = warper.upsample_df(grid, size, precision=warp.EXACT)
    assert np.array_equal(ex, ref), np.abs(ex - ref).max()
    pl = warper.upsample_df(grid, size, precision=warp.PLAIN)
    assert np.abs(pl - ref).max() < 1e-5


@pytest.mark.parametrize("name", ["tps_u8", "tps_u16"])
def test_resample_fixture_exact(warper, name):
    f = dict(np.load(os.path.join(GOLD, name + ".npz")))
    g = int(f["g"])
    grid = warper.grid_from_array(f["df"], g)
    out = warper.resample_tps(f["vol"], grid, precision=warp.EXACT)
    assert np.array_equal(out, f["out"])
    out = warper.resample_tps(f["vol"], grid, precision=warp.PLAIN)
    diff = np.abs(out.astype(int) - f["out"].astype(int))
    assert diff.max() <= 1 and np.mean(diff == 0) >= 0.999


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("interp", [warp.LINEAR, warp.NN])
def test_resample_exact_bitwise(warper, oracle, dtype, interp):
    size = (131, 77, 45)
    g = 4
    vol, sub, tar = case(size, dtype, 40, 5.0, seed=17)
    c, d = oracle.stps_fit(tar, sub)
    gsh = tuple(warp.grid_dim(s, g) for s in size)
    df = oracle.grid_eval(tar, c, d, gsh, g)
    ref = oracle.tps_resample(vol, df, g, interp=interp)
    grid = warper.grid_from_array(df, g)
    got = warper.resample_tps(vol, grid, interp=interp, precision=warp.EXACT)
    assert np.array_equal(got, ref), np.count_nonzero(got != ref)


@pytest.mark.parametrize("dtype,size,g,K,dmax", [
    (np.uint8, (256, 256, 64), 4, 64, 5.0),      # CPU-path config
    (np.uint16, (300, 260, 70), 8, 128, 20.0),
])
def test_chain_plain_and_mixed_vs_oracle(warper, oracle, dtype, size, g, K, dmax):
    vol, sub, tar = case(size, dtype, K, dmax, seed=29, n_blobs=120)
    ref, df_ref, _ = oracle.warp_tps(vol, tar, sub, g=g)
    dense_ref = oracle.dense_df(size, df_ref, g)
    # Plain form: whole chain on the GPU.
    out = warper.warp_tps(vol, tar, sub, f=g, precision=warp.PLAIN)
    diff = np.abs(out.astype(np.int64) - ref.astype(np.int64))
    # Bar: within 1 LSB on >= 99.99 % of voxels (all of them here), and
    # bit-identical on >= 99.9 % (FP32 grid field vs the FP64 oracle).
    assert np.mean(diff <= 1) >= 0.9999 and diff.max() <= 1, diff.max()
    assert np.mean(diff == 0) >= 0.999, np.mean(diff == 0)
    # Stage errors: FP32 grid and its upsample against the FP64 oracle.
    c, d = warper.tps_fit(tar, sub, precision=warp.PLAIN)
    gsh = tuple(warp.grid_dim(s, g) for s in size)
    grid = warper.new_grid(gsh, g, device=False)
    warper.grid_eval(tar, c, d, grid, precision=warp.PLAIN)
    assert np.abs(grid.data - df_ref).max() < DF_TOL
    dense = warper.upsample_df(grid, size, precision=warp.PLAIN)
    assert np.abs(dense - dense_ref).max() < DF_TOL
    # Mixed form: FP16 grid, FP32 blend.
    cm, dm = warper.tps_fit(tar, sub, precision=warp.MIXED)
    g16 = warper.new_grid(gsh, g, storage=warp.F16, device=True)
    warper.grid_eval(tar, cm, dm, g16, precision=warp.MIXED)
    dense16 = warper.upsample_df(g16, size, precision=warp.MIXED)
    assert np.abs(dense16 - dense_ref).max() < DF_TOL
    outm = warper.warp_tps(vol, tar, sub, f=g, precision=warp.MIXED)
    # Intensity errors follow from the <= 1e-2 voxel displacement bound:
    # one level away from the faces, bounds-test flips counted at them.
    check_fast_form(outm, ref, dense_ref, size)
    g16.release()


def test_bricks_reassemble_the_full_output(warper, oracle):
    """Output bricks at offsets, each with its own grid, equal the full warp."""
    import torch
    size = (96, 80, 48)
    g = 4
    vol, sub, tar = case(size, np.uint16, 32, 6.0, seed=41)
    c, d = oracle.stps_fit(tar, sub)
    full_grid = warper.new_grid(tuple(warp.grid_dim(s, g) for s in size), g, device=False)
    warper.grid_eval(tar, c, d, full_grid, precision=warp.EXACT)
    full = warper.resample_tps(vol, full_grid, precision=warp.PLAIN)
    src = torch.from_numpy(vol.view(np.int16)).cuda()
    for (x0, y0, z0), (nx, ny, nz) in [((0, 0, 0), (40, 33, 20)), ((40, 32, 20), (56, 48, 28)),
                                       ((52, 8, 4), (30, 21, 13))]:
        gb = warper.new_grid((warp.grid_dim(nx, g), warp.grid_dim(ny, g), warp.grid_dim(nz, g)), g,
                             origin=(x0, y0, z0), device=True)
        warper.grid_eval(tar, c, d, gb, precision=warp.EXACT)
        dst = torch.zeros((nz, ny, nx), dtype=torch.int16, device="cuda")
        warper.resample_tps(src, gb, dst=dst, precision=warp.PLAIN, dst_origin=(x0, y0, z0))
        torch.cuda.synchronize()
        brick = dst.cpu().numpy().view(np.uint16)
        # Same nodes at the same global positions only when the brick origin
        # is grid-aligned; otherwise the B-spline phase differs.
        if x0 % g == 0 and y0 % g == 0 and z0 % g == 0:
            assert np.array_equal(brick, full[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("interp", [warp.LINEAR, warp.NN])
def test_zero_field_is_identity(warper, dtype, interp):
    """A zero displacement samples every voxel at itself (all weights 0), so
    the output equals the input exactly, the volume's last voxel included.
    Odd extents put that voxel in a dword that runs past the channel's end."""
    rng = np.random.default_rng(7)
    g = 4
    for size in [(67, 31, 23), (129, 17, 13), (64, 9, 6)]:
        vol = rng.integers(1, np.iinfo(dtype).max + 1, size[::-1]).astype(dtype)
        grid = warper.grid_from_array(np.zeros((3,) + tuple(warp.grid_dim(s, g) for s in size[::-1]), np.float32), g)
        for precision in (warp.PLAIN, warp.MIXED):
            out = warper.resample_tps(vol, grid, interp=interp, precision=precision)
            assert np.array_equal(out, vol), (size, precision, np.argwhere(out != vol)[:4])


@pytest.mark.parametrize("precision", [warp.PLAIN, warp.MIXED])
def test_source_window_equals_full_source(warper, oracle, precision):
    """Output boxes resampled from a source window that holds every tap
    equal the same boxes resampled from the whole volume (odd extents; one
    box reaches the volume's far corner)."""
    size = (97, 61, 37)
    g = 4
    vol, sub, tar = case(size, np.uint16, 24, 3.0, seed=43)
    c, d = oracle.stps_fit(tar, sub)
    grid = warper.new_grid(tuple(warp.grid_dim(s, g) for s in size), g, device=False)
    warper.grid_eval(tar, c, d, grid, precision=warp.EXACT)
    full = warper.resample_tps(vol, grid, precision=precision)
    m = int(np.ceil(np.abs(grid.data).max())) + 2  # field bound plus the rounding margin
    for (x0, y0, z0), (nx, ny, nz) in [((33, 21, 9), (64, 40, 28)), ((8, 4, 0), (31, 23, 17)),
                                       ((40, 30, 20), (57, 31, 17))]:
        lo = [max(0, o - m) for o in (x0, y0, z0)]
        hi = [min(s, o + n + m) for o, n, s in zip((x0, y0, z0), (nx, ny, nz), size)]
        win = np.ascontiguousarray(vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])
        dst = np.zeros((nz, ny, nx), np.uint16)
        warper.resample_tps(win, grid, dst=dst, precision=precision, src_origin=tuple(lo), dst_origin=(x0, y0, z0),
                            full=size)
        ref = full[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx]
        assert np.array_equal(dst, ref), ((x0, y0, z0), np.count_nonzero(dst != ref))


def test_linear_df_upsample_and_resample(warper, oracle):
    """TW_BASIS_LINEAR grids (the reference's declared linear DF mode): the
    FP64 reference-order form is bit-identical to the oracle's linear upsample
    and warp; the plain form is within 1e-5 voxel / 1 LSB."""
    size = (131, 77, 45)
    g = 4
    vol, sub, tar = case(size, np.uint16, 40, 5.0, seed=23)
    c, d = oracle.stps_fit(tar, sub)
    gsh = tuple(warp.grid_dim(s, g) for s in size)
    df = oracle.grid_eval(tar, c, d, gsh, g)
    grid = warper.grid_from_array(df, g, basis=warp.BASIS_LINEAR)
    ref_df = oracle.dense_df(size, df, g, kind=1)
    ex = warper.upsample_df(grid, size, precision=warp.EXACT)
    assert np.array_equal(ex, ref_df), np.abs(ex - ref_df).max()
    pl = warper.upsample_df(grid, size, precision=warp.PLAIN)
    assert np.abs(pl - ref_df).max() < 1e-5
    for interp in (warp.LINEAR, warp.NN):
        ref = oracle.tps_resample(vol, df, g, interp=interp | oracle.DF_LINEAR)
        got = warper.resample_tps(vol, grid, interp=interp, precision=warp.EXACT)
        assert np.array_equal(got, ref), (interp, np.count_nonzero(got != ref))
    ref = oracle.tps_resample(vol, df, g, interp=oracle.DF_LINEAR)
    got = warper.resample_tps(vol, grid, precision=warp.PLAIN)
    diff = np.abs(got.astype(np.int64) - ref.astype(np.int64))
    assert diff.max() <= 1 and np.mean(diff == 0) >= 0.999
    # The B-spline reading of the same grid is a different field.
    spl = warper.resample_tps(vol, warper.grid_from_array(df, g), precision=warp.EXACT)
    assert not np.array_equal(spl, ref)


def test_linear_df_chain_vs_oracle(warper, oracle):
    """One-call warp with interp | TW_DF_LINEAR against the oracle chain."""
    size, g = (300, 260, 70), 8
    vol, sub, tar = case(size, np.uint16, 128, 20.0, seed=31, n_blobs=120)
    ref = oracle.warp_tps(vol, tar, sub, g=g, interp=oracle.DF_LINEAR)[0]
    out = warper.warp_tps(vol, tar, sub, f=g, interp=warp.LINEAR | warp.DF_LINEAR, precision=warp.PLAIN)
    diff = np.abs(out.astype(np.int64) - ref.astype(np.int64))
    assert np.mean(diff <= 1) >= 0.9999 and np.mean(diff == 0) >= 0.999, (diff.max(), np.mean(diff == 0))
    outm = warper.warp_tps(vol, tar, sub, f=g, interp=warp.LINEAR | warp.DF_LINEAR, precision=warp.MIXED)
    _, df, _ = oracle.warp_tps(vol, tar, sub, g=g)
    check_fast_form(outm, ref, oracle.dense_df(size, df, g, kind=1), size)
    with pytest.raises(warp.TeraWarpError):
        warper.warp_tps(vol, tar, sub, f=g, interp=2 | warp.DF_LINEAR)


def test_resample_rejects_uncovered_grid(warper):
    vol = np.zeros((20, 20, 20), np.uint8)
    grid = warper.grid_from_array(np.zeros((3, 5, 5, 5), np.float32), 4)
    with pytest.raises(warp.TeraWarpError):
        warper.resample_tps(vol, grid)


def test_plain_config_size_sampled(warper, oracle):
    """1024x1024x256 uint16, K=512, g=8: GPU chain vs oracle on sampled block slabs."""
    import torch
    size = (1024, 1024, 256)
    g = 8
    vol = synth.blob_volume(size[::-1], np.uint16, n_blobs=200, amp=(100, 4000), seed=2)
    sub, tar = synth.landmarks(size, 512, 20.0, seed=2)
    c, d = oracle.stps_fit(tar, sub)
    gsh = tuple(warp.grid_dim(s, g) for s in size)
    df = oracle.grid_eval(tar, c, d, gsh, g)
    grid = warper.grid_from_array(df, g)
    src = torch.from_numpy(vol.view(np.int16)).cuda()
    dst = torch.empty_like(src)
    dgrid = warper.grid_from_array(torch.from_numpy(df).cuda(), g)
    warper.resample_tps(src, dgrid, dst=dst, precision=warp.PLAIN)
    torch.cuda.synchronize()
    out = dst.cpu().numpy().view(np.uint16)
    for zb in (0, 13, 31):
        ref = oracle.tps_resample(vol, df, g, zblocks=(zb, zb + 1))
        sl = slice(zb * g, (zb + 1) * g)
        diff = np.abs(out[sl].astype(np.int64) - ref[sl].astype(np.int64))
        assert diff.max() <= 1 and np.mean(diff == 0) >= 0.999, (zb, diff.max(), np.mean(diff == 0))
    del grid


def test_plain_source_over_2gib_in_slabs(warper, oracle):
    """A uint16 source past 2 GiB (1280x1280x700) runs in z-slabs whose
    source windows come from the grid's per-layer displacement bounds; the
    slices on both sides of each slab boundary match the oracle."""
    import torch
    size = (1280, 1280, 700)
    g = 8
    tile = synth.blob_volume((100, size[1], size[0]), np.uint16, n_blobs=60, amp=(100, 4000), seed=5)
    vol = np.ascontiguousarray(np.tile(tile, (7, 1, 1)))
    assert vol.nbytes > 2 ** 31
    sub, tar = synth.landmarks(size, 256, 20.0, seed=5)
    c, d = oracle.stps_fit(tar, sub)
    gsh = tuple(warp.grid_dim(s, g) for s in size)
    df = oracle.grid_eval(tar, c, d, gsh, g)
    src = torch.from_numpy(vol.view(np.int16)).cuda()
    dst = torch.empty_like(src)
    dgrid = warper.grid_from_array(torch.from_numpy(df).cuda(), g)
    warper.resample_tps(src, dgrid, dst=dst, precision=warp.PLAIN)
    torch.cuda.synchronize()
    out = dst.cpu().numpy().view(np.uint16)
    # The whole mixed chain (FP16 grid, whose layer bounds come from the
    # block tables) on the same volume.
    warper.warp_tps(src, tar, sub, f=g, precision=warp.MIXED, dst=dst)
    torch.cuda.synchronize()
    out_m = dst.cpu().numpy().view(np.uint16)
    del src, dst
    for zb in (0, 31, 32, 63, 64, 87):
        ref = oracle.tps_resample(vol, df, g, zblocks=(zb, zb + 1))
        sl = slice(zb * g, min(size[2], (zb + 1) * g))
        for o in (out, out_m):
            diff = np.abs(o[sl].astype(np.int64) - ref[sl].astype(np.int64))
            assert diff.max() <= 1 and np.mean(diff == 0) >= 0.999, (zb, diff.max(), np.mean(diff == 0))
