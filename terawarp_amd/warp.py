"""Python view of the C ABI (include/terawarp.h).

Volumes are numpy arrays (host memory: the library stages them through the
GPU) or torch tensors on a ROCm device (device memory: no copies), shaped
(c, z, y, x) or (z, y, x), dtype uint8 or uint16 (torch.int16 is accepted as
uint16 storage).  Points are (K, 3) float32 xyz.

Mirrors the reference's call structure: `tps_fit` + `grid_eval` +
`resample_tps` are the three stages of imgwarp_smallmemory_image's GPU
branch (q_imgwarp_tps_quicksmallmemory.cpp:2066, 2143), `resample_affine` is
q_imagewarp_affine_image's (q_warp_affine_tps.cpp:1453).
"""
import ctypes
import math

import numpy as np

from . import _lib as L

PLAIN, MIXED, EXACT = L.TW_PLAIN, L.TW_MIXED, L.TW_EXACT
LINEAR, NN = L.TW_INTERP_LINEAR, L.TW_INTERP_NN
F32, F16 = L.TW_DF_F32, L.TW_DF_F16
# Grid upsampling basis (Grid.basis); DF_LINEAR is OR-ed into the interp
# argument of the one-call warps to select BASIS_LINEAR for their grids.
BASIS_BSPLINE, BASIS_LINEAR = L.TW_BASIS_BSPLINE, L.TW_BASIS_LINEAR
DF_LINEAR = L.TW_DF_LINEAR


class TeraWarpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"terawarp error {code}: {msg}")
        self.code = code


def _torch():
    import torch
    return torch


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _pts(p):
    p = np.ascontiguousarray(p, dtype=np.float32)
    if p.ndim != 2 or p.shape[1] != 3:
        raise ValueError("points must be (K, 3)")
    return p


def _vol(a, origin=(0, 0, 0), full=None):
    """TwVolume for a numpy array or a torch device tensor."""
    if _is_torch(a):
        torch = _torch()
        if not a.is_cuda:
            raise ValueError("torch volumes must live on the GPU; pass numpy arrays for host memory")
        if not a.is_contiguous():
            raise ValueError("volume tensor must be contiguous")
        if a.dtype == torch.uint8:
            dt = L.TW_U8
        elif a.dtype in (torch.int16, getattr(torch, "uint16", torch.int16)):
            dt = L.TW_U16
        else:
            raise ValueError(f"unsupported volume dtype {a.dtype}")
        ptr, mem, shape = a.data_ptr(), L.TW_MEM_DEVICE, tuple(a.shape)
    else:
        if not (isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"]):
            raise ValueError("numpy volume must be C-contiguous")
        if a.dtype == np.uint8:
            dt = L.TW_U8
        elif a.dtype == np.uint16:
            dt = L.TW_U16
        else:
            raise ValueError(f"unsupported volume dtype {a.dtype}")
        ptr, mem, shape = a.ctypes.data, L.TW_MEM_HOST, a.shape
    if len(shape) == 3:
        shape = (1,) + shape
    if len(shape) != 4:
        raise ValueError("volume must be (z, y, x) or (c, z, y, x)")
    nc, nz, ny, nx = shape
    v = L.TwVolume()
    v.data = ptr
    v.nx, v.ny, v.nz, v.nc = nx, ny, nz, nc
    v.x0, v.y0, v.z0 = origin
    full = full or (0, 0, 0)
    v.full_nx, v.full_ny, v.full_nz = full
    v.dtype, v.mem = dt, mem
    return v


def grid_dim(n, f):
    """ceil(n / f) + 3 nodes (q_imgwarp_tps_quicksmallmemory.cpp:592)."""
    return int(math.ceil(n / f)) + 3


def affine_fit(pts_a, pts_b):
    """4x4 M with pts_b ~= M pts_a (DLT with Hartley normalisation)."""
    a, b = _pts(pts_a), _pts(pts_b)
    M = np.zeros((4, 4), np.float64)
    rc = L.load().tw_affine_fit(_fptr(a), _fptr(b), a.shape[0], _dptr(M))
    if rc != L.TW_OK:
        raise TeraWarpError(rc, "affine fit failed (fewer than 4 points or degenerate)")
    return M


ROTATION, TRANSLATION, SCALING = 0, 1, 2


def mode_matrix(mode, para, size_xyz):
    """Target->subject matrix of the README's rotation/translation/scaling modes."""
    p = np.ascontiguousarray(para, np.float64)
    M = np.zeros((4, 4), np.float64)
    rc = L.load().tw_mode_matrix(int(mode), _dptr(p), int(size_xyz[0]), int(size_xyz[1]), int(size_xyz[2]),
                                 _dptr(M))
    if rc != L.TW_OK:
        raise TeraWarpError(rc, f"invalid mode {mode} / parameters {list(para)}")
    return M


class Grid:
    """A displacement grid (3 planes) in host (numpy) or device (torch) memory."""

    def __init__(self, warper, data, gshape_xyz, f, origin, storage, basis=BASIS_BSPLINE):
        self._w = warper
        self.data = data
        self.g = L.TwGrid()
        gx, gy, gz = gshape_xyz
        self.g.gx, self.g.gy, self.g.gz = gx, gy, gz
        self.g.f = f
        self.g.ox, self.g.oy, self.g.oz = origin
        self.g.storage = storage
        self.g.basis = basis
        if _is_torch(data):
            self.g.data = data.data_ptr()
            self.g.mem = L.TW_MEM_DEVICE
        else:
            self.g.data = data.ctypes.data
            self.g.mem = L.TW_MEM_HOST
        self.g.qtab = None

    def release(self):
        if self.g.qtab:
            L.load().tw_grid_release(self._w.ctx, ctypes.byref(self.g))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Warper:
    """One library context (device, HIP stream, cached workspaces)."""

    def __init__(self, device=0, timing=False, tuned=True):
        """tuned: apply the device's stored tunables (terawarp_amd.tune);
        False keeps the library defaults."""
        self.lib = L.load()
        ctx = ctypes.c_void_p()
        rc = self.lib.tw_create(device, ctypes.byref(ctx))
        if rc != L.TW_OK:
            raise TeraWarpError(rc, "tw_create failed (no usable HIP device?)")
        self.ctx = ctx
        self._ext = None
        self.stream_brick = 512
        if timing:
            self.lib.tw_set_timing(self.ctx, 1)
        if tuned:
            from . import tune
            tune.apply(self, tune.load_settings(tune.device_arch(device)))

    def _ordered(self, fn, *args):
        """Call the library with its stream ordered after torch's current
        stream (so torch's writes to the inputs and outputs land first) and
        torch's current stream ordered after the call's work: event waits on
        the device, no host synchronisation."""
        torch = _torch()
        if not torch.cuda.is_available():
            return fn(*args)
        if self._ext is None:
            self._ext = torch.cuda.ExternalStream(self.lib.tw_get_stream(self.ctx))
        cur = torch.cuda.current_stream()
        self._ext.wait_stream(cur)
        rc = fn(*args)
        cur.wait_stream(self._ext)
        return rc

    def close(self):
        if self.ctx:
            self.lib.tw_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != L.TW_OK:
            raise TeraWarpError(rc, self.lib.tw_last_error(self.ctx).decode(errors="replace"))

    def set_stream(self, stream):
        """Launch on a torch.cuda.Stream (or a raw hipStream_t int); None = own stream."""
        ptr = None if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))
        self._check(self.lib.tw_set_stream(self.ctx, ptr))
        self._ext = None

    def set_timing(self, on=True):
        self._check(self.lib.tw_set_timing(self.ctx, 1 if on else 0))

    TUNABLES = {"tile": L.TW_TUNE_TILE, "tz": L.TW_TUNE_TZ, "tz_batch": L.TW_TUNE_TZ_BATCH,
                "fp32_factor_min": L.TW_TUNE_FP32_FACTOR_MIN, "solve_wait_us": L.TW_TUNE_SOLVE_WAIT_US,
                "solve_pipelined": L.TW_TUNE_SOLVE_PIPELINED}

    def set_tune(self, key, value):
        """Set a tunable (tw_set_tune): "tile" (0 auto, 1 gather, 2 lds), "tz"
        / "tz_batch" (gather tile depth of single-volume / persistent
        launches, 0 = from the grid spacing), "fp32_factor_min",
        "solve_wait_us" (bound of the pipelined solve's waits; setting it
        turns that solve back on after a wait ran out); "solve_pipelined" is
        read only (0 once a wait ran out)."""
        self._check(self.lib.tw_set_tune(self.ctx, self.TUNABLES[key], int(value)))

    def get_tune(self, key):
        v = ctypes.c_int64()
        self._check(self.lib.tw_get_tune(self.ctx, self.TUNABLES[key], ctypes.byref(v)))
        return v.value

    def set_tile(self, tile):
        """Fused TPS resample tile: "auto" (the gather tile), "gather" or
        "lds" (source streamed through LDS); outputs are identical."""
        code = {"auto": L.TW_TILE_AUTO, "gather": L.TW_TILE_GATHER, "lds": L.TW_TILE_LDS}[tile]
        self._check(self.lib.tw_set_tile(self.ctx, code))

    def sync(self):
        self._check(self.lib.tw_sync(self.ctx))

    def timings(self):
        t = L.TwTimings()
        self._check(self.lib.tw_get_timings(self.ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in L.TwTimings._fields_ if k != "reserved"}

    # ---- affine --------------------------------------------------------
    def resample_affine(self, src, M, dst=None, out_size_xyz=None, interp=LINEAR, src_origin=(0, 0, 0),
                        dst_origin=(0, 0, 0), full=None):
        M = np.ascontiguousarray(M, np.float64).reshape(-1)[:12].copy()
        if dst is None:
            if _is_torch(src):
                raise ValueError("pass dst for device volumes")
            s4 = src if src.ndim == 4 else src[None]
            nx, ny, nz = out_size_xyz or (s4.shape[3], s4.shape[2], s4.shape[1])
            dst = np.zeros((s4.shape[0], nz, ny, nx), src.dtype)
            if src.ndim == 3:
                dst = dst[0]
        vs, vd = _vol(src, src_origin, full), _vol(dst, dst_origin)
        self._check(self._ordered(self.lib.tw_resample_affine, self.ctx, ctypes.byref(vs), ctypes.byref(vd), _dptr(M), interp))
        return dst

    # ---- TPS -------------------------------------------------------------
    def tps_fit(self, centres, targets, lam=0.2, precision=PLAIN):
        ce, ta = _pts(centres), _pts(targets)
        if ce.shape != ta.shape:
            raise ValueError("centres and targets differ in shape")
        K = ce.shape[0]
        c = np.zeros((K, 4), np.float64)
        d = np.zeros((4, 4), np.float64)
        self._check(self.lib.tw_tps_fit(self.ctx, _fptr(ce), _fptr(ta), K, lam, precision, _dptr(c), _dptr(d)))
        return c, d

    def new_grid(self, gshape_xyz, f, origin=(0, 0, 0), storage=F32, device=True, basis=BASIS_BSPLINE):
        gx, gy, gz = gshape_xyz
        if device:
            torch = _torch()
            dt = torch.float32 if storage == F32 else torch.float16
            data = torch.empty((3, gz, gy, gx), dtype=dt, device="cuda")
        else:
            if storage != F32:
                raise ValueError("host grids are FP32")
            data = np.zeros((3, gz, gy, gx), np.float32)
        return Grid(self, data, gshape_xyz, f, origin, storage, basis)

    def grid_eval(self, centres, c, d, grid, precision=PLAIN):
        ce = _pts(centres)
        c = np.ascontiguousarray(c, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        self._check(self._ordered(self.lib.tw_tps_grid_eval, self.ctx, _fptr(ce), ce.shape[0], _dptr(c), _dptr(d), precision,
                                              ctypes.byref(grid.g)))
        return grid

    def grid_from_array(self, df, f, origin=(0, 0, 0), basis=BASIS_BSPLINE):
        """Wrap an existing (3, gz, gy, gx) float32 numpy array or device tensor."""
        shape = tuple(df.shape)
        return Grid(self, df, (shape[3], shape[2], shape[1]), f, origin, F32, basis)

    def resample_tps(self, src, grid, dst=None, interp=LINEAR, precision=PLAIN, src_origin=(0, 0, 0),
                     dst_origin=(0, 0, 0), full=None):
        if dst is None:
            if _is_torch(src):
                raise ValueError("pass dst for device volumes")
            dst = np.zeros_like(src)
        vs, vd = _vol(src, src_origin, full), _vol(dst, dst_origin)
        self._check(self._ordered(self.lib.tw_resample_tps, self.ctx, ctypes.byref(vs), ctypes.byref(vd), ctypes.byref(grid.g),
                                             interp, precision))
        return dst

    @staticmethod
    def prepare_bricks(srcs, dsts, src_origins, dst_origins, full):
        """The descriptor arrays of resample_bricks, built once for repeated
        calls on the same buffers."""
        n = len(srcs)
        if len(dsts) != n or len(src_origins) != n or len(dst_origins) != n:
            raise ValueError("srcs, dsts and origins differ in length")
        vs = (L.TwVolume * n)(*[_vol(s, o, full) for s, o in zip(srcs, src_origins)])
        vd = (L.TwVolume * n)(*[_vol(d, o) for d, o in zip(dsts, dst_origins)])
        return vs, vd, n

    def resample_bricks(self, srcs, dsts, grid, src_origins, dst_origins, full, interp=LINEAR, precision=MIXED,
                        prepared=None):
        """tw_resample_tps_bricks: device source windows srcs[i] (global origin
        src_origins[i], whole-volume extent full) resampled into the device
        output boxes dsts[i] (global origin dst_origins[i]) against one device
        grid, in one launch.  Asynchronous on the context's stream."""
        vs, vd, n = prepared or self.prepare_bricks(srcs, dsts, src_origins, dst_origins, full)
        self._check(self._ordered(self.lib.tw_resample_tps_bricks, self.ctx, vs, vd, n, ctypes.byref(grid.g),
                                  interp, precision))
        return dsts

    def upsample_df(self, grid, size_xyz, origin=(0, 0, 0), precision=PLAIN):
        nx, ny, nz = size_xyz
        out = np.zeros((3, nz, ny, nx), np.float32)
        self._check(self._ordered(self.lib.tw_upsample_df, self.ctx, ctypes.byref(grid.g), nx, ny, nz, origin[0], origin[1],
                                            origin[2], precision, _fptr(out), L.TW_MEM_HOST))
        return out

    def warp_tps(self, src, centres, targets, f=4, lam=0.2, interp=LINEAR, precision=PLAIN, dst=None):
        ce, ta = _pts(centres), _pts(targets)
        if dst is None:
            if _is_torch(src):
                raise ValueError("pass dst for device volumes")
            dst = np.zeros_like(src)
        vs, vd = _vol(src), _vol(dst)
        self._check(self._ordered(self.lib.tw_warp_tps, self.ctx, ctypes.byref(vs), ctypes.byref(vd), _fptr(ce), _fptr(ta),
                                         ce.shape[0], f, lam, interp, precision))
        return dst

    # ---- batches ---------------------------------------------------------
    def warp_batch(self, items, f=8, lam=0.2, interp=LINEAR, precision=MIXED):
        """Warp many device volumes in one launch per stage (tw_warp_batch).

        items: dicts with "src" and "dst" (device tensors) and either
        "centres" + "targets" ((K, 3) float32; a TPS warp) or "M" (a 3x4 or
        4x4 target->subject matrix; an affine or rigid warp), optionally
        "dst_origin".  Returns (status, done_ms): per-item status codes and the
        device-clock times at which the items' outputs were complete.
        Raises on the first failed item after the batch ran.
        """
        n = len(items)
        arr = (L.TwWarpItem * n)()
        keep = []
        for i, it in enumerate(items):
            w = arr[i]
            w.src = _vol(it["src"])
            w.dst = _vol(it["dst"], it.get("dst_origin", (0, 0, 0)))
            if "M" in it:
                w.kind = L.TW_WARP_AFFINE
                M = np.ascontiguousarray(it["M"], np.float64).reshape(-1)[:12]
                for k in range(12):
                    w.M[k] = float(M[k])
            else:
                ce, ta = _pts(it["centres"]), _pts(it["targets"])
                if ce.shape != ta.shape:
                    raise ValueError(f"item {i}: centres and targets differ in shape")
                keep += [ce, ta]
                w.kind = L.TW_WARP_TPS
                w.K = ce.shape[0]
                w.centres = _fptr(ce)
                w.targets = _fptr(ta)
        status = (ctypes.c_int32 * n)()
        done = np.zeros(n, np.float64)
        rc = self._ordered(self.lib.tw_warp_batch, self.ctx, arr, n, f, lam, interp, precision, status,
                           _dptr(done))
        self._check(rc)
        return [int(x) for x in status], done

    # ---- streaming ----------------------------------------------------------
    def warp_stream(self, size_xyz, dtype, centres, targets, source, f=8, brick=None, lam=0.2, interp=LINEAR,
                    precision=MIXED, threads=16, keep=(), block_base=None, block_origin=False):
        """Warp a volume brick by brick (tw_warp_stream).

        source: a Synth (procedural volume, filled by the library's own
        threads) or a Python callable read(x0, y0, z0, nx, ny, nz) returning
        a (nz, ny, nx) array.  Returns (checksums, kept, stats): per-brick
        tw_checksum64 values, {brick id: (z, y, x) output array} for the
        bricks listed in keep, and the call's statistics.  With block_base,
        the bricks are written as the tera driver's block files instead
        (tw_block_file_write; checksums and kept bricks are then empty).
        block_origin: TW_STREAM_BLOCK_ORIGIN (EXACT positions in the tera
        block warp's block-local order, the bricks being its blocks).
        """
        brick = brick or self.stream_brick  # the tuned brick edge unless given
        ce, ta = _pts(centres), _pts(targets)
        nx, ny, nz = size_xyz
        dt = L.TW_U8 if np.dtype(dtype) == np.uint8 else L.TW_U16
        nb = -(-nx // brick) * -(-ny // brick) * -(-nz // brick)
        sums = np.zeros(nb, np.uint64)
        keep = [int(k) for k in keep]
        kept = {}
        ids = (ctypes.c_int64 * max(1, len(keep)))(*keep)
        ptrs = (ctypes.c_void_p * max(1, len(keep)))()
        for i, b in enumerate(keep):
            bx, by, bz = b % -(-nx // brick), (b // -(-nx // brick)) % -(-ny // brick), \
                b // (-(-nx // brick) * -(-ny // brick))
            shape = (min(brick, nz - bz * brick), min(brick, ny - by * brick), min(brick, nx - bx * brick))
            kept[b] = np.zeros(shape, dtype)
            ptrs[i] = kept[b].ctypes.data
        sink = L.TwChecksumSink(sums.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), nb,
                                np.dtype(dtype).itemsize, len(keep), ids, ptrs)
        d = L.TwStreamDesc()
        d.nx, d.ny, d.nz, d.dtype, d.brick, d.f = nx, ny, nz, dt, brick, f
        d.interp, d.precision, d.threads, d.K, d.lam = interp, precision, threads, ce.shape[0], lam
        d.centres, d.targets = _fptr(ce), _fptr(ta)
        d.flags = L.TW_STREAM_BLOCK_ORIGIN if block_origin else 0
        from .host import SliceDir
        if isinstance(source, Synth):
            d.read = ctypes.cast(self.lib.tw_synth_read, ctypes.c_void_p)
            d.read_user = ctypes.cast(ctypes.byref(source.s), ctypes.c_void_p)
            cb = None
        elif isinstance(source, SliceDir):
            d.read = ctypes.cast(self.lib.tw_slice_dir_read, ctypes.c_void_p)
            d.read_user = source.h
            cb = None
        else:
            def _read(user, x0, y0, z0, bnx, bny, bnz, out):
                try:
                    a = np.ascontiguousarray(source(x0, y0, z0, bnx, bny, bnz), dtype)
                    ctypes.memmove(out, a.ctypes.data, a.nbytes)
                    return L.TW_OK
                except Exception:
                    return L.TW_E_IO
            cb = L.READ_FN(_read)
            d.read = ctypes.cast(cb, ctypes.c_void_p)
        if block_base is not None:
            files = L.TwBlockFiles(str(block_base).encode(), dt, 0)
            d.write = ctypes.cast(self.lib.tw_block_file_write, ctypes.c_void_p)
            d.write_user = ctypes.cast(ctypes.byref(files), ctypes.c_void_p)
        else:
            d.write = ctypes.cast(self.lib.tw_checksum_write, ctypes.c_void_p)
            d.write_user = ctypes.cast(ctypes.byref(sink), ctypes.c_void_p)
        st = L.TwStreamStats()
        self._check(self._ordered(self.lib.tw_warp_stream, self.ctx, ctypes.byref(d), ctypes.byref(st)))
        return sums, kept, {k: getattr(st, k) for k, _ in L.TwStreamStats._fields_}


class Synth:
    """Procedural subject volume (tw_synth): noise hashed from (seed, x, y, z)
    plus Gaussian blobs per cell; any window of it is reproducible."""

    def __init__(self, size_xyz, dtype=np.uint16, seed=1, cell=256, blobs_per_cell=16, sigma=(4.0, 16.0),
                 amp=None, noise=None):
        self.dtype = np.dtype(dtype)
        u8 = self.dtype == np.uint8
        amp = amp or ((20.0, 200.0) if u8 else (100.0, 4000.0))
        self.s = L.TwSynth(seed, size_xyz[0], size_xyz[1], size_xyz[2], L.TW_U8 if u8 else L.TW_U16, cell,
                           blobs_per_cell, 0, sigma[0], sigma[1], amp[0], amp[1],
                           noise if noise is not None else (5.0 if u8 else 40.0))
        self.size = tuple(size_xyz)

    def fill(self, x0, y0, z0, nx, ny, nz):
        out = np.zeros((nz, ny, nx), self.dtype)
        rc = L.load().tw_synth_fill(ctypes.byref(self.s), x0, y0, z0, nx, ny, nz, out.ctypes.data)
        if rc != L.TW_OK:
            raise TeraWarpError(rc, "tw_synth_fill: box outside the volume")
        return out

    def fill_into(self, out, x0, y0, z0, nx, ny, nz, threads=16):
        """Fill the box into `out` (a C-contiguous (nz, ny, nx) array of this
        dtype, e.g. a pinned buffer's view) with `threads` host threads over
        z-ranges (the library call releases the interpreter lock)."""
        from concurrent.futures import ThreadPoolExecutor
        if out.shape != (nz, ny, nx) or out.dtype != self.dtype or not out.flags["C_CONTIGUOUS"]:
            raise ValueError("out must be a C-contiguous (nz, ny, nx) array of the volume's dtype")
        lib = L.load()
        plane = nx * ny * self.dtype.itemsize
        nt = max(1, min(threads, nz))

        def part(t):
            za, zb = nz * t // nt, nz * (t + 1) // nt
            if zb > za:
                return lib.tw_synth_fill(ctypes.byref(self.s), x0, y0, z0 + za, nx, ny, zb - za,
                                         out.ctypes.data + za * plane)
            return L.TW_OK

        with ThreadPoolExecutor(nt) as ex:
            rcs = list(ex.map(part, range(nt)))
        if any(rc != L.TW_OK for rc in rcs):
            raise TeraWarpError(min(rcs), "tw_synth_fill: box outside the volume")
        return out


def checksum64(a):
    """tw_checksum64 of an array's bytes."""
    a = np.ascontiguousarray(a)
    return int(L.load().tw_checksum64(a.ctypes.data, a.nbytes))


def stitch_slices(base, size_xyz, brick, dtype):
    """Write <base>_slice/ZZZZZZZ.raw from the block files <base>_block/ (tw_stitch_slices)."""
    nx, ny, nz = size_xyz
    dt = L.TW_U8 if np.dtype(dtype) == np.uint8 else L.TW_U16
    lib = L.load()
    rc = lib.tw_stitch_slices(str(base).encode(), nx, ny, nz, brick, dt)
    if rc != L.TW_OK:
        raise TeraWarpError(rc, lib.tw_tera_last_error().decode(errors="replace"))


def tera_paths(base, x0=0, y0=0, z0=0, z=None):
    """(block file path of the brick at (x0, y0, z0), slice file path of z)."""
    lib = L.load()
    b = ctypes.create_string_buffer(4096)
    lib.tw_tera_block_path(str(base).encode(), x0, y0, z0, b, 4096)
    s = ctypes.create_string_buffer(4096)
    lib.tw_tera_slice_path(str(base).encode(), z if z is not None else z0, s, 4096)
    return b.value.decode(), s.value.decode()

