"""Seeded synthetic workloads: blob volumes and landmark pairs.

Volumes are sums of 3-D Gaussian blobs plus noise (uniform +-5 for uint8,
Poisson-like value + N(0, sqrt(value)) for uint16), clipped to the type's
range.  Landmarks: subject-space markers, and target-space markers displaced
from them by a smooth field (a sum of low-frequency sinusoids per axis,
|d| <= dmax).  Every generator is a pure function of its seed.
"""
import math

import numpy as np

U8_LIMIT, U16_LIMIT = 255, 65535


def blob_volume(shape_zyx, dtype=np.uint8, n_blobs=200, sigma=(2.0, 8.0), amp=(20.0, 200.0), seed=0,
                noise=None, z_chunk=16):
    """(z, y, x) volume of `n_blobs` Gaussian blobs plus noise."""
    nz, ny, nx = shape_zyx
    rng = np.random.default_rng([seed, 1])
    vol = np.zeros(shape_zyx, np.float32)
    cz = rng.uniform(0, nz - 1, n_blobs)
    cy = rng.uniform(0, ny - 1, n_blobs)
    cx = rng.uniform(0, nx - 1, n_blobs)
    sg = rng.uniform(sigma[0], sigma[1], n_blobs)
    am = rng.uniform(amp[0], amp[1], n_blobs)
    for i in range(n_blobs):
        r = int(math.ceil(3 * sg[i]))
        z0, z1 = max(0, int(cz[i]) - r), min(nz, int(cz[i]) + r + 2)
        y0, y1 = max(0, int(cy[i]) - r), min(ny, int(cy[i]) + r + 2)
        x0, x1 = max(0, int(cx[i]) - r), min(nx, int(cx[i]) + r + 2)
        if z0 >= z1 or y0 >= y1 or x0 >= x1:
            continue
        zz = (np.arange(z0, z1, dtype=np.float32) - cz[i]) ** 2
        yy = (np.arange(y0, y1, dtype=np.float32) - cy[i]) ** 2
        xx = (np.arange(x0, x1, dtype=np.float32) - cx[i]) ** 2
        inv = np.float32(-0.5 / sg[i] ** 2)
        g = (np.exp(inv * zz)[:, None, None] * np.exp(inv * yy)[None, :, None] * np.exp(inv * xx)[None, None, :])
        vol[z0:z1, y0:y1, x0:x1] += np.float32(am[i]) * g
    is16 = np.dtype(dtype) == np.uint16
    if noise is None:
        noise = "poisson" if is16 else "uniform"
    limit = U16_LIMIT if is16 else U8_LIMIT
    out = np.empty(shape_zyx, dtype)
    for z0 in range(0, nz, z_chunk):
        z1 = min(nz, z0 + z_chunk)
        nr = np.random.default_rng([seed, 2, z0])
        slab = vol[z0:z1]
        if noise == "uniform":
            slab = slab + nr.uniform(-5.0, 5.0, slab.shape).astype(np.float32)
        elif noise == "poisson":
            slab = slab + nr.standard_normal(slab.shape).astype(np.float32) * np.sqrt(np.maximum(slab, 1.0))
        out[z0:z1] = np.clip(np.rint(slab), 0, limit).astype(dtype)
    return out


class SmoothField:
    """Sum of `terms` low-frequency sinusoids per axis, |d(p)| <= dmax."""

    def __init__(self, size_xyz, dmax, seed=0, terms=4, cycles=(0.5, 2.0)):
        rng = np.random.default_rng([seed, 3])
        size = np.asarray(size_xyz, np.float64)
        self.freq = np.empty((3, terms, 3))
        for a in range(3):
            for k in range(terms):
                u = rng.normal(size=3)
                u /= np.linalg.norm(u)
                self.freq[a, k] = u * rng.uniform(*cycles) / size
        self.phase = rng.uniform(0, 2 * np.pi, (3, terms))
        amp = rng.uniform(0.5, 1.0, (3, terms))
        # Each axis bounded by dmax/sqrt(3) so that |d| <= dmax.
        amp *= (dmax / math.sqrt(3)) / amp.sum(axis=1, keepdims=True)
        self.amp = amp

    def __call__(self, p):
        p = np.asarray(p, np.float64)
        d = np.zeros_like(p)
        for a in range(3):
            arg = 2 * np.pi * (p @ self.freq[a].T) + self.phase[a]
            d[:, a] = (self.amp[a] * np.sin(arg)).sum(axis=1)
        return d


def landmarks(size_xyz, K, dmax, seed=0, layout="uniform", outliers=0.1):
    """(subject, target) float32 (K, 3) marker sets; target = subject + d(subject).

    layout "uniform": subject markers uniform in the volume;
    layout "lattice": jittered cubic lattice (jitter +-1/4 cell) with a
    fraction `outliers` replaced by uniform points.
    """
    rng = np.random.default_rng([seed, 4])
    size = np.asarray(size_xyz, np.float64)
    if layout == "uniform":
        sub = rng.uniform(0, 1, (K, 3)) * (size - 1)
    elif layout == "lattice":
        n = int(round(K ** (1 / 3)))
        if n ** 3 != K:
            raise ValueError("lattice layout needs a cubic K")
        cell = size / n
        g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
        sub = (g + 0.5 + rng.uniform(-0.25, 0.25, (K, 3))) * cell
        m = rng.random(K) < outliers
        sub[m] = rng.uniform(0, 1, (int(m.sum()), 3)) * (size - 1)
    else:
        raise ValueError(layout)
    field = SmoothField(size_xyz, dmax, seed)
    tar = sub + field(sub)
    return sub.astype(np.float32), tar.astype(np.float32)


def random_affine(size_xyz, seed=0, det_range=(0.8, 1.2), max_shift=8.0):
    """Random target->subject 4x4 with det in det_range, about the volume centre."""
    rng = np.random.default_rng([seed, 5])
    while True:
        A = np.eye(3) + rng.uniform(-0.12, 0.12, (3, 3))
        if det_range[0] <= np.linalg.det(A) <= det_range[1]:
            break
    c = (np.asarray(size_xyz, np.float64) - 1) / 2
    t = c - A @ c + rng.uniform(-max_shift, max_shift, 3)
    M = np.eye(4)
    M[:3, :3] = A
    M[:3, 3] = t
    return M


def batch_instances(seed=5, n=256, dmax=8.0):
    """The uneven-instance batch: axes U[64, 512], uint8 / uint16 50/50, and
    60 % TPS (K U[16, 256], |d| <= dmax), 25 % affine (random 3x4, det in
    [0.8, 1.2]), 15 % rigid (rotation U[-180, 180] deg per axis, integer
    translation U[-32, 32]).  Each spec: size (xyz), dtype, seed, kind and
    either centres / targets (TPS) or the 4x4 target -> subject matrix M."""
    from . import warp
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        size = tuple(int(v) for v in rng.integers(64, 513, 3))
        dtype = np.uint8 if rng.random() < 0.5 else np.uint16
        u = rng.random()
        sp = {"size": size, "dtype": dtype, "seed": seed * 100000 + i}
        if u < 0.6:
            K = int(rng.integers(16, 257))
            sub, tar = landmarks(size, K, dmax, seed=sp["seed"])
            sp.update(kind="tps", centres=tar, targets=sub)
        elif u < 0.85:
            sp.update(kind="affine", M=random_affine(size, seed=sp["seed"]))
        else:
            ang = rng.uniform(-180.0, 180.0, 3)
            tr = rng.integers(-32, 33, 3).astype(np.float64)
            sp.update(kind="rigid", M=warp.mode_matrix(warp.ROTATION, ang, size) @
                      warp.mode_matrix(warp.TRANSLATION, tr, size))
        specs.append(sp)
    return specs


def blob_volume_device(size_xyz, dtype, seed, n_blobs=24):
    """Seeded blob volume generated on the GPU with torch (separable Gaussian
    blobs, sigma 2-8 % of the smallest axis, plus uniform noise): a
    (z, y, x) uint8 or int16-stored uint16 device tensor."""
    import torch
    g = np.random.default_rng(seed)
    nx, ny, nz = size_xyz
    hi = 255.0 if dtype == np.uint8 else 4095.0
    amp = (20.0, 200.0) if dtype == np.uint8 else (100.0, 4000.0)
    ax = [torch.arange(n, device="cuda", dtype=torch.float32) for n in (nz, ny, nx)]
    vol = torch.zeros((nz, ny, nx), device="cuda", dtype=torch.float32)
    smin = min(size_xyz)
    for _ in range(n_blobs):
        c = g.uniform(0, 1, 3) * np.array([nz, ny, nx])
        sg = g.uniform(0.02, 0.08) * smin
        a = g.uniform(*amp)
        e = [torch.exp(-((ax[k] - float(c[k])) ** 2) / (2 * sg * sg)) for k in range(3)]
        vol += a * e[0][:, None, None] * e[1][None, :, None] * e[2][None, None, :]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(int(seed))
    vol += (torch.rand(vol.shape, generator=gen, device="cuda") - 0.5) * 10.0
    vol.clamp_(0, hi)
    if dtype == np.uint8:
        return vol.to(torch.uint8)
    return vol.to(torch.int32).to(torch.int16)
