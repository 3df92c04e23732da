"""Parity bars for the fast (FP32 / FP16) TPS forms against the FP64 oracle.

A fast form's displacement differs from the oracle's by up to DF_TOL voxel.
Away from the volume's faces that moves a trilinear sample by at most one
intensity level.  At a face it can move the sample across the strict bounds
test (position < 0 or > n - 1 gives 0), so the voxel flips between 0 and a
full intensity: those voxels are classified by the oracle's own sample
position and counted separately instead of hiding in a loose percentage.
"""
import numpy as np

DF_TOL = 1e-2  # voxel: the displacement error bound of every fast form


def face_band(dense, size_xyz, origin=(0, 0, 0), eps=DF_TOL):
    """Voxels whose oracle sample position X + d lies within eps of 0 or of
    n - 1 on some axis of a volume of extent size_xyz.  dense: (3, z, y, x)
    per-voxel displacement of a box whose first voxel is at `origin`."""
    band = np.zeros(dense.shape[1:], bool)
    for a, shape in ((0, (1, 1, -1)), (1, (1, -1, 1)), (2, (-1, 1, 1))):
        n = dense.shape[3 - a]
        p = (np.arange(n, dtype=np.float32) + origin[a]).reshape(shape) + dense[a]
        band |= (np.abs(p) <= eps) | (np.abs(p - (size_xyz[a] - 1)) <= eps)
    return band


def check_fast_form(got, ref, dense, size_xyz, min_le1=0.9999, origin=(0, 0, 0)):
    """Every voxel outside the face band within one level of the oracle, and
    at least min_le1 of all voxels (face flips included) within one level.
    Returns a dict of the counts."""
    diff = np.abs(got.astype(np.int64) - ref.astype(np.int64))
    band = face_band(dense, size_xyz, origin)
    off = diff[~band]
    stats = dict(max_off_band=int(off.max()) if off.size else 0, band=int(band.sum()),
                 band_flips=int(np.count_nonzero(diff[band] > 1)), le1=float(np.mean(diff <= 1)),
                 exact=float(np.mean(diff == 0)))
    assert stats["max_off_band"] <= 1, stats
    assert stats["le1"] >= min_le1, stats
    return stats
