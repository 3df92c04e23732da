"""Host side of the streaming driver (no GPU): the procedural volume is a
function of the global coordinate (every window equals the same part of the
whole volume), and tw_checksum64 is the stated position-weighted sum."""
import numpy as np
import pytest

from terawarp_amd import Synth, checksum64


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_synth_windows_match_the_whole_volume(dtype):
    s = Synth((70, 45, 38), dtype, seed=9, cell=16, blobs_per_cell=3, sigma=(1.5, 4.0))
    whole = s.fill(0, 0, 0, 70, 45, 38)
    assert whole.std() > 1.0  # blobs and noise, not a constant
    for (x0, y0, z0, nx, ny, nz) in [(0, 0, 0, 1, 1, 1), (13, 7, 5, 40, 30, 20), (69, 44, 37, 1, 1, 1),
                                     (30, 0, 17, 40, 45, 21)]:
        assert np.array_equal(s.fill(x0, y0, z0, nx, ny, nz), whole[z0:z0 + nz, y0:y0 + ny, x0:x0 + nx])
    other = Synth((70, 45, 38), dtype, seed=10, cell=16, blobs_per_cell=3, sigma=(1.5, 4.0))
    assert not np.array_equal(other.fill(0, 0, 0, 70, 45, 38), whole)


def test_synth_rejects_boxes_outside():
    from terawarp_amd import TeraWarpError
    s = Synth((10, 10, 10), np.uint8)
    with pytest.raises(TeraWarpError):
        s.fill(5, 0, 0, 6, 1, 1)


def _checksum_np(a):
    b = np.ascontiguousarray(a).view(np.uint8).ravel()
    n = b.size
    pad = np.zeros((-n) % 8, np.uint8)
    w = np.concatenate([b, pad]).view("<u8")
    i = np.arange(w.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(w * (2 * i + 1), dtype=np.uint64)) ^ n


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 1000, 4096 + 3])
def test_checksum64_is_the_weighted_word_sum(n):
    a = np.random.default_rng(n).integers(0, 256, n).astype(np.uint8)
    assert checksum64(a) == _checksum_np(a)
    if n >= 16:  # position counts: swapping two words changes it
        b = a.copy()
        b[:8], b[8:16] = a[8:16], a[:8]
        if not np.array_equal(a[:8], a[8:16]):
            assert checksum64(b) != checksum64(a)
