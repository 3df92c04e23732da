"""Write the small golden fixtures under tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

Inputs come from the seeded generators in terawarp_amd.synth; outputs from
the oracle (oracle/tw_oracle.cpp), whose functions are pinned by
tests/test_oracle_*.py.  The reference ships no data of its own, so these
fixtures guard the oracle against regressions and give the GPU tests fixed
cases.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from terawarp_amd import synth  # noqa: E402

CASES = {
    "tps_u8": dict(size=(48, 40, 24), dtype=np.uint8, K=24, g=4, dmax=4.0, seed=11),
    "tps_u16": dict(size=(40, 32, 20), dtype=np.uint16, K=16, g=8, dmax=6.0, seed=12),
}


def make_tps(name, size, dtype, K, g, dmax, seed):
    amp = (20, 200) if dtype == np.uint8 else (100, 4000)
    vol = synth.blob_volume(size[::-1], dtype, n_blobs=12, amp=amp, sigma=(2, 5), seed=seed)
    sub, tar = synth.landmarks(size, K, dmax, seed=seed)
    out, df, (c, d) = oracle.warp_tps(vol, tar, sub, g=g)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), vol=vol, sub=sub, tar=tar, g=g, c=c, d=d, df=df, out=out)


def make_affine():
    size = (40, 36, 28)
    vol = synth.blob_volume(size[::-1], np.uint16, n_blobs=10, amp=(100, 4000), sigma=(2, 5), seed=13)
    M = synth.random_affine(size, seed=13)
    out = oracle.affine_resample(vol, M)
    np.savez_compressed(os.path.join(HERE, "affine_u16.npz"), vol=vol, M=M, out=out)


def main():
    for name, kw in CASES.items():
        make_tps(name, **kw)
    make_affine()
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
