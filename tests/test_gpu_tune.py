"""The autotuner on the GPU: a reduced tuning run (one brick, two tile
depths, no streaming sweep; bricks, batch and the plain-GPU chain) measures every candidate and returns settings
among them; a Warper created afterwards applies the stored settings."""
import json

import pytest

from terawarp_amd import tune
from terawarp_amd.warp import Warper

pytestmark = pytest.mark.gpu


def test_tune_reduced_run_and_apply(warper, tmp_path, monkeypatch):
    old = {k: warper.get_tune(k) for k in ("tile", "tz", "tz_batch")}
    try:
        res = tune.tune(warper, reps=1, tz_candidates=(16, 32), bricks=1, stream=False, log=lambda *a: None)
    finally:
        for k, v in old.items():
            warper.set_tune(k, v)
    for kind in ("bricks", "batch", "warp"):
        times = res["evidence"][kind]["resample_ms"]
        assert set(times) == {"auto/0", "gather/16", "gather/32", "lds/0"}
        assert all(t > 0 for t in times.values())
    s = res["settings"]
    assert s["tile"] in (1, 2) and s["tz"] in (16, 32) and s["tz_batch"] in (16, 32)
    path = tmp_path / "tune.json"
    arch = tune.device_arch()
    path.write_text(json.dumps({arch: {"settings": {"tile": 1, "tz": 24, "tz_batch": 48, "brick": 256}}}))
    monkeypatch.setenv("TERAWARP_TUNE_FILE", str(path))
    w = Warper(0)
    try:
        assert w.get_tune("tile") == 1 and w.get_tune("tz") == 24 and w.get_tune("tz_batch") == 48
        assert w.stream_brick == 256
    finally:
        w.close()
    w = Warper(0, tuned=False)
    try:
        assert w.get_tune("tz") == 0 and w.stream_brick == 512
    finally:
        w.close()
