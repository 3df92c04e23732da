"""The autotuner's persistence (no GPU): stored settings merge the package's
tuning file and the user's file over the library defaults; the evidence
renders as a tuned-vs-fixed table."""
import json

from terawarp_amd import tune


def test_user_file_overrides_defaults(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"gfx_test": {"settings": {"tz": 32, "brick": 256}}}))
    monkeypatch.setenv("TERAWARP_TUNE_FILE", str(path))
    s = tune.load_settings("gfx_test")
    assert s == {"tile": 0, "tz": 32, "tz_batch": 0, "brick": 256}
    assert tune.load_settings("gfx_other") == tune.DEFAULTS


def test_unreadable_user_file_keeps_defaults(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    monkeypatch.setenv("TERAWARP_TUNE_FILE", str(path))
    assert tune.load_settings("gfx_test") == tune.DEFAULTS


def test_report_marks_chosen_and_fixed():
    res = {"settings": {"tile": 1, "tz": 32, "brick": 256},
           "evidence": {"bricks": {"resample_ms": {"auto/0": 10.0, "gather/32": 8.0, "lds/0": 12.0}, "fixed": "auto/0",
                                   "chosen": "gather/32"},
                        "stream": {"wall_ms": {"256": 90.0, "512": 100.0}, "fixed": "512", "size_xyz": [64, 64, 64]}}}
    md = tune.report_md("gfx_test", res)
    assert "| bricks | gather/32 (chosen) | 8.000 | 0.800 |" in md
    assert "| bricks | auto/0 (fixed) | 10.000 | 1.000 |" in md
    assert "| stream 64x64x64 | brick 256 (chosen) | 90.0 | 0.900 |" in md
