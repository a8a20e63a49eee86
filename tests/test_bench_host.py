"""Host logic of bench.py that must hold whatever the device: the streaming
ceiling (hbm_measured) returns a record, never raises, so the bench's JSON
line is printed even where torch or the HIP runtime cannot open a device."""
import math
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_hbm_measured_always_returns_a_record():
    hm = bench.hbm_measured(1 << 20, 1)
    assert set(hm) >= {"copy_GBs", "read_GBs", "source"}
    for key in ("copy_GBs", "read_GBs"):
        v = hm[key]
        assert v is None or (isinstance(v, float) and math.isfinite(v) and v > 0.0), (key, v)
    if hm["copy_GBs"] is None:
        assert hm["read_GBs"] is None and hm["source"] == "not measured"


def test_gpus_above_one_exit_with_a_message(capsys, monkeypatch):
    monkeypatch.setattr(sys, "argv", ["bench.py", "--gpus", "2"])
    try:
        bench.main()
    except SystemExit as e:
        assert e.code == 2
    else:
        raise AssertionError("bench.py --gpus 2 did not exit")
    assert "one GPU" in capsys.readouterr().err
