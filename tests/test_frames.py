"""Frame output (FrameWriter.java's Three.js JSON schema) and run discovery (sim_server.scan_runs).

The number formatting is pinned against Java's String.format(Locale.US, "%.Ng") outputs (every significant
digit kept, fixed notation from 1e-4 up to 10^N, two-digit exponents), the geometry against the derived ends
end1/end2 = coord -/+ L/2 u, and the schema against the reference's field order; FREE slots are left out of
the turnover-coloured frames and notADPRatio = f_ATP + f_ADP-Pi.
"""
import json
import os

import numpy as np
import pytest

from softbox_amd import scenes
from softbox_amd.frames import FrameWriter, _g, resolve_output_dir, scan_runs
from softbox_amd.turnover import TurnoverStore


@pytest.mark.parametrize("x,digits,want", [
    (1.0, 5, "1.0000"), (0.1, 5, "0.10000"), (12345.678, 5, "12346"), (123456.0, 5, "1.2346e+05"),
    (1e-5, 5, "1.0000e-05"), (0.0035, 5, "0.0035000"), (-2.5, 5, "-2.5000"), (0.0, 5, "0.0000"),
    (1.0, 3, "1.00"), (0.333333, 3, "0.333"), (0.01, 6, "0.0100000"), (1.0e-4, 5, "0.00010000"),
])
def test_java_g_formatting(x, digits, want):
    assert _g(x, digits) == want


def test_output_dir_increments(tmp_path):
    d = str(tmp_path / "run")
    assert resolve_output_dir(d) == d
    assert resolve_output_dir(d) == d + ".001"
    assert resolve_output_dir(d) == d + ".002"


def test_frame_schema_and_geometry(tmp_path):
    sc = scenes.network(n_fil=3, segs=4, box=2.0, n_links=0, seed=1)
    w = FrameWriter(str(tmp_path / "f"), 2.0, 2.0, 2.0)
    text = w.frame_json(sc.fs, 0.00125)
    d = json.loads(text)
    assert list(d) == ["frame", "t", "bounds", "segments"] and d["frame"] == 0 and d["t"] == 0.00125
    assert text.startswith('{"frame":0,"t":0.00125000,"bounds":{"xDim":2.0000,"yDim":2.0000,"zDim":2.0000}')
    assert len(d["segments"]) == 12
    seg = d["segments"][5]
    assert list(seg) == ["id", "end1", "end2", "r", "notADPRatio", "cofilinCount", "isBarbedEnd"]
    i = 5
    c = sc.fs.coord.reshape(3, -1)[:, i].astype(np.float64)
    u = sc.fs.uVec.reshape(3, -1)[:, i].astype(np.float64)
    L = float(sc.fs.segLength[i])
    assert np.allclose(seg["end1"], c - 0.5 * L * u, atol=1e-4) and np.allclose(seg["end2"], c + 0.5 * L * u, atol=1e-4)
    assert seg["r"] == 0.0035 and seg["notADPRatio"] == 1.0 and seg["cofilinCount"] == 0
    barbed = [s["isBarbedEnd"] for s in d["segments"]]
    assert barbed == [k % 4 == 3 for k in range(12)]          # end2 free: the last segment of each chain
    w.write_frame(sc.fs, 0.0)
    w.write_frame(sc.fs, 1e-5)
    assert sorted(os.listdir(w.out_dir)) == ["frame_000000.json", "frame_000001.json"]
    assert json.load(open(os.path.join(w.out_dir, "frame_000001.json")))["frame"] == 1


def test_turnover_frames_skip_free_slots_and_carry_the_composition(tmp_path):
    sc = scenes.network(n_fil=2, segs=3, box=2.0, n_links=0, seed=2)
    t = scenes.with_turnover(sc, spare=4)
    ts = t.turn
    C = ts.C
    ts.nucFrac[:C] = 0.5
    ts.nucFrac[C:2 * C] = 0.25
    ts.nucFrac[2 * C:] = 0.25
    ts.filState[1] = -1
    w = FrameWriter(str(tmp_path / "g"), 2.0, 2.0, 2.0)
    d = json.loads(w.frame_json(t.fs, 0.5, ts))
    ids = [s["id"] for s in d["segments"]]
    assert ids == [0, 2, 3, 4, 5]                            # slot 1 and the 4 spare slots are FREE
    s0 = d["segments"][0]
    assert list(s0)[-3:] == ["fATP", "fADPPi", "fADP"]
    assert s0["notADPRatio"] == 0.75 and s0["fATP"] == 0.5 and s0["fADP"] == 0.25


def test_scan_runs_finds_run_folders(tmp_path):
    a = tmp_path / "a" / "run1"
    b = tmp_path / "b"
    hidden = tmp_path / ".git" / "run2"
    for d in (a, b, hidden):
        d.mkdir(parents=True)
        (d / "frame_000000.json").write_text("{}")
    (b / "frame_000001.json").write_text("{}")
    (b / "sub").mkdir()
    (b / "sub" / "frame_000000.json").write_text("{}")       # inside a run folder: not searched
    runs = scan_runs(str(tmp_path))
    assert sorted(r["path"] for r in runs) == ["a/run1", "b"]
    assert {r["path"]: r["frameCount"] for r in runs} == {"a/run1": 1, "b": 2}
