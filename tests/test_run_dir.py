"""srmhd.run_2d.RunDir, the run directory's bookkeeping (CPU only): the
refusal to write into a directory that holds another run (run_2d.py:240-261
refuses the same way), run_meta.json rewritten whole, and the snapshot and
restart cadence (snapshot k at t = k tend / nsnap, to 1e-14; a restart file
every --restart-every steps except at the end)."""

import json
import os

import pytest

from srmhd.run_2d import RunDir


def _meta(**kw):
    m = dict(problem="rotor", n=32, solver="hlld", limiter="mc", cfl=0.4)
    m.update(kw)
    return m


def test_record_rewrites_the_whole_file(tmp_path):
    rd = RunDir(str(tmp_path), _meta(), 0.4, 5, 10)
    rd.record()
    rd.record(steps=7)
    rd.record_cost(9, 9, 3.21, stopped_early=True)
    with open(tmp_path / "run_meta.json") as fh:
        m = json.load(fh)
    assert m["steps"] == 9 and m["steps_here"] == 9 and m["wall_s"] == 3.2
    assert m["s_per_step"] == round(3.21 / 9, 3) and m["stopped_early"] is True
    assert "finished" in m and m["problem"] == "rotor"
    # no temporary file is left behind
    assert sorted(os.listdir(tmp_path)) == ["run_meta.json"]


def test_refuses_a_directory_holding_another_run(tmp_path):
    RunDir(str(tmp_path), _meta(), 0.4, 5, 10).record()
    # the same run (other settings may differ): accepted
    RunDir(str(tmp_path), _meta(cfl=0.2), 0.4, 5, 10).refuse_other_run()
    for key, val in (("n", 64), ("solver", "hlle"), ("limiter", "weno5z"),
                     ("problem", "orszag_tang")):
        with pytest.raises(SystemExit, match=f"{key} .* != "):
            RunDir(str(tmp_path), _meta(**{key: val}), 0.4, 5, 10).refuse_other_run()
    # an empty directory is never refused
    RunDir(str(tmp_path / "new"), _meta(n=64), 0.4, 5, 10).refuse_other_run()


def test_snapshot_and_restart_cadence():
    tend, nsnap = 0.4, 5
    rd = RunDir("unused", _meta(), tend, nsnap, 10)
    # snapshot k is due from t = k tend / nsnap - 1e-14 on, and never past nsnap
    for k in range(1, nsnap + 1):
        tk = k * tend / nsnap
        assert not rd.snap_due(tk - 1e-13, k)
        assert rd.snap_due(tk - 0.5e-14, k) and rd.snap_due(tk, k)
    assert not rd.snap_due(10.0, nsnap + 1)
    assert rd.restart_due(20, False) and not rd.restart_due(21, False)
    assert not rd.restart_due(20, True)
    assert not RunDir("unused", _meta(), tend, nsnap, 0).restart_due(20, False)
