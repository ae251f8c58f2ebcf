"""CPU checks of the learned-guess evaluation reports: summaries and the merge of report halves."""
import json

import numpy as np

from rest_amd import train_guess as tg


def _row(seed, mixed, dE, kind="learned", err=0.05):
    return {"n": 64, "seed": seed, "sad_iterations": 11, "sad_energy": -1.0, "sad_guess_energy": -2.0,
            "sad_seconds": 5.0, "learned_kind": kind, "guard": "error", "learned_guess_energy": -1.5,
            "sad_d1_energy": -1.2, "learned_guess_error": err, "sad_d1_error": 0.25, "learned_electrons": 640.0,
            "learned_iterations": 10, "learned_fp64_iterations": 3, "learned_energy": -1.0 + dE,
            "learned_seconds": 1.7, "dE": dE, "learned_incremental_iterations": 2, "converged": True,
            "mixed": mixed}


def test_summarize_per_mode():
    rows = [_row(1, True, 1e-10), _row(2, True, -3e-10, err=0.09), _row(3, True, 2e-10, kind="sad"),
            _row(1, False, 5e-11)]
    s = tg.summarize(rows)
    mx, fp = s["mixed"], s["fp64"]
    assert mx["clusters"] == 3 and mx["kept"] == 2 and mx["converged"] == 3 and fp["clusters"] == 1
    assert mx["max_abs_dE"] == 3e-10
    assert np.allclose(mx["learned_guess_error_min_mean_max"], [0.05, (0.05 + 0.09 + 0.05) / 3, 0.09])
    assert np.allclose(mx["learned_minus_sad_d1_energy_min_mean_max"], [-0.3, -0.3, -0.3])


def test_merge_joins_report_halves(tmp_path):
    a, b, out = tmp_path / "a.json", tmp_path / "b.json", tmp_path / "all.json"
    a.write_text(json.dumps({"model": "m", "guard": "error", "test": [_row(1, True, 1e-10)]}))
    b.write_text(json.dumps({"model": "m", "guard": "error", "test": [_row(2, True, -2e-10), _row(3, True, 0.0)]}))
    assert tg.main(["--merge", str(a), str(b), "--report", str(out)]) == 0
    d = json.loads(out.read_text())
    assert [r["seed"] for r in d["test"]] == [1, 2, 3]
    assert d["summary"]["mixed"]["clusters"] == 3 and d["summary"]["mixed"]["max_abs_dE"] == 2e-10
