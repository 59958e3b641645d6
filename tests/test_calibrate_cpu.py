"""Resumable calibration runs (microsim_amd/calibrate.py --state): a state file resumes only the run that wrote it,
and the check happens before any device work (so it runs here, without a GPU)."""
import json

import pytest

from microsim_amd import calibrate


def test_state_of_another_run_is_refused(tmp_path):
    state = tmp_path / "state.json"
    run = {"mesh": 64, "seeds": 4, "steps": 10, "iters": 2, "lr": 0.05, "lr_decay": 0.8, "precision": "fp32",
           "offset": 0.2}
    state.write_text(json.dumps({"run": run, "iter": 1}))
    with pytest.raises(SystemExit, match="belongs to another run"):
        calibrate.main(["--mesh", "128", "--seeds", "4", "--steps", "10", "--iters", "2", "--state", str(state)])
