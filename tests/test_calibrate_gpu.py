"""The calibration example (microsim_amd/calibrate.py) at a small size: a few descent iterations on the device
adjoint's gradients lower the loss and the parameter error from the 20% starting offset."""
import json
import sys

import pytest


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_calibration_descends(precision, capsys):
    from microsim_amd import calibrate
    print("[gpu] calibrate 256^2", precision, flush=True)
    sys.stdout.flush()
    assert calibrate.main(["--mesh", "256", "--seeds", "8", "--steps", "200", "--iters", "4",
                           "--precision", precision]) == 0
    out = capsys.readouterr().out
    summary = json.loads([l for l in out.splitlines() if l.startswith('{"summary"')][-1])["summary"]
    assert summary["loss_last"] < 0.5 * summary["loss_first"]
    assert summary["max_rel_param_error_last"] < summary["max_rel_param_error_first"]
