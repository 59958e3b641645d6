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


@pytest.mark.gpu
def test_calibration_config4_scale_descends():
    """The calibration workload at its configured scale: ternary three-phase 2048^2 (256 seeds), 20 000 steps with
    checkpoints every ceil(sqrt(20000)) = 142 steps, compensated FP32 trajectory with FP64 gradient accumulation;
    three descent iterations lower the loss and the largest parameter error from the 20% offsets. (Its per-iteration
    lines go straight to stdout: nothing captures them, so a long run reports as it goes.)"""
    from microsim_amd import calibrate
    print("[gpu] calibrate 2048^2 x 20000 steps, 3 iterations (fp32c)", flush=True)
    summary = {}
    assert calibrate.main(["--mesh", "2048", "--seeds", "256", "--steps", "20000", "--iters", "3",
                           "--precision", "fp32c"], summary_out=summary) == 0
    assert summary["loss_last"] < summary["loss_first"]
    assert summary["max_rel_param_error_last"] < summary["max_rel_param_error_first"]
