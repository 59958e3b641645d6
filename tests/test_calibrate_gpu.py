"""The calibration example (microsim_amd/calibrate.py) at a small size: a few descent iterations on the device
adjoint's gradients lower the loss and the parameter error from the 20% starting offset."""
import sys

import pytest


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32c", "fp32", "fp64"])
def test_calibration_descends(precision):
    from microsim_amd import calibrate
    print("[gpu] calibrate 256^2", precision, flush=True)
    sys.stdout.flush()
    summary = {}
    assert calibrate.main(["--mesh", "256", "--seeds", "8", "--steps", "200", "--iters", "4",
                           "--precision", precision], summary_out=summary) == 0
    assert summary["loss_last"] < 0.5 * summary["loss_first"]
    assert summary["max_rel_param_error_last"] < summary["max_rel_param_error_first"]


@pytest.mark.gpu
def test_calibration_2048_20000_steps_descends():
    """The calibration workload at its configured scale: ternary three-phase 2048^2 (256 seeds), 20 000 steps with
    checkpoints every ceil(sqrt(20000)) = 142 steps, compensated FP32 trajectory with FP64 gradient accumulation;
    one descent iteration (a 20 000-step gradient, about 70 s) lowers the loss and the largest parameter error from
    the 20% offsets."""
    from microsim_amd import calibrate
    print("[gpu] calibrate 2048^2 x 20000 steps, 1 iteration (fp32c)", flush=True)
    summary = {}
    assert calibrate.main(["--mesh", "2048", "--seeds", "256", "--steps", "20000", "--iters", "1",
                           "--precision", "fp32c"], summary_out=summary) == 0
    assert summary["loss_last"] < summary["loss_first"]
    assert summary["max_rel_param_error_last"] < summary["max_rel_param_error_first"]
