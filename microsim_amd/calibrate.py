"""Adjoint-driven calibration: recover perturbed model parameters from a synthetic target microstructure.

The workload is the ternary three-phase seed layout of ``configs.config_ternary``. A forward run with the true
parameters makes the target phi. The calibration starts from the selected parameters offset by +-20% and runs
Adam on their logarithms. The loss is L = mean over interior cells and phases of (phi_T - phi_target)^2, and its
gradient comes from the device adjoint (``adjoint.trajectory``: checkpoints about every sqrt(T) steps, segment
recompute, transposed steps). The step size decays geometrically. The run prints one JSON line per iteration
and a summary with:
- the loss;
- the relative error of every selected parameter;
- the forward and backward wall times;
- the adjoint's device memory for the state copies it keeps.

    python -m microsim_amd.calibrate --mesh 2048 --steps 1000 --iters 15 --precision fp32

A long run can be split over several processes: with ``--state FILE`` the optimiser state (log-parameters, Adam
moments, step size, history) is written after every iteration and read back at the start, and ``--budget-s``
ends a process before an iteration that would not finish within that many seconds; the next process with the same
arguments continues where it stopped (the target is recomputed, bit for bit the same).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np
import torch

from . import _abi, adjoint, configs, host

SELECT = ("dab[0][2]", "dab[1][2]", "tau_ab[0][2]", "tau_ab[1][2]", "D[2][0][0]", "D[2][1][1]")


def _case(mesh: int, seeds: int):
    inp, fills, _ = configs.config_ternary(mesh, seeds, seed=1234, min_spacing=60.0 if mesh >= 1024 else 30.0)
    d = tempfile.mkdtemp(prefix="gp_calib_")
    configs.write_case(d, inp, fills)
    return host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"))


def main(argv=None, summary_out: dict = None) -> int:
    """Run the calibration; the summary is printed and, when summary_out is given, stored in it."""
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--mesh", type=int, default=2048)
    ap.add_argument("--seeds", type=int, default=256)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--iters", type=int, default=15)
    ap.add_argument("--lr", type=float, default=0.05)
    ap.add_argument("--lr-decay", type=float, default=0.8, help="step size factor per iteration")
    ap.add_argument("--precision", choices=("fp32c", "fp32", "fp64"), default="fp32")
    ap.add_argument("--offset", type=float, default=0.2, help="relative offset of the starting parameters")
    ap.add_argument("--state", default=None, help="optimiser state file: read at the start, written every iteration")
    ap.add_argument("--budget-s", type=float, default=0.0,
                    help="stop before an iteration that would end later than this many seconds after the start")
    a = ap.parse_args(argv)
    t_start = time.perf_counter()
    # a state file resumes only the run that wrote it (checked before any device work)
    run_key = {k: getattr(a, k) for k in ("mesh", "seeds", "steps", "iters", "lr", "lr_decay", "precision", "offset")}
    saved = None
    if a.state and os.path.exists(a.state):
        saved = json.load(open(a.state))
        if saved.get("run") != run_key:
            raise SystemExit("calibrate: %s belongs to another run (%s)" % (a.state, saved.get("run")))
    prec = {"fp32c": _abi.GP_FP32C, "fp32": _abi.GP_FP32, "fp64": _abi.GP_FP64}[a.precision]
    case = _case(a.mesh, a.seeds)
    names = adjoint.theta_names(case.N, case.NC + 1)
    idx = [names.index(n) for n in SELECT]
    th_true = adjoint.theta_from_params(case.params)
    signs = np.array([1.0, -1.0, -1.0, 1.0, 1.0, -1.0])
    th_start = th_true.copy()
    th_start[idx] *= 1.0 + a.offset * signs
    every = int(math.ceil(math.sqrt(a.steps)))
    phi0 = torch.tensor(case.phi)
    mu0 = torch.tensor(case.mu)

    t0 = time.perf_counter()
    with torch.no_grad():
        target, _ = adjoint.trajectory(case.params, torch.tensor(th_true), phi0, mu0, a.steps, every, prec)
    target = target[:, 3:-3, 3:-3].clone()
    t_target = time.perf_counter() - t0

    z = torch.tensor(np.log(th_start[idx]), requires_grad=True)
    opt = torch.optim.Adam([z], lr=a.lr)
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=a.lr_decay)
    base = torch.tensor(th_start)
    cells = case.phi.shape[1] * case.phi.shape[2]
    # bytes of one stored state: FP32 planes (plus phi's FP32 low parts and mu's 8-bit extensions in fp32c), or FP64
    state_bytes = {_abi.GP_FP64: 8 * (case.N + case.NC), _abi.GP_FP32: 4 * (case.N + case.NC),
                   _abi.GP_FP32C: 8 * case.N + 5 * case.NC}[prec] * cells
    ncheck = (a.steps + every - 1) // every
    adj_bytes = (ncheck + every) * state_bytes + 3 * (case.N + case.NC) * case.phi.shape[1] * case.phi.shape[2] * 8
    hist = []
    it0 = 0
    if saved is not None:
        st = saved
        with torch.no_grad():
            z.copy_(torch.tensor(st["z"], dtype=z.dtype))
        if st["adam_step"] > 0:
            opt.state[z] = {"step": torch.tensor(float(st["adam_step"])),
                            "exp_avg": torch.tensor(st["exp_avg"], dtype=z.dtype),
                            "exp_avg_sq": torch.tensor(st["exp_avg_sq"], dtype=z.dtype)}
        for g in opt.param_groups:
            g["lr"] = st["lr"]
        hist, it0 = st["hist"], st["iter"]

    def save_state(next_it):
        if not a.state:
            return
        ad = opt.state.get(z, {})
        st = {"run": run_key, "iter": next_it, "z": z.detach().tolist(), "lr": opt.param_groups[0]["lr"],
              "adam_step": int(ad["step"]) if ad else 0,
              "exp_avg": ad["exp_avg"].tolist() if ad else None, "exp_avg_sq": ad["exp_avg_sq"].tolist() if ad else None,
              "hist": hist}
        tmp = a.state + ".tmp"
        json.dump(st, open(tmp, "w"))
        os.replace(tmp, a.state)

    for it in range(it0, a.iters + 1):
        if a.budget_s > 0 and hist:
            per_it = hist[-1]["forward_s"] + hist[-1]["backward_s"]
            if time.perf_counter() - t_start + 1.2 * per_it > a.budget_s:
                print(json.dumps({"stopped_before_iter": it, "state": a.state}), flush=True)
                return 0
        th = base.clone()
        th[idx] = torch.exp(z)
        t1 = time.perf_counter()
        phiT, _ = adjoint.trajectory(case.params, th, phi0, mu0, a.steps, every, prec)
        loss = torch.mean((phiT[:, 3:-3, 3:-3] - target) ** 2)
        t2 = time.perf_counter()
        opt.zero_grad()
        if it < a.iters:
            loss.backward()
        t3 = time.perf_counter()
        err = np.abs(th.detach().numpy()[idx] / th_true[idx] - 1.0)
        rec = {"iter": it, "loss": float(loss.detach()), "max_rel_param_error": float(err.max()),
               "rel_param_error": {n: round(float(e), 5) for n, e in zip(SELECT, err)},
               "forward_s": round(t2 - t1, 3), "backward_s": round(t3 - t2, 3)}
        hist.append(rec)
        print(json.dumps(rec), flush=True)
        if it < a.iters:
            opt.step()
            sched.step()
        # the trajectory's device buffers go with its graph (the last iteration takes no backward)
        del loss, phiT
        save_state(it + 1)
    fw = np.mean([h["forward_s"] for h in hist[:-1]]) if len(hist) > 1 else float("nan")
    bw = np.mean([h["backward_s"] for h in hist[:-1]]) if len(hist) > 1 else float("nan")
    summary = {
        "workload": "ternary 3-phase %dx%d, %d seeds, %d steps, checkpoints every %d, %s" % (
            a.mesh, a.mesh, a.seeds, a.steps, every, a.precision),
        "parameters": list(SELECT), "start_offset": a.offset, "iterations": a.iters, "lr": a.lr,
        "lr_decay": a.lr_decay,
        "loss_first": hist[0]["loss"], "loss_last": hist[-1]["loss"],
        "max_rel_param_error_first": hist[0]["max_rel_param_error"],
        "max_rel_param_error_last": hist[-1]["max_rel_param_error"],
        "mean_forward_s": round(float(fw), 3), "mean_backward_s": round(float(bw), 3),
        "backward_over_forward_wall": round(float(bw / fw), 2) if fw > 0 else None,
        "target_forward_s": round(t_target, 3),
        "adjoint_state_bytes": int(adj_bytes)}
    print(json.dumps({"summary": summary}), flush=True)
    if summary_out is not None:
        summary_out.update(summary)
    return 0


if __name__ == "__main__":
    sys.exit(main())
