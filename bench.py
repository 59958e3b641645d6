"""Benchmark of the grand-potential solver loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mesh 4096] [--seeds 1024] [--modes fp32,fp64]

Workload (the headline configuration): ternary three-phase grand-potential model, two solids + liquid,
4096 x 4096 interior cells (16.8 M), 1024 circular seeds (radius 6-12, Poisson-disc spacing 60, seeded RNG),
parabolic free energies with random SPD Hessians, 4-fold anisotropy, obstacle potential, periodic sides,
dt at 0.8 of the explicit limit; interface energy 0.2 and a liquid fill at driving force 0.04 for both solids,
so the seeds are supercritical and grow (configs.config_ternary). A step is one phi sweep + one mu sweep over
the whole grid; one cell-update is one interior cell advanced by one step.

Prints one JSON line. ``value`` is cell-updates/s of the first mode in --modes with the state resident in HBM,
every step including the reference loop's per-step statistics (``no_stats`` gives the same steps without
them), wall clock of one gp_step call of --steps steps (as a driver issues between two progress reports; the
call ends with one evaluation of the solute and energy sums of the final state); the other modes appear under
``modes``. ``roofline`` is for the whole step:
the algorithmic bytes 2 (N + NC) sizeof(real) per cell-update (phi and mu read once and written once) times every
interior cell, over the step's duration from HIP events recorded on the solver's stream around the whole call
(nothing is recorded between its launches). ``roofline_kernels`` gives each update kernel on its own: the cells one
launch advances (k_cbulk: the bulk cells, k_ilist: the interface cells) over that kernel's mean duration, from a
separate pass with events around every launch. ``late`` repeats the timings after --late-at steps of the same run,
where the seeds have grown (its interface fraction is reported). A streaming-copy kernel is timed on the same device
as a measured bandwidth reference. ``cpu_baseline`` times the serial reference binary (oracle/_ref, built from the
reference sources) on a bounded sample: the same ternary model on 1024 x 1024 with 64 seeds, per-step wall time from
two runs that differ only in their step count. ``configs`` reports FP64 throughput on configurations 0 (binary seed,
200^2) and 1 (ternary, 1024^2) next to the serial reference on the same grids.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
MODES = ("fp32c", "fp32", "fp64")
METRIC = "cell-updates/s & HBM-roofline % (FP64 vs FP32); adjoint/forward time ratio"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(mesh: int, seeds: int, seed: int = 1234):
    from microsim_amd import configs, host
    spacing = 60.0 if mesh >= 1024 else 30.0
    inp, fills, dtl = configs.config_ternary(mesh, seeds, seed=seed, min_spacing=spacing)
    d = tempfile.mkdtemp(prefix="gp_bench_")
    configs.write_case(d, inp, fills)
    t = time.time()
    case = host.load_case(os.path.join(d, "Input.in"), os.path.join(d, "Filling.in"))
    log("workload %dx%d, %d seeds, dt %.6g (0.8 x %.6g), host set-up %.1fs" % (mesh, mesh, seeds, inp.dt, dtl,
                                                                               time.time() - t))
    return case, d


def binary_seed_case(mesh: int, steps: int, radius: float = 60.0):
    """Binary anisotropic seed material (configs.config_c0: c_s 0.2, c_l 0.4, liquid at 0.34) at mesh^2 with a tanh
    seed of the given radius (radius 10 is below the critical radius of this material and melts)."""
    from microsim_amd import configs, host
    inp, _ = configs.config_c0(mesh, steps, steps)
    d = tempfile.mkdtemp(prefix="gp_bench_c0_")
    configs.write_case(d, inp, [])
    phi_i, mu_i = configs.tanh_seed_state(mesh, mesh, radius, 2.0, [2.0 * 1.0 * 0.34])
    return host.state_case(os.path.join(d, "Input.in"), phi_i, mu_i)


def cell_classes(phi: "np.ndarray"):
    """(bulk cells, interface cells) of a halo-inclusive state, as the cell-classified update splits them: a bulk
    cell's 13-cell diamond (Manhattan radius 2) of phase codes is purely one phase."""
    import numpy as np
    N = phi.shape[0]
    code = np.full(phi.shape[1:], -1, dtype=np.int8)
    nz = np.zeros(phi.shape[1:], dtype=np.int8)
    for a in range(N):
        code[phi[a] == 1.0] = a
        nz += (phi[a] != 0.0)
    code[nz != 1] = -1
    c = code[1:-1, 1:-1]  # cells 1 .. rows-2, so the diamond of every interior cell is inside
    ok = c[2:-2, 2:-2] >= 0
    centre = c[2:-2, 2:-2]
    for dx, dy in ((1, 0), (-1, 0), (2, 0), (-2, 0), (0, 1), (0, -1), (0, 2), (0, -2), (1, 1), (1, -1), (-1, 1), (-1, -1)):
        ok &= c[2 + dx:c.shape[0] - 2 + dx, 2 + dy:c.shape[1] - 2 + dy] == centre
    nb = int(ok.sum())
    return nb, int(ok.size) - nb


def _timed_call(s, steps: int, stats: bool):
    """(wall seconds, event ms) of one gp_step call of `steps` steps; the events bracket the whole call on the
    solver's stream (GP_PROFILE_CALLS: nothing is recorded between its launches)."""
    s.profile_read()
    t0 = time.perf_counter()
    st = s.step(steps, stats=stats)
    s.synchronize()
    el = time.perf_counter() - t0
    ev = s.profile_read()["call"][0]
    return el, ev, st


def run_mode(case, precision: int, steps: int, warmup: int, late_at: int = 0):
    """Time `steps` solver steps with the reference loop's per-step statistics (cumulative max/min, sums of
    squared increments) folded on the device every step, as the driver runs them; then the same steps without the
    statistics work. Both timings take the whole gp_step call: wall clock and HIP events on the solver's stream,
    with no per-launch instrumentation. A separate pass brackets every launch with events for the per-kernel split.
    late_at > 0: the run continues to step late_at and the timings are repeated there (the interface fraction
    grows as the seeds grow)."""
    from microsim_amd import _abi
    from microsim_amd.solver import GPSolver
    N, NC = case.N, case.NC
    cells = int(case.params.mesh_x) * int(case.params.mesh_y)
    # algorithmic bytes: the state read and written once per cell-update. FP64: 8 B per field; FP32: 4 B; the
    # compensated mode: 8 B per phi field (two FP32 parts) and 5 B per mu field (FP32 + 8-bit extension)
    b_alg = {0: 2 * (N + NC) * 8, 1: 2 * (N + NC) * 4, 2: 2 * (8 * N + 5 * NC)}[precision]

    def window(s, done):
        phi_a, _ = s.download()
        s.profile_enable(_abi.GP_PROFILE_CALLS)
        el, ev, st = _timed_call(s, steps, True)
        el_ns, ev_ns, _ = _timed_call(s, steps, False)
        # per-kernel split: events around every launch (separate steps, not part of the timings above)
        s.profile_enable(_abi.GP_PROFILE_LAUNCHES)
        s.profile_read()
        s.step(steps, stats=True)
        prof = s.profile_read()
        s.step(steps)
        prof_ns = s.profile_read()
        s.profile_enable(False)
        phi_b, _ = s.download()
        cls = [cell_classes(p) for p in (phi_a, phi_b)]
        nbulk = 0.5 * (cls[0][0] + cls[1][0])
        kcells = {"k_cbulk": nbulk, "k_ilist": cells - nbulk, "k_march": cells, "k_march_stats": cells}
        per_kernel = {}
        for k, (ms, cnt) in prof.items():
            if k in kcells and ms > 0:
                gbs = kcells[k] * b_alg / (ms * 1e-3) / 1e9
                per_kernel[k] = {"ms": round(ms, 4), "cells": int(kcells[k]), "GBps": round(gbs, 1),
                                 "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_per_launch(k, precision)}
        step_ms = ev / steps
        step_gbs = cells * b_alg / (step_ms * 1e-3) / 1e9
        traffic = [traffic_per_launch(k, precision) for k in prof]
        return {
            "at_step": done,
            "value": cells * steps / el,
            "ms_per_step": el * 1e3 / steps,
            "event_ms_per_step": round(step_ms, 5),
            "interface_cell_fraction": round(1.0 - nbulk / cells, 4),
            "kernels_ms": {k: round(v[0], 4) for k, v in prof.items()},
            "kernel_sum_ms": round(sum(ms for ms, cnt in prof.values()), 4),
            "no_stats": {"value": cells * steps / el_ns, "ms_per_step": el_ns * 1e3 / steps,
                         "event_ms_per_step": round(ev_ns / steps, 5),
                         "kernels_ms": {k: round(v[0], 4) for k, v in prof_ns.items()}},
            "roofline": {"bound": "hbm", "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                         "traffic": sum(traffic) if traffic and all(t is not None for t in traffic) else None,
                         "kernel": "whole step (%s), HIP events around one call of %d steps"
                                   % (" + ".join(sorted(prof)), steps),
                         "alg_bytes_per_cell_update": b_alg, "cells_per_launch": cells},
            "roofline_kernels": per_kernel,
            "nonfinite": st["nonfinite"],
        }

    with GPSolver(case.params, precision) as s:
        s.upload(case.phi, case.mu)
        s.step(warmup, stats=True)
        s.synchronize()
        done = warmup
        res = window(s, done)
        done += 4 * steps
        if late_at > done:
            s.step(late_at - done, stats=True)
            done = late_at
            res["late"] = window(s, done)
    return res


def adjoint_ratio(params, phi0, mu0, steps: int, every: int = 0, precision: int = None):
    """Cost of the adjoint against the forward of the same precision (FP64 by default). `steps` plain steps are
    timed, then the same steps recorded for the adjoint (checkpoints every about sqrt(steps)) and the reverse sweep for a
    random linear loss of the final state. From HIP events on the solver's stream: the forward step (its kernels),
    the transposed step (reload of the recomputed state, delta-phi, the adjoint kernels) and the segment
    recompute per step; the adjoint step is the transposed step plus the recompute share. Wall-clock times
    include the host copies of the loss gradient in and of the state and parameter gradients out."""
    import numpy as np
    from microsim_amd import _abi
    from microsim_amd.solver import GPSolver
    rng = np.random.default_rng(5)
    wphi, wmu = rng.standard_normal(phi0.shape), rng.standard_normal(mu0.shape)
    with GPSolver(params, _abi.GP_FP64 if precision is None else precision) as s:
        s.upload(phi0, mu0)
        s.step(2)
        s.upload(phi0, mu0)
        s.synchronize()
        s.profile_enable(True)
        s.profile_read()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        fwd = time.perf_counter() - t0
        pf = s.profile_read()
        s.upload(phi0, mu0)
        s.synchronize()
        t0 = time.perf_counter()
        s.adjoint_forward(steps, every)
        s.synchronize()
        rec = time.perf_counter() - t0
        s.profile_read()
        t0 = time.perf_counter()
        s.adjoint_backward(wphi, wmu)
        bwd = time.perf_counter() - t0
        pb = s.profile_read()
        s.profile_enable(False)
    fwd_step = sum(ms * n for ms, n in pf.values()) / steps
    adj_ms, adj_n = pb.pop("adj_step")
    recompute = sum(ms * n for ms, n in pb.values()) / adj_n
    adj_step = adj_ms + recompute
    return {"steps": steps, "checkpoint_every": every or int(np.ceil(np.sqrt(steps))), "forward_step_ms": round(fwd_step, 4), "transposed_step_ms": round(adj_ms, 4),
            "recompute_ms_per_step": round(recompute, 4), "adjoint_step_ms": round(adj_step, 4),
            "adjoint_over_forward": round(adj_step / fwd_step, 3),
            "wall": {"forward_ms_per_step": round(fwd * 1e3 / steps, 4),
                     "recorded_forward_ms_per_step": round(rec * 1e3 / steps, 4),
                     "reverse_ms_per_step": round(bwd * 1e3 / steps, 4)}}


def copy_bandwidth(nbytes: int = 1 << 31, reps: int = 10):
    """Streaming-copy reference on the same device: y = -x over nbytes (one elementwise kernel reads nbytes and
    writes nbytes), mean over reps. Reported next to the vendor peak."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.neg(x, out=y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.neg(x, out=y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del x, y
    return {"bytes_per_launch": 2 * nbytes, "ms": round(ms, 4), "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def traffic_per_launch(kernel: str, precision: int):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the
    same counters on the copy kernel of known size (profiles/traffic_tern4096.txt); None when not measured."""
    path = os.path.join(REPO, "profiles", "traffic_tern4096.txt")
    if not os.path.exists(path):
        return None
    key = "%s %s bytes_per_launch" % (kernel, {0: "fp64", 1: "fp32", 2: "fp32c"}.get(precision, "fp32"))
    for line in open(path):
        if line.startswith(key):
            return float(line.split("=")[1])
    return None


def precision_errors(case, steps: int, checkpoints=(100, 500, 1000, 2000, 5000, 10000)):
    """Max and L2 (root-mean-square) error against the FP64 mode, over interior cells, for phi, mu and the
    composition c = sum_b c_b(mu) h_b(phi), at each checkpoint up to `steps`, of

      fp32:              the FP32 production mode from the same initial state;
      fp64_rounded_init: the FP64 mode started from the initial state rounded to FP32 (the perturbation that
                         storing the state in FP32 makes before the first step; a floor for any FP32 storage).

    The FP64 mode reproduces the FP64 oracle bit for bit (tests/test_gpu_parity.py), so these are errors against
    the oracle. Each chunk of steps is issued with statistics (non-finite check)."""
    from microsim_amd import _abi, host
    from microsim_amd.solver import GPSolver
    import numpy as np
    t0 = time.perf_counter()
    runs = {"fp64": GPSolver(case.params, _abi.GP_FP64), "fp32": GPSolver(case.params, _abi.GP_FP32),
            "fp32c": GPSolver(case.params, _abi.GP_FP32C),
            "fp64_rounded_init": GPSolver(case.params, _abi.GP_FP64)}
    res = {"steps": steps, "mesh": [int(case.params.mesh_x), int(case.params.mesh_y)], "at": []}
    sl = (slice(None), slice(3, -3), slice(3, -3))
    try:
        runs["fp64"].upload(case.phi, case.mu)
        runs["fp32"].upload(case.phi, case.mu)
        runs["fp32c"].upload(case.phi, case.mu)
        runs["fp64_rounded_init"].upload(case.phi.astype(np.float32).astype(np.float64),
                                          case.mu.astype(np.float32).astype(np.float64))
        done = 0
        for cp in sorted(c for c in set(checkpoints) | {steps} if c <= steps):
            for r in runs.values():
                r.step(cp - done, stats=True)
            done = cp
            ref = [np.ascontiguousarray(a[sl]) for a in runs["fp64"].download()]
            cref = host.composition(case.params, ref[0], ref[1])
            row = {"step": cp}
            for name in ("fp32c", "fp32", "fp64_rounded_init"):
                got = [np.ascontiguousarray(a[sl]) for a in runs[name].download()]
                cgot = host.composition(case.params, got[0], got[1])
                row[name] = {}
                for f, a, b in (("phi", got[0], ref[0]), ("mu", got[1], ref[1]), ("c", cgot, cref)):
                    d = a - b
                    row[name][f] = {"max": float(np.abs(d).max()), "l2": float(np.sqrt(np.mean(d * d)))}
            res["at"].append(row)
            log("precision", json.dumps(row))
    finally:
        for r in runs.values():
            r.close()
    res["seconds"] = round(time.perf_counter() - t0, 1)
    return res


def reference_rate(inp, fills, restart_state, sample_steps: int):
    """Cell-updates/s of the serial reference binary (oracle/_ref, built from the reference sources, 1 core) on the
    given case: the wall-time difference of a (2 + sample_steps)-step and a 2-step run (set-up and the initial
    write cancel). None when the binary is absent or fails."""
    ref = os.path.join(REPO, "oracle", "_ref", "microsim_gp")
    if not os.path.exists(ref):
        return None
    from microsim_amd import configs
    times = {}
    for n in (2, 2 + sample_steps):
        inp.ntimesteps, inp.savet, inp.track_progress = n, n, n
        d = tempfile.mkdtemp(prefix="gp_cpu_")
        configs.write_case(d, inp, fills, restart_state=restart_state)
        t = time.perf_counter()
        r = subprocess.run([ref, "Input.in", "Filling.in", "o"], cwd=d, capture_output=True, text=True)
        times[n] = time.perf_counter() - t
        if r.returncode != 0:
            return None
    per = (times[2 + sample_steps] - times[2]) / sample_steps
    return int(inp.mesh_x) * int(inp.mesh_y) / per


def cpu_baseline(sample_steps: int = 20):
    """Serial reference binary (1 core) on the 1024^2 ternary sample (64 seeds: the seed density of the headline
    workload)."""
    from microsim_amd import configs
    inp, fills, _ = configs.config_ternary(1024, 64, seed=1234)
    v = reference_rate(inp, fills, None, sample_steps)
    if v is None:
        return None
    return {"value": v, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": "serial reference microsim_gp (gcc -O2), ternary 3-phase 1024x1024, 64 seeds (the headline "
                      "workload's seed density), %d steps (difference of a %d- and a 2-step run)"
                      % (sample_steps, 2 + sample_steps)}


def config_rates(cpu_c1=None, c0_steps: int = 2000, c1_steps: int = 200):
    """FP64 throughput of the two reference-scale configurations next to the serial reference on the same grid:
    config 0 (binary two-phase anisotropic seed, 200^2, tanh seed of radius 10 via a restart file, Neumann sides)
    and config 1 (ternary three-phase, 1024^2, 64 seeds, periodic). GPU: wall clock of one gp_step call with the
    per-step statistics."""
    from microsim_amd import _abi, configs, host
    from microsim_amd.solver import GPSolver
    out = {}
    # config 0
    inp, _ = configs.config_c0(200, c0_steps, c0_steps)
    d = tempfile.mkdtemp(prefix="gp_c0_")
    configs.write_case(d, inp, [])
    phi_i, mu_i = configs.tanh_seed_state(200, 200, 10.0, 2.0, [2.0 * 1.0 * 0.34])
    case0 = host.state_case(os.path.join(d, "Input.in"), phi_i, mu_i)
    with GPSolver(case0.params, _abi.GP_FP64) as s:
        s.upload(case0.phi, case0.mu)
        s.step(20, stats=True)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(c0_steps, stats=True)
        s.synchronize()
        el = time.perf_counter() - t0
    g0 = 200 * 200 * c0_steps / el
    c0 = reference_rate(inp, [], (phi_i, mu_i, "o"), 200)
    out["c0_binary200_fp64"] = {"gpu": g0, "gpu_ms_per_step": el * 1e3 / c0_steps, "cpu_reference_1core": c0,
                                "gpu_over_cpu": (g0 / c0) if c0 else None, "unit": "cell-updates/s",
                                "steps": c0_steps}
    # config 1
    case1, _ = make_case(1024, 64)
    with GPSolver(case1.params, _abi.GP_FP64) as s:
        s.upload(case1.phi, case1.mu)
        s.step(5, stats=True)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(c1_steps, stats=True)
        s.synchronize()
        el = time.perf_counter() - t0
    g1 = 1024 * 1024 * c1_steps / el
    out["c1_ternary1024_fp64"] = {"gpu": g1, "gpu_ms_per_step": el * 1e3 / c1_steps,
                                  "cpu_reference_1core": cpu_c1, "gpu_over_cpu": (g1 / cpu_c1) if cpu_c1 else None,
                                  "unit": "cell-updates/s", "steps": c1_steps}
    return out


def oracle_threads(case_small, nthreads: int, steps: int = 3):
    """The FP64 restatement threaded over host cores, on the same 1024^2 sample (extra baseline)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import gp_testlib
    phi, mu = case_small.phi.copy(), case_small.mu.copy()
    gp_testlib.oracle_step(case_small.params, phi, mu, 1, threads=nthreads)
    t = time.perf_counter()
    gp_testlib.oracle_step(case_small.params, phi, mu, steps, threads=nthreads)
    el = time.perf_counter() - t
    return 1024 * 1024 * steps / el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="steps of each timed gp_step call (a driver's call between two progress reports; the call "
                         "ends with one evaluation of the solute/energy invariants)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mesh", type=int, default=4096)
    ap.add_argument("--seeds", type=int, default=1024)
    ap.add_argument("--modes", default="fp32c,fp32,fp64",
                    help="the first is the headline (fp32c: the compensated FP32 production mode)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--late-at", type=int, default=5000,
                    help="repeat the timings after this many steps of the same run (0: skip)")
    ap.add_argument("--no-configs", action="store_true", help="skip the config 0 / config 1 FP64 figures")
    ap.add_argument("--adjoint-steps", type=int, default=20, help="steps of the adjoint timing (0: skip)")
    ap.add_argument("--error-steps", type=int, default=10000,
                    help="steps of the FP32-vs-FP64 error study at the bench mesh and at 1024^2 (0: skip)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs the solver loop on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is a design note, not a measured path)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from microsim_amd import _abi
    PRECISIONS = {"fp64": _abi.GP_FP64, "fp32": _abi.GP_FP32, "fp32c": _abi.GP_FP32C}
    case, _ = make_case(a.mesh, a.seeds)
    copy = copy_bandwidth()
    log("copy kernel", json.dumps(copy))
    res = {}
    for mode in a.modes.split(","):
        prec = PRECISIONS[mode]
        res[mode] = run_mode(case, prec, a.steps, a.warmup, a.late_at)
        log(mode, json.dumps(res[mode]))
    first = a.modes.split(",")[0]
    r0 = res[first]
    out = {
        "metric": METRIC, "value": r0["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r0["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if first == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": "GP ternary 3-phase %dx%d, %d seeds, %s mode" % (a.mesh, a.mesh, a.seeds, first),
                   "mesh": [a.mesh, a.mesh], "phases": 3, "components": 3, "seeds": a.seeds,
                   "parallelism": "single-gpu"},
        "roofline": dict(r0["roofline"], measured_copy_GBps=copy["GBps"]),
        "roofline_kernels": r0["roofline_kernels"],
        "copy_kernel": copy,
        "modes": {m: {k: v for k, v in r.items()} for m, r in res.items()},
    }
    if "fp32" in res and "fp64" in res:
        out["fp32_over_fp64"] = res["fp32"]["value"] / res["fp64"]["value"]
    if "fp32c" in res and "fp64" in res:
        out["fp32c_over_fp64"] = res["fp32c"]["value"] / res["fp64"]["value"]
    if a.adjoint_steps > 0:
        out["adjoint"] = adjoint_ratio(case.params, case.phi, case.mu, a.adjoint_steps)
        out["adjoint"]["workload"] = "this workload, FP64, random linear loss of the final state"
        log("adjoint", json.dumps(out["adjoint"]))
        bcase = binary_seed_case(256, 2000)
        out["adjoint_binary256"] = adjoint_ratio(bcase.params, bcase.phi, bcase.mu, 2000, 100)
        out["adjoint_binary256"]["workload"] = ("binary anisotropic seed (radius 60) 256x256, 2000 steps, "
                                                "checkpoints every 100, FP64")
        log("adjoint binary 256", json.dumps(out["adjoint_binary256"]))
        out["adjoint_binary256_fp32"] = adjoint_ratio(bcase.params, bcase.phi, bcase.mu, 2000, 100, _abi.GP_FP32)
        out["adjoint_binary256_fp32"]["workload"] = ("binary anisotropic seed (radius 60) 256x256, 2000 steps, "
                                                     "checkpoints every 100, FP32 trajectory, FP64 Jacobians and "
                                                     "gradient accumulation")
        log("adjoint binary 256 fp32", json.dumps(out["adjoint_binary256_fp32"]))
    if a.error_steps > 0:
        errs = [precision_errors(case, a.error_steps)]
        if a.mesh != 1024:
            errs.append(precision_errors(make_case(1024, 64)[0], a.error_steps))
        out["fp32_error_vs_fp64"] = errs
        log("fp32 error", json.dumps(errs))
    cb = None
    if not a.no_cpu:
        cb = cpu_baseline()
        if cb is not None:
            out["cpu_baseline"] = cb
            try:
                small, _ = make_case(1024, 64)
                nth = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
                out["cpu_oracle_threads"] = {"value": oracle_threads(small, nth), "cores": nth,
                                             "kind": "port", "unit": "cell-updates/s"}
            except Exception as e:  # extra figure only
                log("threaded oracle skipped:", e)
    if not a.no_configs:
        out["configs"] = config_rates(cb["value"] if cb else None)
        log("configs", json.dumps(out["configs"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
