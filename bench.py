"""Benchmark of the grand-potential solver loop on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mesh 4096] [--seeds 1024] [--modes fp32,fp64]

Workload (the headline configuration): ternary three-phase grand-potential model, two solids + liquid,
4096 x 4096 interior cells (16.8 M), 1024 circular seeds (radius 6-12, Poisson-disc spacing 60, seeded RNG),
parabolic free energies with random SPD Hessians, 4-fold anisotropy, obstacle potential, periodic sides,
dt at 0.8 of the explicit limit. A step is one phi sweep + one mu sweep over the whole grid; one
cell-update is one interior cell advanced by one step.

Prints one JSON line. ``value`` is cell-updates/s of the first mode in --modes with the state resident in HBM;
the other modes appear under ``modes``. ``roofline`` uses the algorithmic bytes 2 (N + NC) sizeof(real) per
cell-update (phi and mu read once and written once) over the mean duration of the step's kernels, timed with
HIP events on the solver's stream inside the timed region. ``cpu_baseline`` times the serial reference binary
(oracle/_ref, built from the reference sources) on a bounded sample: the same ternary model on 1024 x 1024 with
64 seeds, per-step wall time from two runs that differ only in their step count.
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


def run_mode(case, precision: int, steps: int, warmup: int):
    from microsim_amd.solver import GPSolver
    N, NC = case.N, case.NC
    cells = int(case.params.mesh_x) * int(case.params.mesh_y)
    es = 8 if precision == 0 else 4
    with GPSolver(case.params, precision) as s:
        s.upload(case.phi, case.mu)
        s.step(warmup)
        s.synchronize()
        s.profile_enable(True)
        s.profile_read()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        el = time.perf_counter() - t0
        prof = s.profile_read()
        s.profile_enable(False)
        st = s.step(1, stats=True)  # one reporting step after the timed region (non-finite check)
    step_ms = sum(ms for name, (ms, cnt) in prof.items() if name != "k_stats")
    dom = max((k for k in prof if k != "k_stats"), key=lambda k: prof[k][0])
    b_alg = 2 * (N + NC) * es
    achieved = cells * b_alg / (step_ms * 1e-3) / 1e9 if step_ms > 0 else 0.0
    return {
        "value": cells * steps / el,
        "ms_per_step": el * 1e3 / steps,
        "kernels_ms": {k: round(v[0], 4) for k, v in prof.items()},
        "dominant": dom,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "+".join(k for k in prof if k != "k_stats"),
                     "alg_bytes_per_cell_update": b_alg},
        "nonfinite": st["nonfinite"],
    }


def cpu_baseline(sample_steps: int = 20):
    """Serial reference binary (1 core) on the 1024^2 ternary sample; falls back to the oracle port."""
    from microsim_amd import configs
    ref = os.path.join(REPO, "oracle", "_ref", "microsim_gp")
    times = {}
    if os.path.exists(ref):
        for n in (2, 2 + sample_steps):
            inp, fills, _ = configs.config_ternary(1024, 64, seed=1234, steps=n, savet=n)
            inp.track_progress = n
            d = tempfile.mkdtemp(prefix="gp_cpu_")
            configs.write_case(d, inp, fills)
            t = time.perf_counter()
            r = subprocess.run([ref, "Input.in", "Filling.in", "o"], cwd=d, capture_output=True, text=True)
            times[n] = time.perf_counter() - t
            if r.returncode != 0:
                times = {}
                break
        if times:
            per = (times[2 + sample_steps] - times[2]) / sample_steps
            return {"value": 1024 * 1024 / per, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                    "sample": "serial reference microsim_gp (gcc -O2), ternary 3-phase 1024x1024, 64 seeds, "
                              "%d steps (difference of a %d- and a 2-step run)" % (sample_steps, 2 + sample_steps)}
    return None


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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mesh", type=int, default=4096)
    ap.add_argument("--seeds", type=int, default=1024)
    ap.add_argument("--modes", default="fp32,fp64")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs the solver loop on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is a design note, not a measured path)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from microsim_amd import _abi
    case, _ = make_case(a.mesh, a.seeds)
    res = {}
    for mode in a.modes.split(","):
        prec = _abi.GP_FP32 if mode == "fp32" else _abi.GP_FP64
        res[mode] = run_mode(case, prec, a.steps, a.warmup)
        log(mode, json.dumps(res[mode]))
    first = a.modes.split(",")[0]
    r0 = res[first]
    out = {
        "metric": METRIC, "value": r0["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r0["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if first == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": "GP ternary 3-phase %dx%d, %d seeds, %s mode" % (a.mesh, a.mesh, a.seeds, first),
                   "mesh": [a.mesh, a.mesh], "phases": 3, "components": 3, "seeds": a.seeds,
                   "parallelism": "single-gpu"},
        "roofline": r0["roofline"],
        "modes": {m: {k: v for k, v in r.items()} for m, r in res.items()},
    }
    if "fp32" in res and "fp64" in res:
        out["fp32_over_fp64"] = res["fp32"]["value"] / res["fp64"]["value"]
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
