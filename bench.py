#!/usr/bin/env python
"""Forward throughput of the FDTD time loop on one MI355X.

Workload (the 32M-cell forward case): 400x400x200 cells, FP32, dx = 50 nm, Courant 0.99/sqrt(3),
12-cell CPML on all faces, background eps_r = 2.1 with a seeded 200x200x22 design slab
(eps = 2.1 + 10 rho), an x-directed Gaussian-pulse plane wave (1.55 um, 20 % bandwidth) and a
Poynting-flux plane plus a 3-frequency phasor plane on the transmission side. A "step" is one full
FDTD time step (E sweep, H sweep, detectors) of that grid; the timed region runs --steps steps
with all state resident in HBM.

Prints one JSON line: Mcell-updates/s, the HBM roofline of the field-update sweeps (HIP events on
the engine's stream), and a CPU baseline (the NumPy oracle on the host, one core).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(spec_plan):
    """Compulsory HBM bytes of one step of the two-sweep Yee leapfrog (no temporal blocking):
    4 * [N (18 + C_eps + C_mu) + 8 N_slab + 12 pole-cells] for the sweeps, plus detector state
    traffic (phasor: 2 * 8 B * freqs * components per plane cell)."""
    nx, ny, nz = spec_plan["shape"]
    N = nx * ny * nz
    ec = spec_plan["eps_comps"]
    mc = spec_plan["mu_comps"]
    n_slab = 0
    for s in spec_plan["cpml"]:
        shp = list(spec_plan["shape"])
        shp[s["axis"]] = s["hi"] - s["lo"]
        n_slab += shp[0] * shp[1] * shp[2]
    pole_cells = 0
    for a in spec_plan["ade"]:
        cells = 1
        for lo, hi in zip(a["lo"], a["hi"]):
            cells *= hi - lo
        pole_cells += a["num_poles"] * cells
    sweeps = 4 * (N * (18 + ec + mc) + 8 * n_slab + 12 * pole_cells)
    det = 0
    for d in spec_plan["detectors"]:
        cells = 1
        for lo, hi in zip(d["lo"], d["hi"]):
            cells *= hi - lo
        if d["kind"] == 3:
            det += 2 * 8 * d["cos"].shape[0] * len(d["components"]) * cells
    return sweeps, sweeps + det, n_slab


def cpu_baseline(spec, steps=1):
    """Time the NumPy oracle (single core) on the same workload: one warm-up step, then `steps`."""
    import numpy as np

    from oracle import fdtd_oracle as O
    from oracle import scenes as S

    sc = S.build_scene(spec)
    st = O.init_state(sc)
    ds = [O.init_detector_state(d, sc) for d in sc.detectors]
    O.forward_step(st, sc, 0, ds)
    t0 = time.perf_counter()
    for n in range(1, 1 + steps):
        O.forward_step(st, sc, n, ds)
        print(f"[bench] cpu baseline step {n} done", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    cells = int(np.prod(spec["shape"]))
    return cells * steps / dt / 1e6, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the engine runs one solve per GPU and this build has no "
              f"multi-GPU path); --gpus {args.gpus} is not supported.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from fdtdx_amd import engine, fdtd, scenarios

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    spec = scenarios.scenario_c2(T=2000)
    T = spec["T"]
    objects, arrays, params, config = scenarios.build(spec)
    solve = fdtd._solve_for(objects, config, arrays)
    b = solve.buffers()
    stream = engine.current_stream_handle()
    cells = int(np.prod(spec["shape"]))
    print(f"[bench] workload built: {spec['shape']} cells={cells}", file=sys.stderr, flush=True)

    def run(nsteps, timed=False):
        """Advance nsteps steps, cycling through the T-step schedule."""
        done, ms = 0, 0.0
        while done < nsteps:
            t = run.t
            t1 = min(T, t + (nsteps - done))
            if timed:
                ms += engine.forward_timed(solve.plan, b, t, t1, False, stream)
            else:
                engine.forward(solve.plan, b, t, t1, False, stream)
            done += t1 - t
            run.t = 0 if t1 == T else t1
        return ms

    run.t = 0
    run(args.warmup)
    torch.cuda.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    sweep_ms = run(args.steps, timed=True)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(stop)
    print(f"[bench] timed {args.steps} steps: {total_ms:.1f} ms (wall {wall * 1e3:.1f} ms)", file=sys.stderr, flush=True)
    ms_step = total_ms / args.steps
    value = cells * args.steps / (total_ms / 1e3) / 1e6
    sweeps_b, step_b, n_slab = algorithmic_bytes(solve.spec)
    sweep_ms_step = sweep_ms / args.steps
    achieved = sweeps_b / (sweep_ms_step / 1e3) / 1e9
    step_gbs = step_b / (ms_step / 1e3) / 1e9
    out = {
        "metric": "Mcell-updates/s & % HBM roofline",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded design slab)",
        "hbm_roofline_pct": round(100 * step_gbs / HBM_PEAK_GBS, 1),
        "config": {
            "workload": "forward 400x400x200 (32M cells), CPML 12, eps 2.1 + 200x200x22 seeded design slab, "
                        "x plane wave (TFSF), flux + 3-frequency phasor planes",
            "cells": cells, "cpml_cells": n_slab, "schedule_steps": T, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "E+H field-update sweeps (k_update_E, k_update_H)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "bytes_per_step": sweeps_b,
            "sweep_ms_per_step": round(sweep_ms_step, 4),
            "traffic": None,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        del solve, b, arrays
        torch.cuda.empty_cache()
        v, dt = cpu_baseline(spec, args.cpu_steps)
        out["cpu_baseline"] = {"value": round(v, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
                               "sample": f"{args.cpu_steps} step(s) of the same 400x400x200 workload after one "
                                         f"warm-up step, NumPy FP32 oracle restatement ({dt:.1f} s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
