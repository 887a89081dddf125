#!/usr/bin/env python
"""Forward throughput of the FDTD time loop on one MI355X.

Workload (the 32M-cell forward case): 400x400x200 cells, FP32, dx = 50 nm, Courant 0.99/sqrt(3),
12-cell CPML on all faces, background eps_r = 2.1 with a seeded 200x200x22 design slab
(eps = 2.1 + 10 rho), an x-directed Gaussian-pulse plane wave (1.55 um, 20 % bandwidth) and a
Poynting-flux plane plus a 3-frequency phasor plane on the transmission side. A "step" is one full
FDTD time step (E sweep, H sweep, detectors) of that grid; the timed region runs --steps steps
with all state resident in HBM.

Prints one JSON line: Mcell-updates/s (median and best of --reps repetitions), the HBM roofline of
the dominant kernel (the E sweep, k_update_E4 with its listed-tile launch, timed with HIP events on
the engine's stream), a CPU baseline (the oracle's C/OpenMP forward loop on every host thread
OMP_NUM_THREADS allows, over a bounded sample of steps), the 256^3 dispersive forward, the seconds
per gradient of the 200x200x100 differentiation case, and served-vs-serial throughput with p50/p99
latency on a seeded Poisson request mix.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(spec_plan, snap_cells=0):
    """Compulsory HBM bytes of one step of the two-sweep Yee leapfrog (no temporal blocking).

    E sweep: read H (3), E (3), inv_eps (C_eps), write E (3) per cell; psi_E read+write (4) per
    CPML-slab cell; ADE 12 words per pole-cell; the H_prev snapshot (3 words) over the detector
    box. H sweep: read E (3), H (3), inv_mu (C_mu), write H (3); psi_H (4) per slab cell.
    Detector state: phasor read+write of 2 x 8 B per frequency, component and plane cell."""
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
    e_sweep = 4 * (N * (9 + ec) + 4 * n_slab + 12 * pole_cells + 3 * snap_cells)
    h_sweep = 4 * (N * (9 + mc) + 4 * n_slab)
    det = 0
    for d in spec_plan["detectors"]:
        cells = 1
        for lo, hi in zip(d["lo"], d["hi"]):
            cells *= hi - lo
        if d["kind"] == 3:
            det += 2 * 8 * d["cos"].shape[0] * len(d["components"]) * cells
    return e_sweep, h_sweep, det, n_slab


def snapshot_cells(spec_plan):
    """Cells of the H_prev snapshot box: the union of the detector boxes grown by one cell."""
    if not spec_plan["detectors"]:
        return 0
    lo = [min(max(0, d["lo"][a] - 1) for d in spec_plan["detectors"]) for a in range(3)]
    hi = [max(min(spec_plan["shape"][a], d["hi"][a] + 1) for d in spec_plan["detectors"]) for a in range(3)]
    return (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2])


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def copy_calibration(gib=1, reps=20):
    """The streaming ceiling: a 1 GiB torch device copy (the reference dispatch for calibrating
    rocprofv3 FETCH_SIZE / WRITE_SIZE), then the engine's hand-written stream probe
    (fdtdx_stream_probe) at the sweeps' read:write mixes, dword and dwordx4 lanes, with and without
    nontemporal access, over 32M floats (128 MB) per array."""
    import torch

    from fdtdx_amd import engine

    n = gib * 2 ** 30 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    print(json.dumps({"probe": "torch_copy", "copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "ms_per_copy": round(ms, 4), "gb_per_s": round(2 * n * 4 / ms / 1e6, 1)}), flush=True)
    del a, b
    lib = engine.load_library()
    m = 32 * 2 ** 20
    src = torch.rand(7 * m, dtype=torch.float32, device="cuda")
    dst = torch.empty(3 * m, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    h = ctypes.c_void_p(stream.cuda_stream)
    best = {}
    for rd, wr in ((1, 1), (7, 3), (6, 3), (4, 1), (1, 0)):
        for vec in (1, 4):
            for nt in (0, 1):
                def go():
                    rc = lib.fdtdx_stream_probe(rd, wr, vec, nt, ctypes.c_void_p(src.data_ptr()),
                                                ctypes.c_void_p(dst.data_ptr()), m, h)
                    if rc:
                        raise RuntimeError(lib.fdtdx_last_error().decode())
                for _ in range(3):
                    go()
                s.record(stream)
                for _ in range(reps):
                    go()
                e.record(stream)
                torch.cuda.synchronize()
                ms = s.elapsed_time(e) / reps
                gbs = (rd + wr) * m * 4 / ms / 1e6
                key = f"{rd}:{wr}"
                best[key] = max(best.get(key, 0.0), gbs)
                print(json.dumps({"probe": f"stream r{rd} w{wr} vec{vec} nt{nt}", "bytes": (rd + wr) * m * 4,
                                  "ms": round(ms, 4), "gb_per_s": round(gbs, 1)}), flush=True)
    print(json.dumps({"probe": "best", **{k: round(v, 1) for k, v in best.items()}}), flush=True)


def adjoint_bench():
    """Seconds per gradient of the 200x200x100 differentiation case (3000 steps, dipole, seeded
    100x100x20 design region, time-integrated output flux): forward with interface recording
    plus the time-reversed adjoint, from resident state (host set-up excluded)."""
    import torch

    import fdtdx_amd as fx
    from fdtdx_amd import scenarios

    def one(spec):
        objects, arrays, params, config = scenarios.build(spec, gradient=True)
        rho = torch.tensor(params["design"], device=arrays.fields.E.device, requires_grad=True)
        arrays, objects, _ = fx.apply_params(arrays, objects, {"design": rho})
        torch.cuda.synchronize()
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        t0 = time.perf_counter()
        e0.record()
        _, arrays = fx.run_fdtd(arrays, objects, config)
        flux = next(iter(arrays.detector_states["flux"].values()))
        e1.record()
        (-flux.sum()).backward()
        e2.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        assert torch.isfinite(rho.grad).all()
        return wall, e0.elapsed_time(e1) / 1e3, e1.elapsed_time(e2) / 1e3

    one(scenarios.scenario_c3(T=40, shape=(40, 40, 30), pml=6, design=(10, 10, 6)))
    spec = scenarios.scenario_c3()
    wall, fwd, adj = one(spec)
    return {"s_per_grad": round(wall, 3), "forward_s": round(fwd, 3), "adjoint_s": round(adj, 3),
            "config": "200x200x100, CPML 10, dipole, 100x100x20 design (eps 1..12.1), flux objective, "
                      f"{spec['T']} steps, reversible adjoint with interface recording"}


def serving_bench(n=64, step_scale=1.0, seed=0, rate=8.0, solo_cells=1_000_000, solo_max=3, reverse_workers=2):
    """Served (continuous batching) against serial plain solves on the same seeded request mix:
    grid sizes from {64^3, 96^3, 128x128x64, 192^3, 256x256x128}, step counts uniform in
    [500, 4000] scaled by `step_scale`, 30 % gradient requests, Poisson arrivals at `rate` per
    second (above the device's throughput, so the makespan is set by the device). The default is
    the first 64 requests of the 512-request mix; --serving-requests 512 runs the whole mix.
    Requests of at least `solo_cells` cells (128x128x64, 192^3 and 256x256x128) run alone beside
    the batch."""
    import numpy as np
    import torch

    from fdtdx_amd import serving as S

    reqs = S.make_requests(n, rate, seed, step_scale=step_scale)
    warm = S.Request(-1, (24, 24, 24), 8, 1, "dipole", True, 0.0, 4)
    S.run_plain(warm)
    total = torch.cuda.get_device_properties(0).total_memory
    be = S.GpuBackend(use_graph=True, solo_workers=solo_max, reverse_workers=reverse_workers)
    srv = S.Server(be, S.HbmBudget(int(total - 16 * 2 ** 30)), chunk=16, solo_cells=solo_cells, solo_max=solo_max)
    srv.run([warm])
    res, stats = srv.run(reqs)
    be.close()
    served = S.summarize(reqs, res, stats.makespan)
    times, sres = S.run_serial(reqs)
    serial_s = sum(times.values())
    same = all(res[r.rid].status == "ok" and np.array_equal(res[r.rid].flux, sres[r.rid][0]) and
               (sres[r.rid][1] is None or np.array_equal(res[r.rid].grad, sres[r.rid][1])) for r in reqs)
    return {"requests": n, "step_scale": step_scale, "served_req_per_s": round(served["requests_per_s"], 3),
            "serial_req_per_s": round(n / serial_s, 3),
            "speedup": round(served["requests_per_s"] / (n / serial_s), 3),
            "p50_latency_s": round(served["p50_latency_s"], 3), "p99_latency_s": round(served["p99_latency_s"], 3),
            "results_identical_to_serial": bool(same), "arrivals": f"Poisson, {rate} requests/s",
            "steps": f"uniform in [{int(500 * step_scale)}, {int(4000 * step_scale)}]", "gradient_fraction": 0.3,
            "solo_cells": solo_cells, "solo_max": solo_max}


def cpu_baseline(spec, steps=40):
    """The oracle's forward loop on the host: the C/OpenMP restatement (oracle/oracle_c.c) on every
    thread OMP_NUM_THREADS allows, over `steps` steps of the same workload from zero fields (a
    bounded sample of the schedule; about 10-30 s of CPU work). The Python/NumPy restatement is not
    timed: it is the checker, one core, about 2.8 s per step of this grid."""
    import platform

    import numpy as np

    from oracle import fast as F
    from oracle import scenes as S

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    sc = S.build_scene(spec)
    F.run(sc, 2, threads=threads)  # warm-up: page in the tables
    t0 = time.perf_counter()
    F.run(sc, steps, threads=threads)
    dt = time.perf_counter() - t0
    cells = int(np.prod(spec["shape"]))
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": round(cells * steps / dt / 1e6, 2), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "cpu": model,
            "sample": f"steps 0..{steps - 1} of the same 400x400x200 workload from zero fields ({dt:.1f} s), "
                      "C/OpenMP FP32 restatement of the oracle's forward loop (restatement, not the reference: "
                      "the reference's JAX CPU path cannot run in this image)"}


def time_forward(spec, steps, warmup, reps):
    """Build the workload through the public API, warm up, then time `reps` repetitions of `steps`
    steps with plain fdtdx_forward calls bracketed by HIP events on the engine's stream (cycling
    through the T-step schedule); then one event-bracketed run (fdtdx_forward_timed) for the E and
    H sweep kernels' per-launch times."""
    import numpy as np
    import torch

    from fdtdx_amd import engine, fdtd, scenarios

    T = spec["T"]
    objects, arrays, params, config = scenarios.build(spec)
    solve = fdtd._solve_for(objects, config, arrays)
    b = solve.buffers()
    stream = engine.current_stream_handle()
    cells = int(np.prod(spec["shape"]))
    print(f"[bench] workload built: {spec['shape']} cells={cells}", file=sys.stderr, flush=True)
    state = {"t": 0}

    def run(nsteps, timed=False):
        done, ms = 0, [0.0, 0.0]
        while done < nsteps:
            t = state["t"]
            t1 = min(T, t + (nsteps - done))
            if timed:
                e, h = engine.forward_timed(solve.plan, b, t, t1, False, stream)
                ms[0] += e
                ms[1] += h
            else:
                engine.forward(solve.plan, b, t, t1, False, stream)
            done += t1 - t
            state["t"] = 0 if t1 == T else t1
        return ms

    run(warmup)
    torch.cuda.synchronize()
    print(f"[bench] warmup {warmup} steps done", file=sys.stderr, flush=True)
    ts = torch.cuda.current_stream()
    rep_ms = []
    for r in range(reps):
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(ts)
        run(steps)
        stop.record(ts)
        torch.cuda.synchronize()
        rep_ms.append(start.elapsed_time(stop) / steps)
        print(f"[bench] rep {r}: {rep_ms[-1]:.4f} ms/step", file=sys.stderr, flush=True)
    ksteps = min(steps, T)
    e_ms, h_ms = run(ksteps, timed=True)
    torch.cuda.synchronize()
    e_b, h_b, det_b, n_slab = algorithmic_bytes(solve.spec, snapshot_cells(solve.spec))
    sched = {0: "one cell per lane", 1: "four z cells per lane (dwordx4)", 2: "fused", 3: "march"}[solve.plan.schedule]
    kname = ("k_update_E4", "k_update_H4") if solve.plan.schedule == 1 else ("k_update_E", "k_update_H")
    # the quad sweeps run the tiles of sources and other per-cell descriptors in a second launch
    # (k_update_E4s / k_update_H4s); the HIP events bracket both launches of a sweep
    sweep = {"k_update_E4": "E sweep: k_update_E4 + k_update_E4s", "k_update_H4": "H sweep: k_update_H4 + k_update_H4s"}
    e_launch, h_launch = e_ms / ksteps, h_ms / ksteps
    wl = "x".join(str(v) for v in spec["shape"])
    roof = {
        "bound": "hbm",
        "kernel": sweep.get(kname[0], kname[0]),
        "achieved": round(e_b / (e_launch / 1e3) / 1e9, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(e_b / (e_launch / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(kname[0], wl),
        "bytes_per_launch": e_b,
        "ms_per_launch": round(e_launch, 4),
        kname[1]: {"bytes_per_launch": h_b, "ms_per_launch": round(h_launch, 4),
                   "achieved": round(h_b / (h_launch / 1e3) / 1e9, 1),
                   "traffic": measured_traffic(kname[1], wl)},
    }
    rep_sorted = sorted(rep_ms)
    med = rep_sorted[len(rep_sorted) // 2] if len(rep_sorted) % 2 else 0.5 * (
        rep_sorted[len(rep_sorted) // 2 - 1] + rep_sorted[len(rep_sorted) // 2])
    out = {"cells": cells, "cpml_cells": n_slab, "step_bytes": e_b + h_b + det_b, "rep_ms_per_step": rep_ms,
           "median_ms_per_step": med, "best_ms_per_step": rep_sorted[0], "roofline": roof, "schedule": sched}
    del solve, b, arrays, objects
    return out


def c4_bench(reps=5, steps=None):
    """The dispersive + anisotropic forward: 256^3, CPML 12, diagonal eps (2, 4, 9) slab with two
    Lorentz poles, a Drude cube, a dipole and a 5-frequency phasor plane, 4000 steps per
    repetition; roofline against its algorithmic bytes (98 B per cell update)."""
    from fdtdx_amd import scenarios

    spec = scenarios.scenario_c4()
    nsteps = steps or spec["T"]
    f = time_forward(spec, nsteps, 200, reps)
    med = f["median_ms_per_step"]
    gbs = f["step_bytes"] / (med / 1e3) / 1e9
    return {"workload": "forward 256^3, CPML 12, diagonal-eps Lorentz slab + Drude cube, dipole, 5-frequency "
                        f"phasor plane, {nsteps} steps",
            "value": round(f["cells"] / (med / 1e3) / 1e6, 1), "unit": "Mcell-updates/s",
            "ms_per_step": round(med, 4), "best_ms_per_step": round(f["best_ms_per_step"], 4),
            "step_bytes": f["step_bytes"], "bytes_per_cell": round(f["step_bytes"] / f["cells"], 1),
            "hbm_roofline_pct": round(100 * gbs / HBM_PEAK_GBS, 1), "sweep_schedule": f["schedule"],
            "roofline": f["roofline"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--reps", type=int, default=5, help="timed repetitions of --steps steps (median reported)")
    ap.add_argument("--no-c4", action="store_true", help="skip the 256^3 dispersive forward line")
    ap.add_argument("--c4-steps", type=int, default=0,
                    help="steps per repetition of the 256^3 line (default: its 4000-step schedule)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the seconds-per-gradient measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the served-vs-serial measurement")
    ap.add_argument("--serving-requests", type=int, default=64,
                    help="requests of the serving mix (512: the whole mix, several minutes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only time a 1 GiB device-to-device torch copy (the streaming ceiling and the "
                         "reference dispatch for calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
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
    fwd = time_forward(spec, args.steps, args.warmup, args.reps)
    cells = fwd["cells"]
    med = fwd["median_ms_per_step"]
    value = cells / (med / 1e3) / 1e6
    step_b = fwd["step_bytes"]
    out = {
        "metric": "Mcell-updates/s & % HBM roofline",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "reps": args.reps,
        "ms_per_step": round(med, 4),
        "best_ms_per_step": round(fwd["best_ms_per_step"], 4),
        "rep_ms_per_step": [round(v, 4) for v in fwd["rep_ms_per_step"]],
        "best_value": round(cells / (fwd["best_ms_per_step"] / 1e3) / 1e6, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded design slab)",
        "hbm_roofline_pct": round(100 * step_b / (med / 1e3) / 1e9 / HBM_PEAK_GBS, 1),
        "config": {
            "workload": "forward 400x400x200 (32M cells), CPML 12, eps 2.1 + 200x200x22 seeded design slab, "
                        "x plane wave (TFSF), flux + 3-frequency phasor planes",
            "cells": cells, "cpml_cells": fwd["cpml_cells"], "schedule_steps": spec["T"], "parallelism": "single GPU",
            "step_bytes": step_b, "sweep_schedule": fwd["schedule"],
            "timing": f"median (and best) of {args.reps} repetitions of {args.steps} steps, HIP events on the "
                      "engine's stream, plain fdtdx_forward; per-kernel times from a separate event-bracketed run",
        },
        "roofline": fwd["roofline"],
        "cpu_baseline": None,
    }
    torch.cuda.empty_cache()
    if not args.no_c4:
        out["c4"] = c4_bench(args.reps, args.c4_steps or None)
        print(f"[bench] c4 {json.dumps(out['c4'])}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
        print(f"[bench] adjoint {out['adjoint']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_serving:
        out["serving"] = serving_bench(args.serving_requests)
        print(f"[bench] serving {out['serving']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, args.cpu_steps)
        print(f"[bench] cpu baseline {out['cpu_baseline']}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
