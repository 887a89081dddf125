"""Serving: continuous batching of independent FDTD solves on one GPU.

The reference has no serving mode; it runs one solve per jitted call (fdtd/fdtd.py:421-496,
wrapper.py:14-63). Here a request loop admits solves of uneven grid sizes while others run:

* Admission. A request is admitted when a batch slot and its HBM estimate both fit
  (`HbmBudget`); one that can never fit is rejected. Admitted requests are prepared (objects
  placed, arrays and plan built on the device) by worker threads while the GPU keeps running.
* Batching. Prepared solves are attached to a multi-instance batch (`engine.Batch`, one per
  material class). Each chunk of K steps is one hipGraph launch whose persistent kernels advance
  every attached solve at its own step index. Fields, psi, materials and detector states stay
  resident on the device from admission to completion.
* Gradients. A request with `gradient=True` records the CPML interface layers during its
  batched forward, then runs the time-reversed adjoint (fdtdx_reverse_adjoint) on a second stream
  that overlaps the batch, and returns d(sum of flux)/d rho.
* Health. The batch flags a slot whose reduced detector value is NaN/Inf or exceeds a blow-up
  limit; the loop evicts it and reports the request as failed.

Per-cell arithmetic and every detector reduction in a batch are those of the single-solve path,
so a request's result is bit-identical alone, in any mix, and through `run_plain`.

The scheduling logic (`Server`) talks to a backend; `GpuBackend` drives the HIP engine. Tests of
admission, eviction and slot reuse drive the same `Server` with a host-only backend.
"""

from __future__ import annotations

import argparse
import json
import sys
import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

SIZES = ((64, 64, 64), (96, 96, 96), (128, 128, 64), (192, 192, 192), (256, 256, 128))
EPS_A, EPS_B = 1.0, 12.1


@dataclass
class Request:
    rid: int
    shape: tuple
    steps: int
    seed: int
    source: str            # "dipole" or "plane"
    gradient: bool
    arrival: float = 0.0   # seconds after the start of the run
    pml: int = 10

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))


@dataclass
class Result:
    rid: int
    status: str                      # "ok", "rejected", "failed"
    flux: np.ndarray | None = None   # (T, 1) time trace of the output-plane Poynting flux
    grad: np.ndarray | None = None   # d(sum flux)/d rho over the design region (gradient requests)
    arrival: float = 0.0
    done: float = 0.0
    reason: str = ""

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def make_requests(n, rate=None, seed=0, sizes=SIZES, steps=(500, 4000), grad_frac=0.3, step_scale=1.0, pml=10):
    """Seeded request mix: sizes uniform over `sizes`, step counts uniform in `steps` (scaled by
    `step_scale`), dipole or plane-wave source with equal odds, a fraction `grad_frac` asking for
    gradients, and Poisson arrivals at `rate` per second (all at t=0 when rate is None)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        if rate:
            t += float(rng.exponential(1.0 / rate))
        shape = tuple(int(v) for v in sizes[int(rng.integers(len(sizes)))])
        T = int(rng.integers(steps[0], steps[1] + 1))
        T = max(1, int(round(T * step_scale)))
        src = "dipole" if rng.random() < 0.5 else "plane"
        grad = bool(rng.random() < grad_frac)
        out.append(Request(i, shape, T, int(rng.integers(2 ** 31 - 1)), src, grad, t, pml))
    return out


def request_spec(req: Request) -> dict:
    """Scenario of one request: vacuum with a centred random dielectric design region
    (eps = 1 + rho * 11.1, rho ~ U(0, 1) from the request seed), CPML on all faces, a z-dipole or
    an x-directed plane wave (Gaussian pulse at 1.55 um) near the low-x side, and a reduced
    Poynting-flux plane near the high-x side."""
    from .scenarios import gaussian_profile

    nx, ny, nz = req.shape
    p = req.pml
    rng = np.random.default_rng(req.seed)
    dsz = (max(2, nx // 4), max(2, ny // 4), max(2, nz // 8))
    d0 = (nx // 2 - dsz[0] // 2, ny // 2 - dsz[1] // 2, nz // 2 - dsz[2] // 2)
    rho = rng.uniform(0, 1, size=dsz).astype(np.float32)
    src_x, out_x = p + 4, nx - p - 4
    if req.source == "dipole":
        src = dict(kind="dipole", cell=(src_x, ny // 2, nz // 2), polarization=2, profile=gaussian_profile())
    else:
        src = dict(kind="plane", axis=0, index=src_x, direction="+", e_pol=(0, 0, 1), profile=gaussian_profile())
    return dict(
        shape=tuple(req.shape), dx=50e-9, T=int(req.steps), background_eps=1.0,
        design=dict(slc=tuple((d0[a], d0[a] + dsz[a]) for a in range(3)), rho=rho, eps_a=EPS_A, eps_b=EPS_B),
        pml={"thickness": p},
        sources=[src],
        detectors=[dict(name="flux", kind="flux", slc=((out_x, out_x + 1), (p, ny - p), (p, nz - p)),
                        reduce_volume=True, axis=0, direction="+")],
    )


def estimate_bytes(req: Request) -> int:
    """Upper estimate of the device bytes a request holds while admitted: fields (24 B/cell),
    two inverse-permittivity copies (8 B/cell), packed CPML psi and its adjoint-side copy, the
    flux trace and detector workspace; for gradients also the interface record (6 floats per
    interface cell per step), the adjoint fields (24 B/cell), adjoint psi and the gradient."""
    nx, ny, nz = req.shape
    N = nx * ny * nz
    p = req.pml
    faces = 2 * (ny * nz + nx * nz + nx * ny)
    psi = 4 * p * faces
    det = (ny + 2) * (nz + 2)
    b = 4 * (8 * N + 2 * psi + req.steps + 40 * det) + 2 * 1024 * 8 + (1 << 20)
    if req.gradient:
        b += 4 * (req.steps * 6 * faces + 7 * N + psi) + 4 * req.steps
    return int(b)


# Device work of a gradient request relative to its forward alone: the reversible adjoint costs
# about 3.4 forward steps per step on the 200x200x100 differentiation case.
GRADIENT_WORK = 4.4


def estimate_work(req: Request) -> float:
    """Cell updates a request costs the device (forward, plus the adjoint for gradients)."""
    return float(req.cells) * req.steps * (GRADIENT_WORK if req.gradient else 1.0)


class HbmBudget:
    """Byte budget for admission: `capacity` bytes, reserved per admitted request and released at
    completion, eviction or failure."""

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self.used = 0
        self.peak = 0
        self._lock = threading.Lock()

    def try_reserve(self, nbytes: int) -> bool:
        with self._lock:
            if self.used + nbytes > self.capacity:
                return False
            self.used += nbytes
            self.peak = max(self.peak, self.used)
            return True

    def release(self, nbytes: int):
        with self._lock:
            self.used -= nbytes
            assert self.used >= 0


@dataclass
class _Running:
    req: Request
    job: object
    nbytes: int
    slot: int
    first_chunk: int
    finished_chunk: int | None = None


@dataclass
class ServeStats:
    makespan: float = 0.0
    chunks: int = 0
    peak_slots: int = 0
    peak_bytes: int = 0
    evicted: int = 0
    rejected: int = 0
    events: list = field(default_factory=list)  # (time, rid, stage)


class Server:
    """The request loop. `backend` provides estimate / prepare / attach / detach / advance /
    is_complete / wait / flags_of / finish / start_reverse / reverse_done / reverse_result / release
    (and start_solo / solo_done / solo_result when `solo_cells` is set); see `GpuBackend` for their
    meaning.

    solo_cells: requests of at least this many cells skip the batch and run alone (forward and, for
    gradients, the adjoint) on a worker stream beside it, up to `solo_max` at a time. The batch
    sweeps run one cell per thread, the single-solve sweeps four z cells per lane, which is faster
    on grids large enough to fill the device by themselves; both give the same results bit for bit.
    None (the default) batches every request."""

    def __init__(self, backend, budget: HbmBudget, max_slots: int = 64, chunk: int = 16, prep_workers: int = 2,
                 max_inflight: int = 2, clock=time.perf_counter, idle_sleep: float = 2e-4,
                 switch_interval: float = 5e-4, policy: str = "sjf", max_active_cells: int | None = 48_000_000,
                 prep_ahead: int | None = 8, progress: int = 0, solo_cells: int | None = None, solo_max: int = 2):
        self.backend = backend
        self.budget = budget
        self.max_slots = int(max_slots)
        self.chunk = int(chunk)
        self.prep_workers = int(prep_workers)
        self.max_inflight = int(max_inflight)
        self.clock = clock
        self.idle_sleep = idle_sleep
        self.switch_interval = switch_interval
        if policy not in ("fifo", "sjf"):
            raise ValueError(f"unknown policy {policy!r}")
        self.policy = policy
        self.max_active_cells = max_active_cells
        self.prep_ahead = prep_ahead
        self.progress = int(progress)  # print a line every `progress` finished requests (0: quiet)
        self.solo_cells = solo_cells
        self.solo_max = int(solo_max)

    def _key(self, req: Request):
        """Service order: arrival order ("fifo") or least device work first ("sjf")."""
        if self.policy == "sjf":
            return (estimate_work(req), req.arrival, req.rid)
        return (req.arrival, req.rid)

    def run(self, requests) -> tuple[dict, ServeStats]:
        be = self.backend
        stats = ServeStats()
        results: dict[int, Result] = {}
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue: list[Request] = []
        preparing = {}                        # future -> (req, nbytes)
        ready: deque = deque()                # (req, job, nbytes)
        slots: list[_Running | None] = [None] * self.max_slots
        inflight: deque = deque()             # (chunk id, token, [slot ids finishing in it])
        reversing: list = []                  # (_Running, handle)
        solo: list = []                       # (_Running, handle): requests running alone
        deferred: list = []                   # (last chunk that may touch it, _Running, Result)
        chunk_id = 0
        pool = ThreadPoolExecutor(max_workers=max(1, self.prep_workers))
        # preparation threads hold the GIL in long Python stretches; switch often so the loop can
        # keep the device queue filled
        old_switch = sys.getswitchinterval()
        sys.setswitchinterval(self.switch_interval)
        t0 = self.clock()

        def now():
            return self.clock() - t0

        def finish(run: _Running, status: str, flux=None, grad=None, reason=""):
            results[run.req.rid] = Result(run.req.rid, status, flux, grad, run.req.arrival, now(), reason)
            stats.events.append((now(), run.req.rid, "done"))
            be.release(run.job)
            self.budget.release(run.nbytes)
            if self.progress and len(results) % self.progress == 0:
                print(f"[serve] {len(results)} of {len(requests)} requests done at {now():.1f} s", flush=True)

        def retired_through():
            return inflight[0][0] - 1 if inflight else chunk_id - 1

        try:
            while pending or queue or preparing or ready or inflight or reversing or solo or deferred or any(slots):
                progressed = False
                t = now()
                while pending and pending[0].arrival <= t:
                    queue.append(pending.popleft())
                    progressed = True
                # admission: first fit in service order, bounded by slots (running + in preparation)
                # and by the preparation lookahead
                reserved = sum(s is not None for s in slots) + len(preparing) + len(ready) + len(solo)
                queue.sort(key=self._key)
                for req in list(queue):
                    nb = be.estimate(req)
                    if nb > self.budget.capacity:
                        queue.remove(req)
                        results[req.rid] = Result(req.rid, "rejected", arrival=req.arrival, done=now(),
                                                  reason=f"needs {nb} B, budget {self.budget.capacity} B")
                        stats.rejected += 1
                        progressed = True
                        continue
                    if reserved >= self.max_slots:
                        break
                    if self.prep_ahead is not None and len(preparing) + len(ready) >= self.prep_ahead:
                        break
                    if self.budget.try_reserve(nb):
                        queue.remove(req)
                        stats.events.append((now(), req.rid, "admit"))
                        preparing[pool.submit(be.prepare, req)] = (req, nb)
                        reserved += 1
                        progressed = True
                stats.peak_bytes = max(stats.peak_bytes, self.budget.used)
                for fut in [f for f in preparing if f.done()]:
                    req, nb = preparing.pop(fut)
                    progressed = True
                    try:
                        ready.append((req, fut.result(), nb))
                        stats.events.append((now(), req.rid, "prepared"))
                    except Exception as exc:  # preparation failed: report, free the reservation
                        self.budget.release(nb)
                        results[req.rid] = Result(req.rid, "failed", arrival=req.arrival, done=now(),
                                                  reason=f"prepare: {exc!r}")
                # attach prepared requests to free slots, in service order, while the attached
                # unfinished solves hold fewer cells than the cap (enough to fill the device)
                if len(ready) > 1:
                    ready = deque(sorted(ready, key=lambda item: self._key(item[0])))
                # large requests run alone beside the batch (solo_cells), as soon as a worker is free
                if self.solo_cells is not None:
                    for item in list(ready):
                        req, job, nb = item
                        if req.cells < self.solo_cells or len(solo) >= self.solo_max:
                            continue
                        ready.remove(item)
                        run = _Running(req, job, nb, -1, chunk_id)
                        solo.append((run, be.start_solo(job)))
                        stats.events.append((now(), req.rid, "solo"))
                        progressed = True
                    held = deque(item for item in ready if item[0].cells >= self.solo_cells)
                    ready = deque(item for item in ready if item[0].cells < self.solo_cells)
                else:
                    held = deque()
                while ready and None in slots:
                    if self.max_active_cells is not None:
                        active = sum(s.req.cells for s in slots if s is not None and s.finished_chunk is None)
                        if active > 0 and active + ready[0][0].cells > self.max_active_cells:
                            break
                    req, job, nb = ready.popleft()
                    sl = slots.index(None)
                    be.attach(sl, job)
                    slots[sl] = _Running(req, job, nb, sl, chunk_id)
                    stats.events.append((now(), req.rid, "attach"))
                    progressed = True
                ready.extend(held)  # solo-sized requests wait for a solo worker
                stats.peak_slots = max(stats.peak_slots, sum(s is not None for s in slots))
                # one chunk for every attached solve that still has steps to run
                active = [s for s in slots if s is not None and s.finished_chunk is None]
                if active and len(inflight) <= self.max_inflight:
                    token = be.advance(self.chunk)
                    ending = []
                    for s in active:
                        if be.step_index(s.slot, s.job) >= s.req.steps:
                            s.finished_chunk = chunk_id
                            ending.append(s.slot)
                    inflight.append((chunk_id, token, ending))
                    chunk_id += 1
                    stats.chunks += 1
                    progressed = True
                # retire completed chunks (block on the oldest when the pipeline is full)
                while inflight and (len(inflight) > self.max_inflight or be.is_complete(inflight[0][1])):
                    cid, token, ending = inflight.popleft()
                    be.wait(token)
                    progressed = True
                    for s in list(slots):
                        if s is None or s.first_chunk > cid:
                            continue
                        f = be.flags_of(s.slot, s.job)
                        if f:
                            # chunks launched before the detach still advance this solve: its
                            # buffers are released once they have completed
                            be.detach(s.slot, s.job)
                            slots[s.slot] = None
                            stats.evicted += 1
                            why = "non-finite detector value" if f & 1 else "detector value above the blow-up limit"
                            deferred.append((chunk_id - 1, s, f"evicted: {why}"))
                    for sl in ending:
                        s = slots[sl]
                        if s is None or s.finished_chunk != cid:
                            continue
                        be.detach(sl, s.job)
                        slots[sl] = None
                        stats.events.append((now(), s.req.rid, "forward_done"))
                        if s.req.gradient:
                            reversing.append((s, be.start_reverse(s.job, token)))
                        else:
                            flux, _ = be.finish(s.job)
                            finish(s, "ok", flux=flux)
                for item in [d for d in deferred if d[0] <= retired_through()]:
                    deferred.remove(item)
                    finish(item[1], "failed", reason=item[2])
                    progressed = True
                for item in list(solo):
                    s, h = item
                    if be.solo_done(h):
                        solo.remove(item)
                        try:
                            flux, grad = be.solo_result(h)
                            finish(s, "ok", flux=flux, grad=grad)
                        except Exception as exc:
                            finish(s, "failed", reason=f"solo: {exc!r}")
                        progressed = True
                for item in list(reversing):
                    s, h = item
                    if be.reverse_done(h):
                        reversing.remove(item)
                        try:
                            flux, grad = be.reverse_result(h)
                            finish(s, "ok", flux=flux, grad=grad)
                        except Exception as exc:
                            finish(s, "failed", reason=f"reverse: {exc!r}")
                        progressed = True
                if not progressed:
                    time.sleep(self.idle_sleep)
        finally:
            pool.shutdown(wait=True)
            sys.setswitchinterval(old_switch)
        stats.makespan = now()
        return results, stats


# ---------------------------------------------------------------- the GPU backend
class _Job:
    __slots__ = ("req", "solve", "buffers", "rho", "flux", "batch", "nbytes")


class GpuBackend:
    """Drives the HIP engine: requests are prepared into resident device state on worker threads
    (each on its own stream), advanced in multi-instance batches by hipGraph chunks on one batch
    stream, and their adjoints run on worker streams that overlap the batch."""

    def __init__(self, device=None, use_graph: bool = True, blowup_limit: float = 0.0, reverse_workers: int = 2,
                 timing: bool = False, solo_workers: int = 2):
        import torch

        from . import engine

        self.torch = torch
        self.engine = engine
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self.stream = torch.cuda.Stream(device=self.device)
        self.handle = engine.c_void_p(self.stream.cuda_stream)
        self.use_graph = use_graph
        self.blowup = float(blowup_limit)
        self.batches = {}
        self._local = threading.local()
        self.rev_pool = ThreadPoolExecutor(max_workers=max(1, reverse_workers))
        self.solo_pool = ThreadPoolExecutor(max_workers=max(1, solo_workers))
        self.timing = timing
        self.chunk_events = []          # (start, end) per chunk when timing
        self.prep_s = []                # wall time of each preparation
        self.reverse_s = []             # wall time of each adjoint job

    def close(self):
        self.rev_pool.shutdown(wait=True)
        self.solo_pool.shutdown(wait=True)
        for b in self.batches.values():
            b.close()
        self.batches.clear()

    def _thread_stream(self):
        s = getattr(self._local, "stream", None)
        if s is None:
            s = self.torch.cuda.Stream(device=self.device)
            self._local.stream = s
        return s

    def estimate(self, req):
        return estimate_bytes(req)

    def prepare(self, req):
        import fdtdx_amd as fx

        from .fdtd import Solve
        from .scenarios import build

        torch = self.torch
        st = self._thread_stream()
        t0 = time.perf_counter()
        with torch.cuda.stream(st):
            objects, arrays, params, config = build(request_spec(req), gradient=req.gradient)
            rho = None
            if req.gradient:
                rho = torch.tensor(params["design"], device=arrays.fields.E.device, requires_grad=True)
                arrays, objects, _ = fx.apply_params(arrays, objects, {"design": rho})
            solve = Solve(objects, config, arrays, record_steps=req.steps if req.gradient else 0)
            arrays.reset()
            job = _Job()
            job.req, job.solve, job.rho = req, solve, rho
            job.buffers = solve.buffers()
            job.flux = solve.det_tensors()[0]
            job.batch = None
            st.synchronize()
        self.prep_s.append(time.perf_counter() - t0)
        return job

    def _batch_for(self, plan):
        key = (plan.spec["eps_comps"], plan.spec["mu_comps"])
        b = self.batches.get(key)
        if b is None:
            b = self.engine.Batch(key[0], key[1], self.blowup)
            self.batches[key] = b
        return b

    def attach(self, slot, job):
        b = self._batch_for(job.solve.plan)
        b.attach(slot, job.solve.plan, job.buffers, job.solve._keep, 0, job.req.steps, record=job.req.gradient)
        job.batch = b

    def detach(self, slot, job):
        job.batch.detach(slot)

    def advance(self, steps):
        if self.timing:
            ev0 = self.torch.cuda.Event(enable_timing=True)
            ev0.record(self.stream)
        for b in self.batches.values():
            if b._held:
                b.advance(steps, self.handle, graph=self.use_graph)
        ev = self.torch.cuda.Event(enable_timing=self.timing)
        ev.record(self.stream)
        if self.timing:
            self.chunk_events.append((ev0, ev))
        return ev

    def batch_busy_s(self) -> float:
        """Summed device time of the timed chunks (call after they completed)."""
        return sum(a.elapsed_time(b) for a, b in self.chunk_events) / 1e3

    def step_index(self, slot, job):
        return job.batch.step_index(slot)

    def is_complete(self, token):
        return token.query()

    def wait(self, token):
        token.synchronize()

    def flags_of(self, slot, job):
        return int(job.batch.flags()[slot])

    def finish(self, job):
        return job.flux.detach().cpu().numpy(), None

    def start_reverse(self, job, token):
        """Run the adjoint of a finished gradient request on a worker stream. The caller has
        host-waited `token` (the chunk that ended the request's forward), so neither this call nor
        the worker touches the batch stream's events: the loop thread may be capturing a new chunk
        graph on that stream at any moment, and HIP refuses a wait on (hipErrorStreamCaptureIsolation)
        or a query of (hipErrorCapturedEvent) an event whose stream is capturing."""
        del token
        return self.rev_pool.submit(self._reverse, job)

    def _reverse(self, job):
        from .reversible import reverse_gradient

        torch = self.torch
        st = self._thread_stream()
        t0 = time.perf_counter()
        with torch.cuda.stream(st):
            a = job.solve.arrays
            cot = torch.ones_like(job.flux)
            g_ie = reverse_gradient(job.solve, a.inv_permittivities, a.fields.E, a.fields.H, [cot], job.req.steps,
                                    stream=st)
            torch.autograd.backward(a.inv_permittivities, g_ie)
            grad = job.rho.grad.detach().cpu().numpy()
            flux = job.flux.detach().cpu().numpy()
            st.synchronize()
        self.reverse_s.append(time.perf_counter() - t0)
        return flux, grad

    def start_solo(self, job):
        """Run a prepared request alone on a worker stream: the single-solve forward (with the
        interface record for a gradient request, then its adjoint as in _reverse); the same kernels
        and order of operations as run_plain, so the same results bit for bit."""
        return self.solo_pool.submit(self._solo, job)

    def _solo(self, job):
        torch = self.torch
        st = self._thread_stream()
        with torch.cuda.stream(st):
            job.solve.forward(0, job.req.steps, record=job.req.gradient)
        if job.req.gradient:
            return self._reverse(job)
        with torch.cuda.stream(st):
            flux = job.flux.detach().cpu().numpy()
            st.synchronize()
        return flux, None

    def solo_done(self, h):
        return h.done()

    def solo_result(self, h):
        return h.result()

    def reverse_done(self, h):
        return h.done()

    def reverse_result(self, h):
        return h.result()

    def release(self, job):
        job.solve = job.buffers = job.rho = job.flux = None


def run_plain(req: Request, timings: dict | None = None):
    """The one-solve-at-a-time path through the public API (place_objects, apply_params, run_fdtd,
    and for gradients loss.backward() through the reversible adjoint); returns (flux, grad).
    `timings`, when given, receives the set-up time ("setup_s") of the solve."""
    import torch

    import fdtdx_amd as fx

    from .scenarios import build

    t0 = time.perf_counter()
    objects, arrays, params, config = build(request_spec(req), gradient=req.gradient)
    rho = None
    if req.gradient:
        rho = torch.tensor(params["design"], device=arrays.fields.E.device, requires_grad=True)
        arrays, objects, _ = fx.apply_params(arrays, objects, {"design": rho})
    if timings is not None:
        torch.cuda.synchronize()
        timings["setup_s"] = time.perf_counter() - t0
    _, arrays = fx.run_fdtd(arrays, objects, config)
    flux = next(iter(arrays.detector_states["flux"].values()))
    grad = None
    if req.gradient:
        flux.sum().backward()
        grad = rho.grad.detach().cpu().numpy()
    out = flux.detach().cpu().numpy()
    torch.cuda.synchronize()
    return out, grad


def run_serial(requests, setup: dict | None = None, progress: int = 0):
    """Plain solves back to back; returns per-request wall time and results (and, in `setup`,
    each request's set-up time). progress > 0 prints a line every `progress` requests."""
    times, results = {}, {}
    t0 = time.perf_counter()
    for q, r in enumerate(requests):
        t = time.perf_counter()
        tm = {}
        results[r.rid] = run_plain(r, tm)
        times[r.rid] = time.perf_counter() - t
        if setup is not None:
            setup[r.rid] = tm["setup_s"]
        if progress and (q + 1) % progress == 0:
            print(f"[serial] {q + 1} of {len(requests)} requests done at {time.perf_counter() - t0:.1f} s", flush=True)
    return times, results


def fifo_latencies(requests, durations):
    """Latencies of a single FIFO server with the given service times and the requests' arrivals."""
    free = 0.0
    out = {}
    for r in sorted(requests, key=lambda q: (q.arrival, q.rid)):
        start = max(free, r.arrival)
        free = start + durations[r.rid]
        out[r.rid] = free - r.arrival
    return out


def stage_summary(stats: ServeStats) -> dict:
    """Mean time between consecutive stages of a request (admit -> prepared -> attach ->
    forward_done -> done), and when the last preparation and the last attach happened."""
    at = {}
    for t, rid, st in stats.events:
        at.setdefault(rid, {})[st] = t
    out = {}
    order = ["admit", "prepared", "attach", "forward_done", "done"]
    for a, b in zip(order, order[1:]):
        d = [v[b] - v[a] for v in at.values() if a in v and b in v]
        if d:
            out[f"{a}->{b}_mean_s"] = round(float(np.mean(d)), 4)
    for st in ("prepared", "attach"):
        ts = [v[st] for v in at.values() if st in v]
        if ts:
            out[f"last_{st}_s"] = round(max(ts), 3)
    return out


def percentile(vals, q):
    return float(np.percentile(np.asarray(vals, np.float64), q)) if len(vals) else float("nan")


def summarize(requests, results, makespan):
    ok = [results[r.rid] for r in requests if r.rid in results and results[r.rid].status == "ok"]
    fwd_cells = sum(r.cells * r.steps for r in requests if r.rid in results and results[r.rid].status == "ok")
    lat = [x.latency for x in ok]
    return dict(requests=len(requests), completed=len(ok), makespan_s=makespan,
                requests_per_s=len(ok) / makespan if makespan > 0 else float("nan"),
                fwd_mcell_updates_per_s=fwd_cells / makespan / 1e6 if makespan > 0 else float("nan"),
                p50_latency_s=percentile(lat, 50), p99_latency_s=percentile(lat, 99))


def main(argv=None):
    ap = argparse.ArgumentParser(description="Continuous-batching FDTD serving run against serial plain solves")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--rate", type=float, default=None, help="Poisson arrival rate per second (default: all at t=0)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--step-scale", type=float, default=1.0, help="scale the [500, 4000] step range")
    ap.add_argument("--chunk", type=int, default=16)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--reserve-gib", type=float, default=16.0, help="HBM kept out of the admission budget")
    ap.add_argument("--budget-gib", type=float, default=None, help="admission budget (default: device HBM - reserve)")
    ap.add_argument("--serial", action="store_true", help="also time serial plain solves and compare results")
    ap.add_argument("--serial-max", type=int, default=None, help="time serial solves on the first N requests only")
    ap.add_argument("--sizes", type=str, default=None,
                    help="grid sizes to draw from, e.g. 64x64x64,96x96x96 (default: the five serving sizes)")
    ap.add_argument("--grad-frac", type=float, default=0.3)
    ap.add_argument("--policy", choices=("sjf", "fifo"), default="sjf")
    ap.add_argument("--max-active-cells", type=int, default=48_000_000)
    ap.add_argument("--prep-ahead", type=int, default=8)
    ap.add_argument("--prep-workers", type=int, default=2)
    ap.add_argument("--inflight", type=int, default=2, help="chunks kept queued on the device")
    ap.add_argument("--solo-cells", type=int, default=None,
                    help="requests of at least this many cells run alone beside the batch (default: batch all)")
    ap.add_argument("--out", type=str, default=None)
    ap.add_argument("--progress", type=int, default=32, help="print a line every N finished requests (0: quiet)")
    a = ap.parse_args(argv)
    import torch

    sizes = SIZES if not a.sizes else tuple(tuple(int(v) for v in z.split("x")) for z in a.sizes.split(","))
    reqs = make_requests(a.requests, a.rate, a.seed, sizes=sizes, grad_frac=a.grad_frac, step_scale=a.step_scale)
    total = torch.cuda.get_device_properties(0).total_memory
    cap = int(a.budget_gib * 2 ** 30) if a.budget_gib else int(total - a.reserve_gib * 2 ** 30)
    print(f"[serve] {len(reqs)} requests, budget {cap / 2**30:.1f} GiB, chunk {a.chunk}, "
          f"graph {not a.no_graph}", flush=True)
    # warm up: compile-free engine, but the first build imports and initialises the host modules
    warm = Request(-1, (24, 24, 24), 8, 1, "dipole", True, 0.0, 4)
    run_plain(warm)
    be = GpuBackend(use_graph=not a.no_graph, timing=True)
    srv = Server(be, HbmBudget(cap), chunk=a.chunk, prep_workers=a.prep_workers, max_inflight=a.inflight,
                 policy=a.policy, max_active_cells=a.max_active_cells, prep_ahead=a.prep_ahead,
                 progress=a.progress, solo_cells=a.solo_cells)
    srv.run([warm])
    be.chunk_events.clear()
    be.prep_s.clear()
    be.reverse_s.clear()
    res, stats = srv.run(reqs)
    busy = be.batch_busy_s()
    be.close()
    out = dict(served=summarize(reqs, res, stats.makespan), chunks=stats.chunks, peak_slots=stats.peak_slots,
               peak_budget_gib=stats.peak_bytes / 2 ** 30, rejected=stats.rejected, evicted=stats.evicted,
               batch_busy_s=busy, prep_s_total=sum(be.prep_s), reverse_s_total=sum(be.reverse_s),
               stages=stage_summary(stats),
               rate=a.rate, step_scale=a.step_scale, seed=a.seed, chunk=a.chunk, policy=a.policy,
               max_active_cells=a.max_active_cells, sizes=[list(z) for z in sizes], grad_frac=a.grad_frac)
    print(f"[serve] served: {json.dumps(out['served'])}", flush=True)
    if a.serial:
        sub = reqs[: a.serial_max] if a.serial_max else reqs
        setup = {}
        times, sres = run_serial(sub, setup, progress=a.progress)
        serial_total = sum(times.values())
        lat = fifo_latencies(sub, times)
        mism = 0
        for r in sub:
            f, g = sres[r.rid]
            x = res.get(r.rid)
            if x is None or x.status != "ok" or not np.array_equal(f, x.flux) or (
                    g is not None and not np.array_equal(g, x.grad)):
                mism += 1
        ser = dict(requests=len(sub), total_s=serial_total, setup_s_total=sum(setup.values()),
                   requests_per_s=len(sub) / serial_total,
                   fwd_mcell_updates_per_s=sum(r.cells * r.steps for r in sub) / serial_total / 1e6,
                   p50_latency_s=percentile(list(lat.values()), 50), p99_latency_s=percentile(list(lat.values()), 99),
                   mismatches_vs_served=mism)
        out["serial"] = ser
        if a.serial_max is None:
            out["speedup_requests_per_s"] = out["served"]["requests_per_s"] / ser["requests_per_s"]
        print(f"[serve] serial: {json.dumps(ser)}", flush=True)
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
