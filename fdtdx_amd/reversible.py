"""Reversible gradients through the HIP engine (reference: src/fdtdx/fdtd/fdtd.py:39-418).

Forward: reset, run all steps recording the PML interface layers (update.py:1140-1178).
Backward: fdtdx_reverse_adjoint walks the steps in reverse, restores the interfaces, reconstructs
the fields with frozen CPML psi (backward.py:62-135) and applies the transposed step, which is
what the reference's jax.vjp of forward() computes (fdtd.py:215-251). The gradient returned is
for inv_permittivities only, as in the reference (fdtd.py:324-333).

Slices (GradientConfig.num_checkpoints_reversible = k, fdtd.py:106-121, 295-313): the run is cut
at round(i T / (k+1)); the forward keeps the full field state (E, H, psi) at every interior cut,
and the reverse resets its reconstructed state to it on reaching the cut, so each slice
reconstructs with the psi of its own end frozen.

Record budget (GradientConfig.max_record_bytes, not in the reference): when the interface record
(every step, or the slots a Recorder's time filter keeps, in its formats) does not fit, the
forward keeps the full state at the step of each record chunk's first slot instead, and the
backward regenerates each chunk's slots just before reversing it, by re-running the forward
steps that write them from that checkpoint into a one-chunk record buffer (with scratch fields
and detector states, so the reverse state and the outputs are untouched). The reverse pass then
sees the same record values and the same resets, so the gradient is bit-identical.
"""

from __future__ import annotations

import numpy as np

from . import engine


def slice_boundaries(T: int, num_checkpoints: int) -> list[int]:
    """fdtd.py:20-36: round(i * T / k) for k = num_checkpoints + 1 slices."""
    k = num_checkpoints + 1
    return [round(i * T / k) for i in range(k + 1)]


class _Addr:
    """A device address standing in for a tensor in engine.make_buffers."""

    def __init__(self, addr: int):
        self.addr = int(addr)

    def data_ptr(self):
        return self.addr


class Schedule:
    """Where the forward keeps state checkpoints and how the record is held.

    bounds: slice boundaries (reset points are the interior ones).
    stored: the whole record is kept; otherwise chunk starts (a superset of the slice boundaries)
    split the reverse pass into record chunks whose slots fit `chunk_steps` slots, each
    regenerated from the checkpoint at the step of its first slot.

    `record` is the bytes of one step's record when every step is kept in its own slot, or the
    plan (slot_of, slot_arr, slot_step, record_slot_bytes) when a time filter keeps only some
    steps: a reversed step n reads the slot at or before n and, when n itself is not kept, the
    next one (LinearReconstructEveryK, time_filter.py:233-257)."""

    def __init__(self, T: int, num_checkpoints: int, record, budget: int | None):
        self.T = T
        if num_checkpoints > 0 and num_checkpoints + 1 > T:  # fdtd.py:115-119
            raise Exception("num_checkpoints_reversible must be <= time_steps_total - 1 "
                            f"(got num_checkpoints_reversible={num_checkpoints}, time_steps_total={T})")
        self.bounds = slice_boundaries(T, num_checkpoints)
        if isinstance(record, (int, np.integer)):
            self.slot_bytes = int(record)
            self.slot_of = np.arange(T)
            self.arr = np.arange(T)
            self.slot_step = np.arange(T)
        else:
            self.slot_bytes = int(record.record_slot_bytes)
            self.slot_of, self.arr, self.slot_step = record.slot_of, record.slot_arr, record.slot_step
        nslots = len(self.slot_step)
        total = nslots * self.slot_bytes
        self.stored = budget is None or total <= budget or self.slot_bytes == 0 or T == 0
        if self.stored:
            self.chunk_steps = nslots
            self.starts = list(self.bounds[:-1])
        else:
            cap = max(2, int(budget // self.slot_bytes))
            self.chunk_steps = cap
            starts = set()
            for a, b in zip(self.bounds[:-1], self.bounds[1:]):
                c = a
                while c < b:
                    d = c + 1
                    while d < b and self._last(d) - int(self.arr[c]) + 1 <= cap:
                        d += 1
                    starts.add(c)
                    c = d
            self.starts = sorted(starts)
        # state kept by the forward at these step indices (0 is the zero initial state)
        runs = set() if self.stored else {self.run_range(c, d)[0] for a, b in zip(self.bounds[:-1], self.bounds[1:])
                                          for c, d in self.chunks(a, b)}
        self.keep = sorted((set(self.bounds[1:-1]) | runs) - {0})

    def _last(self, n: int) -> int:
        """The last slot reversed step n reads."""
        return int(self.slot_of[n]) if self.slot_of[n] >= 0 else int(self.arr[n]) + 1

    def run_range(self, c: int, d: int):
        """Forward steps [r0, r1) that regenerate the slots chunk [c, d) reads, and the first of
        those slots."""
        lo = int(self.arr[c])
        hi = self._last(d - 1)
        return int(self.slot_step[lo]), int(self.slot_step[hi]) + 1, lo

    def chunks(self, a: int, b: int):
        """Record chunks of slice [a, b) from the last to the first."""
        cs = [c for c in self.starts if a <= c < b]
        ends = cs[1:] + [b]
        return list(zip(cs, ends))[::-1]


def _state(solve, E, H, psi):
    return E.clone(), H.clone(), psi.clone()


def run_forward(solve, inv_eps, sched: Schedule):
    """Forward steps 0 .. T-1 from a reset state, keeping the schedule's checkpoints; the record
    is written only when the schedule keeps it whole. Returns {step: (E, H, psi)}."""
    a = solve.arrays
    ckpts = {}
    t = 0
    for c in sched.keep + [sched.T]:
        if c > t:
            solve.forward(t, c, record=sched.stored, inv_eps=inv_eps)
        if c < sched.T:
            ckpts[c] = _state(solve, a.fields.E, a.fields.H, a.psi)
        t = c
    return ckpts


def reverse_gradient(solve, inv_eps, E, H, cotangents, T, stream=None, sched: Schedule | None = None,
                     ckpts: dict | None = None, psi=None):
    """Run fdtdx_reverse_adjoint from the final fields (E, H, overwritten by the reconstruction)
    and the solve's psi and interface record, for detector cotangents `cotangents` (one per
    state entry the engine writes, in `solve.state_keys` order, None for none); returns
    dL/d inv_eps. Work is enqueued on `stream` (a
    torch stream; default the current one). `sched` and `ckpts` come from run_forward; without
    them the run is one slice with the record kept whole."""
    import torch

    a = solve.arrays
    if sched is None:
        sched = Schedule(T, 0, 0, None)
        ckpts = {}
    if len(cotangents) != len(solve.state_keys):
        raise ValueError(f"expected {len(solve.state_keys)} detector cotangents, got {len(cotangents)}")
    # each engine detector of an accumulating group reads its entry's cotangent (that of a sum)
    by_key = dict(zip(solve.state_keys, cotangents))
    cotangents = [by_key[k] for k in solve.det_keys]
    psi_w = (a.psi if psi is None else psi).clone()
    lamE = torch.zeros_like(E)
    lamH = torch.zeros_like(H)
    lam_psi = torch.zeros_like(a.psi)
    grad = torch.zeros(inv_eps.shape, dtype=torch.float32, device=E.device)
    cots = []
    for g in cotangents:
        if g is None:
            cots.append(None)
        elif g.is_complex():
            cots.append(torch.view_as_real(g.contiguous()).contiguous())
        else:
            cots.append(g.contiguous().float())
    ie = inv_eps.detach().contiguous()
    im = None if np.isscalar(a.inv_permeabilities) else a.inv_permeabilities.detach().contiguous()
    handle = engine.current_stream_handle() if stream is None else engine.c_void_p(stream.cuda_stream)
    plan = solve.plan
    slot_bytes = plan.record_slot_bytes
    if not sched.stored:
        # scratch state for regenerating a chunk's record, and its one-chunk record buffer
        E2, H2, psi2 = torch.empty_like(E), torch.empty_like(H), torch.empty_like(a.psi)
        scratch = {k: torch.empty_like(t) for k, t in zip(solve.state_keys, solve.state_tensors())}
        det_scratch = [scratch[k] for k in solve.det_keys]
        chunk_rec = torch.empty(max(1, sched.chunk_steps * slot_bytes), dtype=torch.uint8, device=E.device)
    for s in reversed(range(len(sched.bounds) - 1)):
        lo, hi = sched.bounds[s], sched.bounds[s + 1]
        if hi < T:  # reset the reconstructed state at the slice end (fdtd.py:295-313)
            cE, cH, cP = ckpts[hi]
            E.copy_(cE)
            H.copy_(cH)
            psi_w.copy_(cP)
        if sched.stored:
            b, keep = engine.make_buffers(E, H, ie, im, psi_w, a.ade_state, solve.det_tensors(), solve.record,
                                          solve.workspace)
            engine.reverse_adjoint(plan, b, lamE, lamH, lam_psi, grad, cots, hi, lo, handle)
            continue
        for c, d in sched.chunks(lo, hi):
            r0, r1, first = sched.run_range(c, d)
            if r0 == 0:
                E2.zero_()
                H2.zero_()
                psi2.zero_()
            else:
                cE, cH, cP = ckpts[r0]
                E2.copy_(cE)
                H2.copy_(cH)
                psi2.copy_(cP)
            base = _Addr(chunk_rec.data_ptr() - first * slot_bytes)  # slot i at base + i * slot_bytes
            assert base.addr > 0
            bf, keep_f = engine.make_buffers(E2, H2, ie, im, psi2, a.ade_state, det_scratch, base, solve.workspace)
            engine.forward(plan, bf, r0, r1, True, handle)
            b, keep = engine.make_buffers(E, H, ie, im, psi_w, a.ade_state, solve.det_tensors(), base,
                                          solve.workspace)
            engine.reverse_adjoint(plan, b, lamE, lamH, lam_psi, grad, cots, d, c, handle)
    return grad


def _autograd_fn():
    import torch

    class ReversibleFDTD(torch.autograd.Function):
        @staticmethod
        def forward(ctx, inv_eps, solve, sched):
            a = solve.arrays
            a.reset()
            ckpts = run_forward(solve, inv_eps, sched)
            ctx.solve = solve
            ctx.sched = sched
            ctx.ckpts = ckpts
            ctx.final = (a.fields.E.clone(), a.fields.H.clone(), a.psi.clone())
            ctx.save_for_backward(inv_eps)
            outs = tuple(t.clone() for t in solve.state_tensors())
            return outs

        @staticmethod
        def backward(ctx, *grads):
            (inv_eps,) = ctx.saved_tensors
            E, H, psi = ctx.final
            grad = reverse_gradient(ctx.solve, inv_eps, E.clone(), H.clone(), grads, ctx.sched.T, sched=ctx.sched,
                                    ckpts=ctx.ckpts, psi=psi)
            return grad, None, None

    return ReversibleFDTD


_FN = None


def _stream_handle():
    return engine.current_stream_handle()


def record_budget(config) -> int:
    import torch

    gc = config.gradient_config
    if gc is not None and getattr(gc, "max_record_bytes", None) is not None:
        return int(gc.max_record_bytes)
    free, _ = torch.cuda.mem_get_info()
    return int(free // 2)


def reversible_run(arrays, objects, config, progress_callback=None):
    global _FN
    from .fdtd import _solve_for

    T = config.time_steps_total
    solve = _solve_for(objects, config, arrays, record_steps=0)
    ie = arrays.inv_permittivities
    if not getattr(ie, "requires_grad", False):
        # the forward keeps the interface record when it fits the budget, so that full_backward can
        # replay it (fdtd.py:175-212 records whenever a recorder is configured)
        keep = solve.record_bytes(T) <= record_budget(config)
        solve.ensure_record(T if keep else 0)
        arrays.reset()
        solve.forward(0, T, record=keep)
        return T, arrays
    gc = config.gradient_config
    k = 0 if gc is None else int(gc.num_checkpoints_reversible)
    sched = Schedule(T, k, solve.plan, record_budget(config))
    solve.ensure_record(T if sched.stored else 0)
    if _FN is None:
        _FN = _autograd_fn()
    outs = _FN.apply(ie, solve, sched)
    new_states = {nm: dict(st) for nm, st in arrays.detector_states.items()}
    for (nm, key), t in zip(solve.state_keys, outs):
        new_states[nm][key] = t
    arrays.detector_states = new_states
    return T, arrays
