"""Reversible gradients through the HIP engine (reference: src/fdtdx/fdtd/fdtd.py:39-418).

Forward: reset, run all steps recording the PML interface layers (update.py:1140-1178).
Backward: fdtdx_reverse_adjoint walks the steps in reverse, restores the interfaces, reconstructs
the fields with frozen CPML psi (backward.py:62-135) and applies the transposed step, which is
what the reference's jax.vjp of forward() computes (fdtd.py:215-251). The gradient returned is
for inv_permittivities only, as in the reference (fdtd.py:324-333).
"""

from __future__ import annotations

import numpy as np

from . import engine


def _autograd_fn():
    import torch

    class ReversibleFDTD(torch.autograd.Function):
        @staticmethod
        def forward(ctx, inv_eps, solve, T):
            a = solve.arrays
            a.reset()
            solve.forward(0, T, record=True, inv_eps=inv_eps)
            ctx.solve = solve
            ctx.T = T
            ctx.final = (a.fields.E.clone(), a.fields.H.clone())
            ctx.save_for_backward(inv_eps)
            outs = tuple(t.clone() for t in solve.det_tensors())
            return outs

        @staticmethod
        def backward(ctx, *grads):
            (inv_eps,) = ctx.saved_tensors
            E, H = ctx.final
            grad = reverse_gradient(ctx.solve, inv_eps, E.clone(), H.clone(), grads, ctx.T)
            return grad, None, None

    return ReversibleFDTD


def reverse_gradient(solve, inv_eps, E, H, cotangents, T, stream=None):
    """Run fdtdx_reverse_adjoint from the final fields (E, H, overwritten by the reconstruction)
    and the solve's psi and interface record, for detector cotangents `cotangents` (one per
    forward detector, None for none); returns dL/d inv_eps. Work is enqueued on `stream` (a
    torch stream; default the current one)."""
    import torch

    a = solve.arrays
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
    b, keep = engine.make_buffers(E, H, ie, im, a.psi, a.ade_state, solve.det_tensors(), solve.record,
                                  solve.workspace)
    handle = engine.current_stream_handle() if stream is None else engine.c_void_p(stream.cuda_stream)
    engine.reverse_adjoint(solve.plan, b, lamE, lamH, lam_psi, grad, cots, T, 0, handle)
    return grad


_FN = None


def reversible_run(arrays, objects, config, progress_callback=None):
    global _FN
    from .fdtd import _solve_for

    T = config.time_steps_total
    solve = _solve_for(objects, config, arrays, record_steps=T)
    ie = arrays.inv_permittivities
    if not getattr(ie, "requires_grad", False):
        arrays.reset()
        solve.forward(0, T, record=False)
        return T, arrays
    if _FN is None:
        _FN = _autograd_fn()
    outs = _FN.apply(ie, solve, T)
    names = [d.name for d in objects.forward_detectors]
    new_states = dict(arrays.detector_states)
    for nm, t in zip(names, outs):
        key = next(iter(arrays.detector_states[nm].keys()))
        new_states[nm] = {key: t}
    arrays.detector_states = new_states
    return T, arrays
