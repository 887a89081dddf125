"""fdtdx_forward_noalloc (the forward with its workspace in the library's static device scratch)
against fdtdx_forward through the C ABI: fields, psi and every detector state bit for bit, on two
calls in a row, the second on a smaller grid, so the second call runs on scratch the first one
left behind; and a plan whose forward needs more than the scratch is rejected."""

import numpy as np
import pytest

from fdtdx_amd import scenarios as W

pytestmark = pytest.mark.gpu


def _run(spec, noalloc):
    import torch

    from fdtdx_amd import engine
    from fdtdx_amd import fdtd as F
    import scenarios

    objects, arrays, params, config = scenarios.build(spec)
    s = F.Solve(objects, config, arrays)
    arrays.reset()
    b = s.buffers()
    stream = engine.current_stream_handle()
    print(f"case shape={spec['shape']} T={spec['T']} noalloc={noalloc}", flush=True)
    if noalloc:
        engine.forward_noalloc(s.plan, b, 0, spec["T"], False, stream)
    else:
        engine.forward(s.plan, b, 0, spec["T"], False, stream)
    torch.cuda.synchronize()
    return [arrays.fields.E.cpu().numpy(), arrays.fields.H.cpu().numpy(), arrays.psi.cpu().numpy()] + [
        t.cpu().numpy() for t in s.det_tensors()]


def test_noalloc_forward_equals_forward_on_two_calls():
    specs = [W.scenario_c2(T=60, shape=(48, 40, 32), pml=6, slab=(12, 12, 8)),  # flux + phasor planes
             W.scenario_c1(T=50, n=24, pml=5, cube=6)]                           # smaller: energy + snapshot
    ref = [_run(sp, False) for sp in specs]
    got = [_run(sp, True) for sp in specs]
    for r, g in zip(ref, got):
        assert np.abs(r[0]).max() > 0
        assert len(r) == len(g)
        for a, b in zip(g, r):
            assert np.array_equal(a, b)


def test_noalloc_forward_rejects_a_workspace_beyond_the_scratch():
    from fdtdx_amd import engine

    spec = W.scenario_c1(T=4, n=128, pml=4, cube=8)  # full-volume energy + field detectors: 2M-cell snapshot
    with pytest.raises(engine.EngineError, match="static scratch"):
        _run(spec, True)
