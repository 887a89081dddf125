"""Serving on the GPU: every request's result (flux trace, and the design gradient for gradient
requests) is bit-identical when served alone, in a mix with staggered admission, with and without
hipGraph chunks, and through the plain one-solve path. Health flags evict a solve whose flux
blows up or turns non-finite without disturbing the others."""

import numpy as np
import pytest

from fdtdx_amd import serving as S

pytestmark = pytest.mark.gpu

SHAPES = ((24, 24, 24), (32, 28, 20), (40, 32, 24), (28, 36, 30))


def _mix(n=8, seed=5):
    """Two long solves (one with a gradient) that are still running when the short ones arrive."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shape = SHAPES[i % len(SHAPES)]
        steps = (3000, 2500)[i] if i < 2 else int(rng.integers(30, 110))
        out.append(S.Request(i, shape, steps, 1000 + i, "dipole" if i % 2 == 0 else "plane", i % 3 == 1,
                             arrival=0.002 * i, pml=4))
    return out


def _serve(reqs, chunk=7, graph=True, max_slots=64, blowup=0.0, poison=None, solo_cells=None):
    be = S.GpuBackend(use_graph=graph, blowup_limit=blowup)
    if poison is not None:
        prep = be.prepare

        def prepare(req):
            job = prep(req)
            if req.rid in poison:  # a NaN permittivity cell next to the flux plane
                nx, ny, nz = req.shape
                job.solve.arrays.inv_permittivities.detach()[:, nx - req.pml - 5, ny // 2, nz // 2] = float("nan")
                import torch

                torch.cuda.synchronize()
            return job

        be.prepare = prepare
    try:
        srv = S.Server(be, S.HbmBudget(8 << 30), max_slots=max_slots, chunk=chunk, solo_cells=solo_cells)
        res, stats = srv.run(reqs)
    finally:
        be.close()
    return res, stats


@pytest.fixture(scope="module")
def plain():
    reqs = _mix()
    print("[serving] plain solves of", [(r.rid, r.shape, r.steps, r.source, r.gradient) for r in reqs], flush=True)
    return reqs, {r.rid: S.run_plain(r) for r in reqs}


def _same(res, ref, rids):
    for rid in rids:
        x = res[rid]
        f, g = ref[rid]
        assert x.status == "ok", (rid, x.reason)
        assert np.array_equal(x.flux, f), rid
        if g is None:
            assert x.grad is None
        else:
            assert np.array_equal(x.grad, g), rid


def test_mix_equals_plain_solves(plain):
    reqs, ref = plain
    assert any(np.abs(ref[r.rid][0]).max() > 0 for r in reqs)
    res, stats = _serve(reqs, chunk=7, graph=True)
    print(f"[serving] mix: chunks={stats.chunks} peak_slots={stats.peak_slots}", flush=True)
    assert stats.peak_slots >= 2
    _same(res, ref, [r.rid for r in reqs])


def test_mix_with_solo_requests_equals_plain_solves(plain):
    """Requests of the two largest shapes run alone (the single-solve four-cells-per-lane sweeps, and
    for a gradient its adjoint) on worker streams beside the batch: every result still equals the
    plain solve's bit for bit."""
    reqs, ref = plain
    big = 28 * 36 * 30  # the two largest shapes, a gradient request among them
    res, stats = _serve(reqs, chunk=7, graph=True, solo_cells=big)
    solo = {rid for _, rid, stage in stats.events if stage == "solo"}
    print(f"[serving] solo mix: solo={sorted(solo)} chunks={stats.chunks}", flush=True)
    assert solo == {r.rid for r in reqs if r.cells >= big}
    assert any(reqs[rid].gradient for rid in solo)
    _same(res, ref, [r.rid for r in reqs])


def test_reverse_runs_while_batch_stream_captures(plain):
    """A gradient request's adjoint may start while the loop thread has a chunk graph in capture
    on the batch stream (a new grid key appears whenever slot membership changes the tile
    count). The backend opens such a capture itself here, on the batch stream and in the
    engine's relaxed mode, and starts the reverse inside it: the adjoint must not touch the
    capturing stream or its events, and its result must equal the plain solve's."""
    import ctypes

    reqs, ref = plain
    r = next(q for q in reqs if q.gradient and q.steps < 400)
    print("[serving] reverse during capture of rid", r.rid, r.shape, r.steps, flush=True)
    be = S.GpuBackend(use_graph=True)
    try:
        job = be.prepare(r)
        be.attach(0, job)
        token = None
        while be.step_index(0, job) < r.steps:
            token = be.advance(7)
        be.wait(token)
        be.detach(0, job)
        hip = ctypes.CDLL("libamdhip64.so")
        stream = ctypes.c_void_p(be.stream.cuda_stream)
        assert hip.hipStreamBeginCapture(stream, 2) == 0  # hipStreamCaptureModeRelaxed
        graph = ctypes.c_void_p()
        try:
            flux, grad = be.start_reverse(job, token).result(timeout=100)
        finally:
            rc = hip.hipStreamEndCapture(stream, ctypes.byref(graph))
            if graph.value:
                hip.hipGraphDestroy(graph)
        assert rc == 0
        f, g = ref[r.rid]
        assert np.array_equal(flux, f)
        assert np.array_equal(grad, g)
    finally:
        be.close()


def test_alone_equals_plain_solves(plain):
    reqs, ref = plain
    for r in reqs[:3]:
        res, _ = _serve([r], chunk=11, graph=True)
        _same(res, ref, [r.rid])


def test_mix_without_graphs_and_with_few_slots(plain):
    reqs, ref = plain
    res, stats = _serve(reqs, chunk=5, graph=False, max_slots=3)
    assert stats.peak_slots <= 3
    _same(res, ref, [r.rid for r in reqs])


def test_nonfinite_solve_is_evicted(plain):
    reqs, ref = plain
    fwd = [r for r in reqs if not r.gradient][:4]
    bad = fwd[1].rid
    res, stats = _serve(fwd, chunk=6, poison={bad})
    assert res[bad].status == "failed" and "non-finite" in res[bad].reason
    assert stats.evicted == 1
    _same(res, ref, [r.rid for r in fwd if r.rid != bad])


def test_blowup_limit_flags_every_solve(plain):
    reqs, _ = plain
    fwd = sorted([r for r in reqs if not r.gradient], key=lambda r: -r.steps)[:2]
    res, stats = _serve(fwd, chunk=6, blowup=1e-30)
    assert all(res[r.rid].status == "failed" and "blow-up" in res[r.rid].reason for r in fwd)
