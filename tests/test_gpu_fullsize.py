"""GPU parity at the benchmark configurations' own sizes, against oracle fixtures made on the CPU
by tests/golden/make_goldens.py (FP32 run: the checker; FP64 run: the FP32 method's own error,
printed beside it).

  c1  64^3 dipole + dielectric cube, CPML 10, 1000 steps: energy trace, final Ez, final E / H
  c2  400x400x200 plane wave + seeded design slab, CPML 12, 2000 steps: flux trace, 3-frequency
      phasor plane, final E / H
  c3  200x200x100 dipole + seeded design region, CPML 10, 3000 steps: the reversible gradient of
      the summed output flux against the FP32 oracle's reversible adjoint (whole design region)
      and against FP64 central differences at 16 seeded voxels
  c4  256^3 dipole, diagonal-eps Lorentz slab + Drude cube, CPML 12, 4000 steps: 5-frequency
      phasor plane, final E / H

Full fields do not fit a fixture: the final E and H are compared on 4096 seeded samples (rel-L2)
and through the L2 norm of every (component, x-plane), the phasor planes on seeded samples and
through the norm of every (frequency, component, row). Every comparison with the FP32 oracle is
held to rel-L2 < 1e-5 (the gradient: 1e-3, the north-star tolerance for gradients).
"""

import os

import numpy as np
import pytest

from fdtdx_amd import scenarios as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    p = os.path.join(GOLDEN, name)
    if not os.path.exists(p):
        pytest.skip(f"fixture {name} not generated (python tests/golden/make_goldens.py)")
    return np.load(p)


def rel(a, b):
    a = np.asarray(a).astype(np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b).astype(np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _plane_norms(F):
    import torch

    return torch.sqrt(torch.sum(F.double() ** 2, dim=(2, 3))).cpu().numpy()


def _phasor_row_norms(ph):
    import torch

    a = torch.abs(ph[0].to(torch.complex128))
    a = a[:, :, 0] if a.shape[2] == 1 else a
    return torch.sqrt(torch.sum(a * a, dim=-1)).cpu().numpy()


def _run(spec):
    import torch

    import fdtdx_amd as fx
    import scenarios

    objects, arrays, params, config = scenarios.build(spec)
    print(f"case shape={spec['shape']} T={spec['T']}", flush=True)
    _, arrays = fx.run_fdtd(arrays, objects, config)
    torch.cuda.synchronize()
    return arrays


def _check_fields(arrays, g, tol=1e-5):
    E, H = arrays.fields.E, arrays.fields.H
    idx = g["sample_index"]
    for name, F in (("E", E), ("H", H)):
        s = F.reshape(-1)[__import__("torch").as_tensor(idx, device=F.device)].double().cpu().numpy()
        r32 = rel(s, g[f"f32_{name}_samples"])
        r64 = rel(s, g[f"f64_{name}_samples"])
        pn = _plane_norms(F)
        rp = rel(pn, g[f"f32_{name}_plane_norms"])
        print(f"  {name}: samples rel-L2 vs FP32 oracle {r32:.3e} (vs FP64 {r64:.3e}); plane norms {rp:.3e}",
              flush=True)
        assert r32 < tol and rp < tol


def test_c1_fullsize_matches_oracle():
    g = _load("c1.npz")
    arrays = _run(W.scenario_c1())
    _check_fields(arrays, g)
    en = arrays.detector_states["energy"]["energy"].double().cpu().numpy()
    print(f"  energy rel-L2 vs FP32 {rel(en, g['f32_energy']):.3e} (vs FP64 {rel(en, g['f64_energy']):.3e})",
          flush=True)
    assert rel(en, g["f32_energy"]) < 1e-5
    ez = arrays.detector_states["ez"]["fields"].double().cpu().numpy().reshape(-1)
    assert rel(ez[g["ez_index"]], g["f32_ez_samples"]) < 1e-5
    assert abs(np.linalg.norm(ez) - float(g["f32_ez_norm"])) <= 1e-5 * float(g["f32_ez_norm"])


def test_c2_fullsize_matches_oracle():
    g = _load("c2.npz")
    arrays = _run(W.scenario_c2())
    _check_fields(arrays, g)
    flux = arrays.detector_states["flux"]["poynting_flux"].double().cpu().numpy()
    print(f"  flux rel-L2 vs FP32 {rel(flux, g['f32_flux']):.3e} (vs FP64 {rel(flux, g['f64_flux']):.3e})", flush=True)
    assert rel(flux, g["f32_flux"]) < 1e-5
    ph = arrays.detector_states["phasor"]["phasor"]
    s = ph.reshape(-1)[__import__("torch").as_tensor(g["phasor_index"], device=ph.device)].cpu().numpy()
    print(f"  phasor samples rel-L2 vs FP32 {rel(s, g['f32_phasor_samples']):.3e} "
          f"(vs FP64 {rel(s, g['f64_phasor_samples']):.3e})", flush=True)
    assert rel(s, g["f32_phasor_samples"]) < 1e-5
    assert rel(_phasor_row_norms(ph), g["f32_phasor_row_norms"]) < 1e-5


def test_c4_fullsize_matches_oracle():
    g = _load("c4.npz")
    arrays = _run(W.scenario_c4())
    _check_fields(arrays, g)
    ph = arrays.detector_states["phasor"]["phasor"]
    s = ph.reshape(-1)[__import__("torch").as_tensor(g["phasor_index"], device=ph.device)].cpu().numpy()
    print(f"  phasor samples rel-L2 vs FP32 {rel(s, g['f32_phasor_samples']):.3e} "
          f"(vs FP64 {rel(s, g['f64_phasor_samples']):.3e})", flush=True)
    assert rel(s, g["f32_phasor_samples"]) < 1e-5
    assert rel(_phasor_row_norms(ph), g["f32_phasor_row_norms"]) < 1e-5


def test_c3_fullsize_gradient_matches_oracle_and_fd():
    import torch

    import fdtdx_amd as fx

    gg = _load("c3_grad.npz")
    gf = _load("c3_fd.npz")
    spec = W.scenario_c3()
    objects, arrays, params, config = W.build(spec, gradient=True)
    rho = torch.tensor(np.asarray(spec["design"]["rho"], np.float32), device="cuda", requires_grad=True)
    arrays2, objects2, _ = fx.apply_params(arrays, objects, {"design": rho})
    print(f"case c3 shape={spec['shape']} T={spec['T']}", flush=True)
    _, arrays3 = fx.run_fdtd(arrays2, objects2, config)
    flux = arrays3.detector_states["flux"]["poynting_flux"]
    flux.sum().backward()
    torch.cuda.synchronize()
    g = rho.grad.double().cpu().numpy()
    f = flux.double().detach().cpu().numpy()
    print(f"  flux rel-L2 vs FP32 {rel(f, gg['f32_flux']):.3e} (vs FP64 {rel(f, gf['f64_flux']):.3e})", flush=True)
    assert rel(f, gg["f32_flux"]) < 1e-5
    rg = rel(g, gg["f32_grad_rho"])
    print(f"  gradient rel-L2 vs the FP32 oracle's reversible adjoint {rg:.3e}", flush=True)
    assert rg < 1e-3
    worst = 0.0
    for idx, fd in zip(gf["fd_voxels"], gf["f64_fd"]):
        e = abs(g[tuple(idx)] - fd) / abs(fd)
        worst = max(worst, e)
        print(f"  voxel {tuple(int(v) for v in idx)} gpu {g[tuple(idx)]:.6e} fd64 {fd:.6e} rel {e:.2e}", flush=True)
    assert worst < 1e-3
