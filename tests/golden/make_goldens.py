"""Write the full-size oracle fixtures of the four benchmark configurations (run in the build
container, on the CPU; the GPU tests in tests/test_gpu_fullsize.py read the .npz files).

    python tests/golden/make_goldens.py c1 c2 c4 c3fd   # C oracle (oracle/fast.py), all cores
    python tests/golden/make_goldens.py c3grad          # NumPy oracle's reversible adjoint, FP32

Each configuration is run in FP32 (the checker) and in FP64 (the error of the FP32 method,
reported beside it). Full fields do not fit a small fixture, so a fixture holds what pins them
without them: the detector outputs whole where they are small (energy and flux traces), seeded
samples of the final E and H and of large detector states, the L2 norm of every (component,
x-plane) of the final fields, and of every (frequency, component, row) of phasor planes.

  c1      64^3, CPML 10, 16^3 cube eps 2.25, z-dipole, 1000 steps: energy trace, final Ez
  c2      400x400x200, CPML 12, seeded design slab, plane wave, 2000 steps: flux trace, 3-frequency
          phasor plane
  c3grad  200x200x100, CPML 10, dipole, 3000 steps: d(sum of output flux)/d rho over the
          100x100x20 design region by the reversible adjoint (whole, float32), and the flux trace
  c3fd    the same case: FP64 central differences (h = 1e-3) of the summed flux at 16 seeded
          voxels, and the FP64 flux trace
  c4      256^3, CPML 12, diagonal-eps Lorentz slab and Drude cube, dipole, 4000 steps:
          5-frequency phasor plane
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fdtdx_amd import scenarios as W  # noqa: E402  (scenario specs are plain dicts)
from oracle import fdtd_oracle as O  # noqa: E402
from oracle import fast as F  # noqa: E402
from oracle import scenes as S  # noqa: E402

SAMPLE_SEED = 20261203
NSAMPLES = 4096
FD_SEED = 16
FD_VOXELS = 16
FD_H = 1e-3


def sample_index(size, seed=SAMPLE_SEED, n=NSAMPLES):
    return np.random.default_rng(seed).integers(0, size, n).astype(np.int64)


def plane_norms(F_):
    """L2 norm of every (component, x-plane) of a (3, nx, ny, nz) field, in float64."""
    return np.sqrt(np.sum(np.square(F_.astype(np.float64)), axis=(2, 3)))


def phasor_row_norms(ph):
    """L2 norm of every (frequency, component, row along the plane's first axis) of a phasor state
    (1, F, C, *plane) with one singleton plane axis."""
    a = np.abs(ph[0].astype(np.complex128))
    a = a.reshape(a.shape[0], a.shape[1], -1, a.shape[-1]) if a.shape[2] != 1 else a[:, :, 0]
    return np.sqrt(np.sum(a * a, axis=-1))


def field_record(prefix, st, out):
    idx = sample_index(st.E.size)
    out["sample_index"] = idx
    out[f"{prefix}E_samples"] = st.E.reshape(-1)[idx].astype(np.float64)
    out[f"{prefix}H_samples"] = st.H.reshape(-1)[idx].astype(np.float64)
    out[f"{prefix}E_plane_norms"] = plane_norms(st.E)
    out[f"{prefix}H_plane_norms"] = plane_norms(st.H)


def run_both(spec, fn):
    out = {}
    for tag, dt in (("f32_", np.float32), ("f64_", np.float64)):
        t0 = time.time()
        st, ds = F.run(S.build_scene(spec, dt), spec["T"])
        print(f"  {tag[:-1]} run {time.time() - t0:.0f} s", flush=True)
        field_record(tag, st, out)
        fn(tag, st, ds, out)
    return out


def golden_c1():
    spec = W.scenario_c1()

    def det(tag, st, ds, out):
        out[f"{tag}energy"] = ds[0].astype(np.float64)
        ez = ds[1]
        out["ez_index"] = sample_index(ez.size, SAMPLE_SEED + 1)
        out[f"{tag}ez_samples"] = ez.reshape(-1)[out["ez_index"]].astype(np.float64)
        out[f"{tag}ez_norm"] = np.float64(np.linalg.norm(ez.astype(np.float64)))

    return run_both(spec, det)


def golden_c2():
    spec = W.scenario_c2()

    def det(tag, st, ds, out):
        out[f"{tag}flux"] = ds[0].astype(np.float64)
        ph = ds[1]
        out["phasor_index"] = sample_index(ph.size, SAMPLE_SEED + 2)
        out[f"{tag}phasor_samples"] = ph.reshape(-1)[out["phasor_index"]].astype(np.complex128)
        out[f"{tag}phasor_row_norms"] = phasor_row_norms(ph)

    return run_both(spec, det)


def golden_c4():
    spec = W.scenario_c4()

    def det(tag, st, ds, out):
        ph = ds[0]
        out["phasor_index"] = sample_index(ph.size, SAMPLE_SEED + 4)
        out[f"{tag}phasor_samples"] = ph.reshape(-1)[out["phasor_index"]].astype(np.complex128)
        out[f"{tag}phasor_row_norms"] = phasor_row_norms(ph)

    return run_both(spec, det)


def _design_grad(spec, grad_inv_eps):
    """Chain rule of the continuous design (initialization.py:404-417): inv_eps = 1 / (eps_a +
    rho (eps_b - eps_a)), so d/d rho = -(eps_b - eps_a) / eps^2 * d/d inv_eps."""
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    return grad_inv_eps[(0, *slc)].astype(np.float64) * (-(d["eps_b"] - d["eps_a"]) / eps**2)


def golden_c3grad():
    spec = W.scenario_c3()
    t0 = time.time()
    sc = S.build_scene(spec, np.float32)
    grad, ds, _ = O.reversible_gradient(sc, spec["T"], lambda ds: [np.ones_like(ds[0])])
    print(f"  reversible gradient {time.time() - t0:.0f} s", flush=True)
    return {"f32_grad_rho": _design_grad(spec, grad).astype(np.float32), "f32_flux": ds[0].astype(np.float64)}


def fd_voxels(shape, n=FD_VOXELS, seed=FD_SEED):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, s, n) for s in shape], axis=1).astype(np.int64)


def golden_c3fd():
    spec = W.scenario_c3()
    d = spec["design"]
    rho = np.asarray(d["rho"], np.float64)

    def loss(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        _, ds = F.run(S.build_scene(s2, np.float64), spec["T"])
        return ds[0]

    t0 = time.time()
    flux64 = loss(rho)
    print(f"  base run {time.time() - t0:.0f} s", flush=True)
    vox = fd_voxels(rho.shape)
    fd = np.zeros(len(vox))
    for v, idx in enumerate(vox):
        vals = []
        for sgn in (+1, -1):
            r = rho.copy()
            r[tuple(idx)] += sgn * FD_H
            vals.append(float(np.sum(loss(r))))
        fd[v] = (vals[0] - vals[1]) / (2 * FD_H)
        print(f"  voxel {tuple(idx)} fd {fd[v]:.6e} ({time.time() - t0:.0f} s)", flush=True)
    return {"f64_flux": flux64.astype(np.float64), "fd_voxels": vox, "fd_h": np.float64(FD_H), "f64_fd": fd}


JOBS = {"c1": golden_c1, "c2": golden_c2, "c4": golden_c4, "c3grad": golden_c3grad, "c3fd": golden_c3fd}
FILES = {"c1": "c1.npz", "c2": "c2.npz", "c4": "c4.npz", "c3grad": "c3_grad.npz", "c3fd": "c3_fd.npz"}


def main(argv):
    for name in argv or list(JOBS):
        print(f"[golden] {name}", flush=True)
        t0 = time.time()
        out = JOBS[name]()
        np.savez_compressed(os.path.join(HERE, FILES[name]), **out)
        print(f"[golden] {name} written in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
