"""Re-export of the package workload builder for tests."""

from fdtdx_amd.scenarios import *  # noqa: F401,F403
from fdtdx_amd.scenarios import build  # noqa: F401

import numpy as np

from fdtdx_amd import scenarios as W

PERIODIC_XY = {"min_x": "periodic", "max_x": "periodic", "min_y": "periodic", "max_y": "periodic"}


def periodic_spec(T=80, shape=(18, 16, 30), pml=5, design=(6, 6, 4), seed=4):
    """Periodic in x and y, CPML in z: a dipole below a seeded design block and a z-normal flux
    plane spanning the whole periodic cross-section."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0, 1, size=design).astype(np.float32)
    d0 = (nx // 2 - design[0] // 2, ny // 2 - design[1] // 2, nz // 2 - design[2] // 2)
    return dict(shape=shape, dx=50e-9, T=T, background_eps=1.0,
                design=dict(slc=tuple((d0[a], d0[a] + design[a]) for a in range(3)), rho=rho, eps_a=1.0, eps_b=12.1),
                pml={"thickness": pml, "types": dict(PERIODIC_XY)},
                sources=[dict(kind="dipole", cell=(nx // 2, ny // 2, pml + 3), polarization=0,
                              profile=W.gaussian_profile(wavelength=0.4e-6, bandwidth=0.8))],
                detectors=[dict(name="flux", kind="flux", slc=((0, nx), (0, ny), (nz - pml - 3, nz - pml - 2)),
                                reduce_volume=True, axis=2, direction="+")])


def bloch_spec(T=120, shape=(20, 16, 28), pml=5, theta=(0.9, -1.7)):
    """Bloch phases theta on x and y (complex fields), CPML in z: an x-polarised dipole beside a
    dielectric block, with flux, phasor and energy detectors spanning the periodic cross-section
    (their interpolation stencils cross the Bloch faces), a reduced complex phasor, a field
    detector (real part) and an unreduced three-component flux box inside."""
    nx, ny, nz = shape
    dx = 50e-9
    f0 = 299792458.0 / 0.6e-6
    faces = {f: "bloch" for f in ("min_x", "max_x", "min_y", "max_y")}
    zt = nz - pml - 3
    return dict(shape=shape, dx=dx, T=T, background_eps=1.0,
                boxes=[dict(slc=((6, 13), (3, 11), (12, 18)), eps=(2.5, 3.0, 2.0))],
                pml={"thickness": pml, "types": faces},
                bloch_vector=(theta[0] / (nx * dx), theta[1] / (ny * dx), 0.0),
                sources=[dict(kind="dipole", cell=(2, ny // 2, pml + 4), polarization=0,
                              profile=W.gaussian_profile(wavelength=0.6e-6, bandwidth=0.8))],
                detectors=[dict(name="flux", kind="flux", slc=((0, nx), (0, ny), (zt, zt + 1)), reduce_volume=True,
                                axis=2, direction="+"),
                           dict(name="ph", kind="phasor", slc=((0, nx), (0, ny), (zt - 1, zt)),
                                frequencies=[0.9 * f0, 1.1 * f0], components=("Ex", "Ey", "Hx")),
                           dict(name="en", kind="energy", slc=((0, nx), (0, ny), (pml + 2, nz - pml - 2)),
                                reduce_volume=True),
                           dict(name="phr", kind="phasor", slc=((5, 12), (4, 10), (8, 14)), frequencies=[f0],
                                components=("Ez", "Hy"), reduce_volume=True),
                           dict(name="fld", kind="field", slc=((4, 9), (5, 11), (10, 13)), components=("Ex", "Hz")),
                           dict(name="fl3", kind="flux", slc=((6, 7), (6, 9), (14, 16)), reduce_volume=False,
                                keep_all_components=True, axis=0, direction="-")])


def detector_zoo_spec(T=90):
    """scenario_c3 with one detector of each kind and option behind the design region: apodized and
    reduced phasors, a reduced field detector, a reduced energy detector and an unreduced flux plane
    keeping all three components."""
    spec = W.scenario_c3(T=T, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    dt = 0.99 * spec["dx"] / (299792458.0 * np.sqrt(3.0))
    f0 = 299792458.0 / 1.55e-6
    spec["detectors"] = spec["detectors"] + [
        dict(name="ph", kind="phasor", slc=((15, 16), (6, 16), (6, 14)), frequencies=[0.9 * f0, 1.1 * f0],
             components=("Ez", "Hy"), apodization=dict(kind="tukey", start_time=0.0, end_time=T * dt, alpha=1.0)),
        dict(name="phr", kind="phasor", slc=((13, 17), (8, 14), (7, 13)), frequencies=[f0], components=("Ez", "Hx"),
             reduce_volume=True, apodization=dict(kind="gaussian", center_time=0.6 * T * dt, sigma_time=0.2 * T * dt)),
        dict(name="fld", kind="field", slc=((13, 17), (9, 13), (8, 12)), components=("Ey", "Ez", "Hy"),
             reduce_volume=True),
        dict(name="en", kind="energy", slc=((12, 18), (6, 16), (6, 14)), reduce_volume=True),
        dict(name="fl3", kind="flux", slc=((14, 15), (7, 15), (6, 14)), reduce_volume=False, keep_all_components=True,
             axis=0, direction="-"),
    ]
    return spec


def detector_zoo_losses(ds, weights):
    """One loss per detector of detector_zoo_spec and its cotangent (torch's convention for the
    complex phasors: dL/dRe + i dL/dIm)."""
    fl, ph, phr, fld, en, fl3 = ds
    losses = [np.sum(fl), np.sum(np.abs(ph) ** 2), np.sum(np.abs(phr) ** 2), np.sum(weights["fld"] * fld),
              np.sum(en), np.sum(weights["fl3"] * fl3)]
    cots = [np.ones_like(fl), 2 * ph, 2 * phr, weights["fld"].astype(fld.dtype), np.ones_like(en),
            weights["fl3"].astype(fl3.dtype)]
    return losses, cots


def surface_zoo_spec(T=90):
    """scenario_c3 with the closed-surface and phasor-flux detectors: the net flux through all six
    faces of a box around the dipole, the phasor flux through a plane behind the design region,
    and the net phasor flux through the x and z faces of a box around the design region."""
    spec = W.scenario_c3(T=T, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    f0 = 299792458.0 / 1.55e-6
    spec["detectors"] = spec["detectors"] + [
        dict(name="cf", kind="closed_flux", slc=((6, 11), (8, 15), (7, 14))),
        dict(name="pf", kind="phasor_flux", slc=((16, 17), (6, 16), (6, 14)), frequencies=[0.9 * f0, 1.1 * f0],
             direction="+"),
        dict(name="cp", kind="closed_phasor", slc=((8, 17), (7, 15), (6, 14)), frequencies=[f0], axes=(0, 2),
             orientation="inward"),
    ]
    return spec


def surface_zoo_losses(ds, names, spec, dx):
    """Loss per spec detector of surface_zoo_spec (sum of flux, sum of net flux, sum of phasor flux,
    sum of net phasor flux) and, per loss, the cotangent of each oracle detector (names: the
    oracle's detector names, the closed phasor expanded into its faces)."""
    from oracle import fdtd_oracle as O

    by = dict(zip(names, ds))
    w = dx**2
    losses, cots = {}, {}
    losses["flux"], cots["flux"] = float(np.sum(by["flux"])), {"flux": np.ones_like(by["flux"])}
    losses["cf"], cots["cf"] = float(np.sum(by["cf"])), {"cf": np.ones_like(by["cf"])}
    pf = by["pf"][0]
    fl = O.phasor_poynting_flux(pf, 0, 1, w)
    losses["pf"] = float(np.sum(fl))
    cots["pf"] = {"pf": O.phasor_poynting_flux_cotangent(pf, 0, 1, w, np.ones_like(fl))[None]}
    faces = {tuple(n.split(":phasor_axis")[1].split("_")): by[n] for n in names if n.startswith("cp:")}
    faces = {(int(a), side): st for (a, side), st in faces.items()}
    net = O.closed_phasor_net_flux(faces, w, inward=True)
    losses["cp"] = float(np.sum(net))
    cots["cp"] = {f"cp:phasor_axis{a}_{side}": O.phasor_poynting_flux_cotangent(
        st[0], a, -(1 if side == "max" else -1), w, np.ones(st.shape[1]))[None] for (a, side), st in faces.items()}
    return losses, cots


# non-symmetric permittivity and permeability tensors (row-major) with real positive eigenvalues:
# fine for one-step algebra, but a non-reciprocal medium with gain, so long runs grow without bound
TENSOR_EPS = (3.0, 0.6, 0.2, -0.3, 2.5, 0.4, 0.1, 0.5, 4.0)
TENSOR_MU = (1.5, 0.2, 0.0, -0.1, 1.2, 0.3, 0.2, 0.0, 2.0)
# symmetric positive-definite (physical, lossless) tensors for full runs
SYM_TENSOR_EPS = (3.0, 0.6, 0.2, 0.6, 2.5, 0.4, 0.2, 0.4, 4.0)
SYM_TENSOR_MU = (1.5, 0.2, 0.0, 0.2, 1.2, 0.3, 0.0, 0.3, 2.0)


def tensor_spec(T=120, n=28, pml=6, periodic=False, walls=None):
    """A z-dipole inside a cube of symmetric full-tensor permittivity and permeability (so the
    dipole injects into all three E components), CPML (or periodic x and y), a field snapshot and a
    flux plane."""
    spec = W.scenario_c1(T=T, n=n, pml=pml, cube=n // 3)
    c = n // 2
    spec["boxes"] = [dict(slc=((c - 4, c + 4), (c - 3, c + 5), (c - 5, c + 3)), eps=SYM_TENSOR_EPS,
                          mu=SYM_TENSOR_MU)]
    spec["detectors"] = [
        dict(name="ez", kind="field", slc=((pml, n - pml), (c, c + 1), (pml, n - pml)), components=("Ex", "Ez", "Hy")),
        dict(name="flux", kind="flux", slc=((n - pml - 3, n - pml - 2), (pml, n - pml), (pml, n - pml)),
             reduce_volume=True, axis=0, direction="+"),
    ]
    types = dict(PERIODIC_XY) if periodic else {}
    if walls:
        types.update(walls)
    spec["pml"] = dict(spec["pml"], types=types)
    return spec


# full conductivity tensors (S/m, row-major): a symmetric positive-definite electric one (a lossy
# anisotropic conductor) and a magnetic one; with the symmetric tensors above both damp every mode
SIGMA_E_TENSOR = (3.0e4, 1.0e4, -0.5e4, 1.0e4, 2.0e4, 0.8e4, -0.5e4, 0.8e4, 2.5e4)
SIGMA_H_TENSOR = (2.0e9, 0.5e9, 0.0, 0.5e9, 1.5e9, 0.3e9, 0.0, 0.3e9, 1.0e9)


def tensor_loss_spec(T=120, n=28, pml=6, magnetic=True):
    """tensor_spec with full electric (and magnetic) conductivity tensors on the tensor cube and on
    a second, diagonal-permittivity box: inside both the update is A avg(F) +- B avg(K) with
    per-cell matrices."""
    spec = tensor_spec(T=T, n=n, pml=pml)
    c = n // 2
    box = dict(spec["boxes"][0], sigma_E=SIGMA_E_TENSOR)
    if magnetic:
        box["sigma_H"] = SIGMA_H_TENSOR
    spec["boxes"] = [box, dict(slc=((c + 4, c + 9), (c - 6, c), (c - 2, c + 4)), eps=(2.0, 3.0, 2.5),
                               sigma_E=(1.0e4, 0.0, 0.2e4, 0.0, 1.5e4, 0.0, 0.2e4, 0.0, 0.8e4))]
    return spec


def oriented_ade_spec(T=150, n=28, pml=6, sigma=False):
    """ADE poles in the full-tensor update: a diagonal-permittivity box with an oriented Lorentz pole
    (u off every axis) touching a box of symmetric full-tensor permittivity with an isotropic
    Lorentz pole and a Drude pole oriented along y; optionally a full conductivity tensor on the
    second box. A dipole in front, CPML, a field plane and a flux plane."""
    spec = tensor_spec(T=T, n=n, pml=pml)
    c = n // 2
    w0, w1 = 2 * np.pi * 1.6e14, 2 * np.pi * 2.1e14
    u = (1.0, 2.0, 2.0)
    b1 = dict(slc=((c - 5, c + 1), (c - 4, c + 4), (c - 3, c + 3)), eps=(2.0, 2.5, 3.0),
              poles=[("lorentz", w0, 0.08 * w0, 1.2, u)])
    b2 = dict(slc=((c + 1, c + 6), (c - 4, c + 3), (c - 4, c + 2)), eps=SYM_TENSOR_EPS,
              poles=[("lorentz", w1, 0.05 * w1, 0.6), ("drude", 1.2e15, 5e13, (0.0, 1.0, 0.0))])
    if sigma:
        b2["sigma_E"] = SIGMA_E_TENSOR
    spec["boxes"] = [b1, b2]
    spec["sources"] = [dict(spec["sources"][0], cell=(c - 8, c, c))]
    return spec


def symmetry_spec(T=90, shape=(36, 24, 20), pml=5, symmetry=(0, 1, -1), e_pol=(0, 0, 1)):
    """An x-directed plane wave (z-polarized by default) through a slab and a design block centred
    on the y and z mid-planes, which are mirror planes of the fields; the default symmetry is
    magnetic across y (E_z and H_y even, E_y odd) and electric across z, and a y-polarized wave fits
    (0, -1, +1). Detectors span the whole y-z cross-section: a reduced flux plane, an x-plane field
    snapshot and a phasor plane."""
    nx, ny, nz = shape
    rng = np.random.default_rng(2)
    cy, cz = ny // 2, nz // 2
    half = rng.uniform(0, 1, size=(4, 3, 2))
    rho = np.concatenate([half[:, ::-1], half], axis=1)  # mirror-symmetric in y ...
    rho = np.concatenate([rho[:, :, ::-1], rho], axis=2).astype(np.float32)  # ... and in z
    f0 = 299792458.0 / 1.55e-6
    return dict(shape=shape, dx=50e-9, T=T, background_eps=1.0, symmetry=symmetry,
                boxes=[dict(slc=((nx // 2 + 3, nx // 2 + 6), (cy - 5, cy + 5), (cz - 3, cz + 3)), eps=2.25)],
                design=dict(slc=((nx // 2 - 3, nx // 2 + 1), (cy - 3, cy + 3), (cz - 2, cz + 2)), rho=rho,
                            eps_a=1.0, eps_b=6.0),
                pml={"thickness": pml},
                sources=[dict(kind="plane", axis=0, index=pml + 3, direction="+", e_pol=e_pol,
                              profile=W.gaussian_profile())],
                detectors=[dict(name="flux", kind="flux", slc=((nx - pml - 4, nx - pml - 3), (0, ny), (0, nz)),
                                reduce_volume=True, axis=0, direction="+"),
                           dict(name="snap", kind="field", slc=((nx - pml - 6, nx - pml - 5), (0, ny), (0, nz)),
                                components=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")),
                           dict(name="ph", kind="phasor", slc=((nx - pml - 5, nx - pml - 4), (0, ny), (0, nz)),
                                frequencies=[f0], components=("Ey", "Ez", "Hy", "Hz"))])


def full_spec(spec):
    """The same scenario without the symmetry reduction."""
    out = dict(spec)
    out.pop("symmetry", None)
    return out


def energy_slices_spec(T=100, n=24, pml=5):
    """The dipole-in-a-cube scenario with EnergyDetector(as_slices=True) twice: collapsed-axis
    means over a box around the cube, and planes at given cells of a second box."""
    spec = W.scenario_c1(T=T, n=n, pml=pml, cube=6)
    c = n // 2
    spec["detectors"] = [
        dict(name="mean", kind="energy", slc=((c - 5, c + 4), (c - 4, c + 5), (c - 3, c + 6)), as_slices=True),
        dict(name="planes", kind="energy", slc=((pml, n - pml), (pml + 1, n - pml), (pml, n - pml - 1)), as_slices=True,
             x_slice_index=4, y_slice_index=6, z_slice_index=5),
    ]
    return spec


def backward_spec(T=80):
    """scenario_c3 with a forward field detector and, over the same box behind the design region,
    inverse (backward-time) field, phasor and reduced energy detectors, all point-sampled."""
    spec = W.scenario_c3(T=T, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    f0 = 299792458.0 / 1.55e-6
    box = ((13, 16), (8, 13), (7, 12))
    spec["detectors"] = [
        dict(name="fwd", kind="field", slc=box, components=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"),
             exact_interpolation=False),
        dict(name="ifld", kind="field", slc=box, components=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"),
             exact_interpolation=False, inverse=True),
        dict(name="iph", kind="phasor", slc=box, frequencies=[0.9 * f0, 1.1 * f0], components=("Ez", "Hy"),
             exact_interpolation=False, inverse=True),
        dict(name="ien", kind="energy", slc=box, reduce_volume=True, exact_interpolation=False, inverse=True),
    ]
    return spec


def mode_spec(T=151, divergence=5.0):
    """The small plane-wave case with a Gaussian mode-overlap plane (3 frequencies, 0.35 um radius,
    a diverging reference wavefront) beside its flux and phasor planes."""
    from fdtdx_amd import constants

    spec = W.scenario_c2(T=T, shape=(60, 40, 32), pml=6, slab=(20, 20, 6))
    f0 = constants.c / 1.55e-6
    spec["detectors"] = spec["detectors"] + [
        dict(name="mode", kind="mode_gaussian", slc=((48, 49), (0, 40), (0, 32)), frequencies=[0.9 * f0, f0, 1.1 * f0],
             mode_radius=0.35e-6, divergence_angle=divergence)]
    return spec


def oracle_modes(spec, det_name="mode"):
    """The oracle's unit-flux Gaussian reference modes of a mode_gaussian detector of a spec (x-normal
    plane), one (mode_E, mode_H) per frequency, on cell-centre coordinates about the domain centre."""
    import numpy as np

    from oracle import fdtd_oracle as O
    from oracle import scenes as S

    d = next(x for x in spec["detectors"] if x["name"] == det_name)
    (i0, _), (j0, j1), (k0, k1) = d["slc"]
    nx, ny, nz = spec["shape"]
    dx = spec["dx"]
    u, v = np.meshgrid((np.arange(j0, j1) + 0.5) * dx - ny * dx / 2, (np.arange(k0, k1) + 0.5) * dx - nz * dx / 2,
                       indexing="ij")
    ie = S.build_inv_eps(spec, np.float32)[:, i0:i0 + 1, j0:j1, k0:k1].astype(np.float64)
    n = float(np.sqrt(np.mean(1.0 / ie)))
    out = []
    for f in d["frequencies"]:
        mE, mH = O.gaussian_mode(u[None], v[None], 0, d["mode_radius"], d.get("direction", "+"), 1, n,
                                 299792458.0 / f, d.get("divergence_angle", 0.0))
        out.append(O.unit_flux(mE, mH, 0))
    return out


def waveguide_spec(T=240, shape=(60, 34, 26), pml=6, wavelength=1.55e-6):
    """A strip waveguide along x (eps 12.1, 7 x 4 cells, off centre so no two of its modes are
    degenerate) in eps 2.1 cladding, its fundamental mode launched by a mode source three cells
    inside the low-x CPML, reduced flux planes behind and ahead of the source, a mode-overlap plane
    (3 frequencies) and a phasor plane of the transverse fields ahead."""
    from fdtdx_amd import constants

    nx, ny, nz = shape
    f0 = constants.c / wavelength
    core = ((0, nx), (ny // 2 - 4, ny // 2 + 3), (nz // 2 - 2, nz // 2 + 2))
    src_x = pml + 3
    return dict(
        shape=shape, dx=50e-9, T=T, background_eps=2.1, pml={"thickness": pml},
        boxes=[dict(slc=core, eps=12.1)],
        sources=[dict(kind="mode", axis=0, index=src_x, direction="+", mode_index=0,
                      profile=W.gaussian_profile(wavelength=wavelength, bandwidth=0.2))],
        detectors=[
            dict(name="back", kind="flux", slc=((src_x - 2, src_x - 1), (pml, ny - pml), (pml, nz - pml)),
                 axis=0, direction="-", reduce_volume=True),
            dict(name="flux", kind="flux", slc=((nx - pml - 8, nx - pml - 7), (pml, ny - pml), (pml, nz - pml)),
                 axis=0, direction="+", reduce_volume=True),
            dict(name="mode", kind="mode", slc=((nx - pml - 6, nx - pml - 5), (0, ny), (0, nz)),
                 frequencies=[0.95 * f0, f0, 1.05 * f0], direction="+"),
            dict(name="ph", kind="phasor", slc=((nx - pml - 6, nx - pml - 5), (pml, ny - pml), (pml, nz - pml)),
                 frequencies=[f0], components=("Ey", "Ez", "Hy", "Hz")),
        ])


def with_oracle_mode(spec, dtype=None):
    """The spec with each mode source's cross-section (E, H, neff) attached for the oracle: the mode
    of the source plane of the oracle's own permittivity raster, from the mode solver."""
    import numpy as np

    from fdtdx_amd import modes as M
    from oracle import scenes as S

    spec = dict(spec)
    ie = S.build_inv_eps(spec, np.float32).astype(np.float64)
    srcs = []
    for s in spec["sources"]:
        if s["kind"] == "mode":
            slc = [slice(None)] * 3
            slc[s["axis"]] = slice(s["index"], s["index"] + 1)
            E, H, neff = M.compute_mode(s["profile"]["center_frequency"], ie[(slice(None), *slc)], 1.0, spec["dx"],
                                        direction=s.get("direction", "+"), mode_index=s.get("mode_index", 0),
                                        filter_pol=s.get("filter_pol"))
            s = dict(s, mode=(E, H, neff))
        srcs.append(s)
    spec["sources"] = srcs
    return spec


def oracle_solver_modes(spec, det_name="mode"):
    """Reference modes of a solver-based mode-overlap detector (x-normal plane): one unit-flux
    (mode_E, mode_H) per frequency from the mode solver on the oracle's permittivity raster."""
    import numpy as np

    from fdtdx_amd import modes as M
    from oracle import scenes as S

    d = next(x for x in spec["detectors"] if x["name"] == det_name)
    sl = tuple(slice(*r) for r in d["slc"])
    ie = S.build_inv_eps(spec, np.float32)[(slice(None), *sl)].astype(np.float64)
    out = []
    for f in d["frequencies"]:
        mE, mH, _ = M.compute_mode(f, ie, 1.0, spec["dx"], direction=d.get("direction", "+"),
                                   mode_index=d.get("mode_index", 0), filter_pol=d.get("filter_pol"))
        out.append((mE, mH))
    return out
