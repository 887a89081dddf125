"""GPU parity of the forward time loop (HIP engine through the C ABI) against the CPU oracle."""

import numpy as np
import pytest

from oracle import fdtd_oracle as O
from oracle import scenes as S
from fdtdx_amd import scenarios as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a).astype(np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b).astype(np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# Sweep schedules: the two in-place sweeps with four z cells per lane (the default when nz is a
# multiple of 4) and with one, the LDS-tiled fused kernel and the register-march kernel with 4
# and 2 rows per wavefront (short x chunks so chunk halos are exercised).
MODES = {
    "quad": {"FDTDX_SWEEP": "quad"},
    "two": {"FDTDX_SWEEP": "two"},
    "fused": {"FDTDX_SWEEP": "fused"},
    "march4": {"FDTDX_SWEEP": "march", "FDTDX_MARCH_TY": "4", "FDTDX_MARCH_XC": "5"},
    "march2": {"FDTDX_SWEEP": "march", "FDTDX_MARCH_TY": "2", "FDTDX_MARCH_XC": "3"},
}


def _set_mode(monkeypatch, mode):
    for k in ("FDTDX_SWEEP", "FDTDX_MARCH_TY", "FDTDX_MARCH_XC"):
        monkeypatch.delenv(k, raising=False)
    for k, v in MODES[mode].items():
        monkeypatch.setenv(k, v)


@pytest.fixture(params=list(MODES))
def sweep(request, monkeypatch):
    _set_mode(monkeypatch, request.param)
    return request.param


def _run(spec):
    import torch

    import fdtdx_amd as fx
    import scenarios

    objects, arrays, params, config = scenarios.build(spec)
    print(f"case shape={spec['shape']} T={spec['T']}", flush=True)
    _, arrays = fx.run_fdtd(arrays, objects, config)
    torch.cuda.synchronize()
    return objects, arrays


def test_dipole_cube_small_matches_oracle(sweep):
    spec = W.scenario_c1(T=200, n=32, pml=6, cube=8)
    objects, arrays = _run(spec)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H))
    assert rel(E, st.E) < 1e-5
    assert rel(H, st.H) < 1e-5
    en = arrays.detector_states["energy"]["energy"].cpu().numpy()
    assert rel(en, ds[0]) < 1e-5
    ez = arrays.detector_states["ez"]["fields"].cpu().numpy()
    assert rel(ez, ds[1]) < 1e-5


def test_plane_wave_design_small_matches_oracle(sweep):
    spec = W.scenario_c2(T=151, shape=(60, 40, 30), pml=6, slab=(20, 20, 6))
    objects, arrays = _run(spec)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    print("E", rel(E, st.E))
    assert rel(E, st.E) < 1e-5
    assert rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    flux = arrays.detector_states["flux"]["poynting_flux"].cpu().numpy()
    assert rel(flux, ds[0]) < 1e-5
    ph = arrays.detector_states["phasor"]["phasor"].cpu().numpy()
    assert rel(ph, ds[1]) < 1e-5


def test_dispersive_anisotropic_small_matches_oracle(sweep):
    spec = W.scenario_c4(T=120, n=40, pml=6, slab=(16, 16, 6), cube=6)
    objects, arrays = _run(spec)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    print("E", rel(E, st.E))
    assert rel(E, st.E) < 1e-5
    assert rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    ph = arrays.detector_states["phasor"]["phasor"].cpu().numpy()
    assert rel(ph, ds[0]) < 1e-5


def test_quad_ade_tiles_match_the_one_cell_path(monkeypatch):
    """ADE tiles of the quad schedule (listed for the E half-step only; the H half-step sweeps
    them four cells per lane): fields, ADE state and phasors equal the one-cell schedule's bit for
    bit, and the oracle's within 1e-5."""
    import scenarios

    from fdtdx_amd import fdtd as F

    spec = W.scenario_c4(T=120, n=40, pml=6, slab=(16, 16, 8), cube=8)
    out = {}
    for mode in ("quad", "two"):
        _set_mode(monkeypatch, mode)
        objects, arrays, params, config = scenarios.build(spec)
        F.run_fdtd(arrays, objects, config)
        solve = F._solve_for(objects, config, arrays)
        out[mode] = (solve.plan.schedule, arrays.fields.E.cpu().numpy(), arrays.fields.H.cpu().numpy(),
                     arrays.ade_state.cpu().numpy(), arrays.detector_states["phasor"]["phasor"].cpu().numpy())
    assert out["quad"][0] != out["two"][0]
    for a, b in zip(out["quad"][1:], out["two"][1:]):
        np.testing.assert_array_equal(a, b)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    assert rel(out["quad"][1], st.E) < 1e-5 and rel(out["quad"][4], ds[0]) < 1e-5


def test_single_pass_equals_two_sweep_bitwise(monkeypatch):
    """The single-pass kernels perform the two-sweep kernels' float operations in the same order:
    fields, psi, detector states and the interface record agree bit for bit (z split into two
    lane runs, partial y tiles and x chunks, an odd step count so the ping-pong copies back)."""
    import torch

    from fdtdx_amd import fdtd as F
    import scenarios

    spec = W.scenario_c2(T=37, shape=(50, 45, 70), pml=5, slab=(20, 16, 8))
    out = {}
    for mode in MODES:
        _set_mode(monkeypatch, mode)
        objects, arrays, params, config = scenarios.build(spec)
        s = F.Solve(objects, config, arrays, record_steps=spec["T"])
        arrays.reset()
        print(f"mode {mode}", flush=True)
        s.forward(0, spec["T"], record=True)
        torch.cuda.synchronize()
        out[mode] = [arrays.fields.E.cpu().numpy(), arrays.fields.H.cpu().numpy(), arrays.psi.cpu().numpy(),
                     s.record.cpu().numpy()] + [t.cpu().numpy() for t in s.det_tensors()]
    for mode in MODES:
        for a, b in zip(out[mode], out["two"]):
            assert np.array_equal(a, b), mode


@pytest.mark.parametrize("mode", ["two", "fused"])
def test_conductor_faces_match_oracle(mode, monkeypatch):
    """PEC / PMC faces (pec.py:70-77, pmc.py:62-75) on three faces, CPML on the others; a
    single-pass request falls back to the two sweeps, which apply the faces."""
    _set_mode(monkeypatch, mode)
    spec = W.scenario_c1(T=150, n=28, pml=5, cube=6)
    spec["pml"] = dict(spec["pml"], types={"max_x": "pec", "min_y": "pmc", "max_z": "pec"})
    spec["sources"][0]["profile"] = W.gaussian_profile(wavelength=0.4e-6, bandwidth=0.8)
    objects, arrays = _run(spec)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H))
    assert rel(E, st.E) < 1e-5 and rel(H, st.H) < 1e-5
    assert not E[1:, -1].any() and not H[0, :, 0].any() and not H[2, :, 0].any() and not E[:2, :, :, -1].any()
    assert rel(arrays.detector_states["energy"]["energy"].cpu().numpy(), ds[0]) < 1e-5


@pytest.mark.parametrize("mode", ["two", "fused"])
def test_periodic_unit_cell_plane_wave_matches_oracle(mode, monkeypatch):
    """Periodic y and z (wrap padding, update.py:92-137) around an x-directed plane wave with a
    design slab: the classic unit cell. Flux and phasor planes span the periodic cross-section, so
    their interpolation stencils wrap too; a single-pass request falls back to the two sweeps."""
    _set_mode(monkeypatch, mode)
    spec = W.scenario_c2(T=151, shape=(60, 24, 20), pml=6, slab=(20, 12, 6))
    spec["pml"] = dict(spec["pml"], types={"min_y": "periodic", "max_y": "periodic", "min_z": "periodic",
                                           "max_z": "periodic"})
    objects, arrays = _run(spec)
    sc = S.build_scene(spec)
    assert sc.periodic == (False, True, True)
    st, ds = O.run(sc, spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H), flush=True)
    assert rel(E, st.E) < 1e-5 and rel(H, st.H) < 1e-5
    assert rel(arrays.detector_states["flux"]["poynting_flux"].cpu().numpy(), ds[0]) < 1e-5
    assert rel(arrays.detector_states["phasor"]["phasor"].cpu().numpy(), ds[1]) < 1e-5


def test_periodic_dipole_matches_oracle():
    """Periodic x and y with CPML in z; a dipole radiates across the wrapped faces."""
    from scenarios import periodic_spec

    spec = periodic_spec(T=120)
    objects, arrays = _run(spec)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    print("E", rel(E, st.E), flush=True)
    assert rel(E, st.E) < 1e-5 and rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    # the field has crossed the wrapped faces: both edge planes of x carry it
    assert np.abs(E[:, 0]).max() > 1e-3 * np.abs(E).max() and np.abs(E[:, -1]).max() > 1e-3 * np.abs(E).max()
    assert rel(arrays.detector_states["flux"]["poynting_flux"].cpu().numpy(), ds[0]) < 1e-5


@pytest.mark.parametrize("mode", ["two", "march4"])
def test_conductivity_matches_oracle(mode, monkeypatch):
    """Electric (diagonal) and magnetic (isotropic) conductivity in two overlapping boxes, one
    reaching into the CPML (update.py:300-354, 741-750); a single-pass request falls back to the
    two sweeps, which apply the loss."""
    _set_mode(monkeypatch, mode)
    spec = W.scenario_c1(T=150, n=28, pml=5, cube=6)
    spec["boxes"] = spec["boxes"] + [
        dict(slc=((16, 26), (8, 20), (2, 12)), eps=2.0, sigma_E=(3e4, 1e4, 2e4), sigma_H=2e9),
        dict(slc=((4, 10), (4, 28), (15, 22)), eps=1.5, sigma_E=(5e3, 8e3, 6e3))]
    objects, arrays = _run(spec)
    sc = S.build_scene(spec)
    assert sc.sigma_E is not None and sc.sigma_H is not None
    st, ds = O.run(sc, spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H), flush=True)
    assert rel(E, st.E) < 1e-5 and rel(H, st.H) < 1e-5
    assert rel(arrays.detector_states["energy"]["energy"].cpu().numpy(), ds[0]) < 1e-5


def test_conductivity_with_dispersion_matches_oracle():
    """Loss divides after the ADE correction (update.py:338-350): conductivity on the dispersive
    box of the anisotropic Lorentz/Drude case."""
    spec = W.scenario_c4(T=120, n=40, pml=6, slab=(16, 16, 6), cube=6)
    spec["boxes"] = [dict(b, sigma_E=(2e4, 1e4, 3e4)) if b.get("poles") else b for b in spec["boxes"]]
    objects, arrays = _run(spec)
    sc = S.build_scene(spec)
    assert sc.sigma_E is not None and sc.ade is not None
    st, ds = O.run(sc, spec["T"])
    E = arrays.fields.E.cpu().numpy()
    print("E", rel(E, st.E), flush=True)
    assert rel(E, st.E) < 1e-5 and rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    assert rel(arrays.detector_states["phasor"]["phasor"].cpu().numpy(), ds[0]) < 1e-5


def test_plane_source_in_dispersive_medium_matches_oracle():
    """A broadband plane wave launched inside a Lorentz medium (effective carrier permittivity and
    the filtered E-side table, linear_polarization.py:191-210, 330-351)."""
    spec = W.scenario_c2(T=151, shape=(60, 30, 24), pml=6, slab=(20, 12, 6))
    w = 2 * np.pi * 299792458.0 / 1.55e-6
    spec["boxes"] = [dict(slc=((0, 40), (0, 30), (0, 24)), eps=2.1, poles=[("lorentz", 1.4 * w, 0.1 * w, 0.8)])]
    objects, arrays = _run(spec)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    print("E", rel(E, st.E), flush=True)
    assert rel(E, st.E) < 1e-5 and rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    assert rel(arrays.detector_states["flux"]["poynting_flux"].cpu().numpy(), ds[0]) < 1e-5


def test_detector_options_match_oracle():
    """Reduced field and phasor detectors, apodized phasors (Tukey and Gaussian windows) and a flux
    plane keeping all three components, against the oracle."""
    from scenarios import detector_zoo_spec

    spec = detector_zoo_spec(T=120)
    objects, arrays = _run(spec)
    _, ds = O.run(S.build_scene(spec), spec["T"])
    keys = dict(flux="poynting_flux", ph="phasor", phr="phasor", fld="fields", en="energy", fl3="poynting_flux")
    for q, d in enumerate(spec["detectors"]):
        got = arrays.detector_states[d["name"]][keys[d["name"]]].cpu().numpy()
        print(d["name"], got.shape, "rel", rel(got, ds[q]), flush=True)
        assert got.shape == ds[q].shape
        assert rel(got, ds[q]) < 1e-5


def _surface_states(objects, arrays, spec):
    """Detector states of surface_zoo_spec in the oracle's (expanded) detector order."""
    out = []
    for d in spec["detectors"]:
        st = arrays.detector_states[d["name"]]
        if d["kind"] == "closed_phasor":
            for a in d["axes"]:
                for side in ("min", "max"):
                    out.append(st[f"phasor_axis{a}_{side}"])
        else:
            out.append(next(iter(st.values())))
    return [t.detach().cpu().numpy() for t in out]


def test_surface_detectors_match_oracle():
    """The net flux through the six faces of a box (one reduced flux detector per face, the later
    ones accumulating), the phasors of a phasor-flux plane and of the faces of a closed phasor
    surface against the oracle; the post-processed phasor flux and net phasor flux (computed on the
    device) against the oracle's post-processing of its own phasors."""
    import fdtdx_amd as fx
    from scenarios import surface_zoo_spec

    spec = surface_zoo_spec(T=120)
    objects, arrays = _run(spec)
    sc = S.build_scene(spec)
    _, ds = O.run(sc, spec["T"])
    got = _surface_states(objects, arrays, spec)
    for d, g, w in zip(sc.detectors, got, ds):
        print(d["name"], g.shape, "rel", rel(g, w), flush=True)
        assert g.shape == w.shape
        assert rel(g, w) < 1e-5
    dx2 = float(np.float32(spec["dx"] ** 2))
    pf = next(o for o in objects.detectors if isinstance(o, fx.PhasorPoyntingFluxDetector))
    by = {d["name"]: w for d, w in zip(sc.detectors, ds)}
    host = pf.compute_poynting_flux(arrays.detector_states["pf"]).cpu().numpy()
    want = O.phasor_poynting_flux(by["pf"][0], 0, 1, dx2)
    print("phasor flux", host, want, flush=True)
    assert rel(host, want) < 1e-5
    cp = next(o for o in objects.detectors if isinstance(o, fx.ClosedSurfacePhasorPoyntingFluxDetector))
    net = cp.compute_net_flux(arrays.detector_states["cp"]).cpu().numpy()
    faces = {(a, side): by[f"cp:phasor_axis{a}_{side}"] for a in (0, 2) for side in ("min", "max")}
    want = O.closed_phasor_net_flux(faces, dx2, inward=True)
    print("net phasor flux", net, want, flush=True)
    assert rel(net, want) < 1e-5


def test_batch_accumulating_detectors_equal_the_plain_solve():
    """A closed surface (face detectors accumulating into one state) advanced in a multi-instance
    batch equals the same solve run through fdtdx_forward, bit for bit; a plan whose accumulating
    detector follows no matching one is rejected."""
    import torch

    from fdtdx_amd import engine
    from fdtdx_amd.fdtd import Solve
    from scenarios import surface_zoo_spec

    spec = surface_zoo_spec(T=100)
    objects, arrays, params, config = W.build(spec)
    solve = Solve(objects, config, arrays)
    arrays.reset()
    solve.forward(0, spec["T"])
    torch.cuda.synchronize()
    plain = [t.clone() for t in solve.state_tensors()]
    arrays.reset()
    torch.cuda.synchronize()
    b = engine.Batch(solve.plan.spec["eps_comps"], solve.plan.spec["mu_comps"])
    try:
        buf = solve.buffers()
        b.attach(0, solve.plan, buf, solve._keep, 0, spec["T"])
        b.advance(spec["T"], engine.current_stream_handle(), graph=False)
        torch.cuda.synchronize()
        assert b.step_index(0) == spec["T"]
        for (name, key), x, y in zip(solve.state_keys, solve.state_tensors(), plain):
            print(name, key, "equal", bool(torch.equal(x, y)), flush=True)
            assert torch.equal(x, y), (name, key)
    finally:
        b.close()
    cf = [d for d in solve.spec["detectors"] if d["name"] == "cf"]
    with pytest.raises(engine.EngineError, match="accumulate"):
        engine.Plan(dict(solve.spec, detectors=[dict(cf[1])]))


@pytest.mark.parametrize("variant", ["cpml", "periodic_walls"])
def test_full_tensor_matches_oracle(variant):
    """Non-symmetric full-tensor permittivity and permeability (the curl sweep, then the tensor
    sweep with four-point means of the off-diagonal curl components) with a dipole injecting into
    all three components: CPML on every face, or periodic x and y with a PEC and a PMC face on z."""
    from scenarios import tensor_spec

    spec = tensor_spec(periodic=variant != "cpml", walls=None if variant == "cpml" else {"max_z": "pec", "min_z": "pmc"})
    objects, arrays = _run(spec)
    assert arrays.inv_permittivities.shape[0] == 9
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H), flush=True)
    assert rel(E, st.E) < 1e-5
    assert rel(H, st.H) < 1e-5
    for q, name in enumerate(("ez", "flux")):
        got = next(iter(arrays.detector_states[name].values())).cpu().numpy()
        print(name, rel(got, ds[q]), flush=True)
        assert rel(got, ds[q]) < 1e-5


@pytest.mark.parametrize("symmetry,e_pol", [((0, 1, -1), (0, 0, 1)), ((0, -1, 1), (0, 1, 0))])
def test_symmetry_reduced_run_matches_oracle(symmetry, e_pol):
    """A mirror-symmetry-reduced run (PEC face on the electric plane, CPML on the far faces only,
    detector stencils reading the electric plane's mirror halo) against the oracle's reduced run:
    fields, flux, x-plane snapshot and phasors."""
    from scenarios import symmetry_spec

    spec = symmetry_spec(symmetry=symmetry, e_pol=e_pol, T=120)
    objects, arrays = _run(spec)
    assert tuple(arrays.fields.E.shape) == (3, 36, 12, 10)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    print("E", rel(arrays.fields.E.cpu().numpy(), st.E), flush=True)
    assert rel(arrays.fields.E.cpu().numpy(), st.E) < 1e-5
    assert rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    for q, name in enumerate(("flux", "snap", "ph")):
        got = next(iter(arrays.detector_states[name].values())).cpu().numpy()
        print(name, rel(got, ds[q]), flush=True)
        assert rel(got, ds[q]) < 1e-5


def test_energy_slices_match_oracle():
    """EnergyDetector(as_slices=True): collapsed-axis means (one thread per plane cell, float64
    sums) and planes at given cells, against the oracle."""
    from scenarios import energy_slices_spec

    spec = energy_slices_spec()
    objects, arrays = _run(spec)
    _, ds = O.run(S.build_scene(spec), spec["T"])
    q = 0
    for name in ("mean", "planes"):
        for key in ("XY Plane", "XZ Plane", "YZ Plane"):
            got = arrays.detector_states[name][key].cpu().numpy()
            want = ds[q].reshape(got.shape)
            print(name, key, rel(got, want), flush=True)
            assert rel(got, want) < 1e-5
            q += 1


def test_single_frequency_dipole_matches_oracle():
    """The reference's default temporal profile, SingleFrequencyProfile (profile.py:254-275), with
    a wave-character phase shift, driving the dipole of the cube case."""
    from fdtdx_amd import constants

    spec = W.scenario_c1(T=150, n=28, pml=6, cube=8)
    spec["sources"][0]["profile"] = dict(kind="single", frequency=constants.c / 1.55e-6, phase_shift=0.4)
    objects, arrays = _run(spec)
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    print("single-frequency E", rel(E, st.E), flush=True)
    assert np.abs(st.E).max() > 0
    assert rel(E, st.E) < 1e-5
    assert rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    assert rel(arrays.detector_states["energy"]["energy"].cpu().numpy(), ds[0]) < 1e-5


def _quad_cases():
    c1 = W.scenario_c1(T=60, n=32, pml=8, cube=8)                                  # z slabs on quad edges
    c2 = W.scenario_c2(T=50, shape=(40, 36, 28), pml=5, slab=(12, 12, 8))          # z slabs inside quads
    c4 = W.scenario_c4(T=50, n=36, pml=8, slab=(12, 12, 8), cube=4)                # diagonal eps, ADE tiles
    lossy = W.scenario_c1(T=60, n=28, pml=4, cube=6)                               # conductivity, walls, mu
    lossy["pml"] = dict(lossy["pml"], types={"max_x": "pec", "min_y": "pmc"})
    lossy["boxes"] = lossy["boxes"] + [
        dict(slc=((14, 24), (8, 20), (4, 12)), eps=2.0, sigma_E=(3e4, 1e4, 2e4), sigma_H=2e9, mu=(1.5, 2.0, 2.5))]
    per = W.scenario_c2(T=50, shape=(48, 24, 20), pml=6, slab=(16, 12, 8))        # periodic y, z
    per["pml"] = dict(per["pml"], types={"min_y": "periodic", "max_y": "periodic", "min_z": "periodic",
                                         "max_z": "periodic"})
    return {"cube_aligned": c1, "plane_unaligned": c2, "dispersive": c4, "lossy_walls_mu": lossy, "periodic": per}


@pytest.mark.parametrize("case", ["cube_aligned", "plane_unaligned", "dispersive", "lossy_walls_mu", "periodic"])
def test_quad_sweeps_equal_one_cell_sweeps_bitwise(case, monkeypatch):
    """Four z cells per lane (dwordx4 loads and stores, z neighbours from the lane's registers,
    quad CPML) against one cell per lane: fields, psi, the interface record and every detector
    state agree bit for bit, with z slabs on and inside quad boundaries, ADE, conductivity, PEC /
    PMC faces, magnetic materials and periodic axes."""
    import torch

    from fdtdx_amd import fdtd as F
    import scenarios

    spec = _quad_cases()[case]
    assert spec["shape"][2] % 4 == 0
    out = {}
    for mode in ("two", "quad"):
        _set_mode(monkeypatch, mode)
        objects, arrays, params, config = scenarios.build(spec)
        rec = 0 if case == "dispersive" else spec["T"]
        s = F.Solve(objects, config, arrays, record_steps=rec)
        assert s.plan.quad == (mode == "quad")
        arrays.reset()
        print(f"case {case} mode {mode}", flush=True)
        s.forward(0, spec["T"], record=bool(rec))
        torch.cuda.synchronize()
        out[mode] = [arrays.fields.E.cpu().numpy(), arrays.fields.H.cpu().numpy(), arrays.psi.cpu().numpy()]
        out[mode] += [s.record.cpu().numpy()] if rec else []
        out[mode] += [t.cpu().numpy() for t in s.det_tensors()]
    assert np.abs(out["two"][0]).max() > 0
    for a, b in zip(out["quad"], out["two"]):
        assert np.array_equal(a, b)


def test_gaussian_mode_overlap_matches_oracle():
    """GaussianModeOverlapDetector (detectors/mode.py:666-798): the engine's 6-component phasor
    plane and the overlap with the diverging Gaussian reference at three frequencies against the
    oracle's phasors and its own overlap (mode_overlap, unit-flux gaussian_mode)."""
    from scenarios import mode_spec, oracle_modes

    spec = mode_spec()
    objects, arrays = _run(spec)
    det = next(o for o in objects.detectors if o.name == "mode")
    ov = det.compute_overlap(arrays.detector_states["mode"]).cpu().numpy()
    st, ds = O.run(S.build_scene(spec), spec["T"])
    ph = ds[2]
    assert rel(arrays.detector_states["mode"]["phasor"].cpu().numpy(), ph) < 1e-5
    ref = np.array([O.mode_overlap(ph[0, f], mE, mH, 0) for f, (mE, mH) in enumerate(oracle_modes(spec))])
    print("mode overlaps", ov, ref, flush=True)
    assert np.abs(ref).min() > 0
    assert rel(ov, ref) < 1e-5


def test_mode_source_and_mode_overlap_match_oracle():
    """ModePlaneSource (mode.py:21-306) and the solver-based ModeOverlapDetector
    (detectors/mode.py:448-597) in a strip waveguide: the engine launches the fundamental mode of
    its own solver; the oracle injects the same mode (solved on its own permittivity raster) through
    its TFSF plane tables. E, H, both flux traces, the phasor planes and the overlaps at three
    frequencies agree within 1e-5; the launch is one-way (backward flux below 1 % of forward) and
    the mode carries over 97 % of the phasor power at the overlap plane."""
    from scenarios import oracle_solver_modes, waveguide_spec, with_oracle_mode

    spec = waveguide_spec(T=600)
    objects, arrays = _run(spec)
    det = next(o for o in objects.detectors if o.name == "mode")
    ov = det.compute_overlap(arrays.detector_states["mode"]).cpu().numpy()
    ospec = with_oracle_mode(spec)
    st, ds = O.run(S.build_scene(ospec), spec["T"])
    assert rel(arrays.fields.E.cpu().numpy(), st.E) < 1e-5
    assert rel(arrays.fields.H.cpu().numpy(), st.H) < 1e-5
    back = arrays.detector_states["back"]["poynting_flux"].cpu().numpy()
    fwd = arrays.detector_states["flux"]["poynting_flux"].cpu().numpy()
    assert rel(back, ds[0]) < 1e-5 and rel(fwd, ds[1]) < 1e-5
    assert rel(arrays.detector_states["mode"]["phasor"].cpu().numpy(), ds[2]) < 1e-5
    ref = np.array([O.mode_overlap(ds[2][0, f], mE, mH, 0) for f, (mE, mH) in enumerate(oracle_solver_modes(spec))])
    print("mode overlaps", ov, ref, "back/fwd", back.sum() / fwd.sum(), flush=True)
    assert rel(ov, ref) < 1e-5
    assert abs(back.sum()) < 0.01 * fwd.sum()
    p = ds[3][0, 0]  # Ey, Ez, Hy, Hz phasors of the carrier frequency on the core's plane
    power = 0.5 * np.real(np.sum(p[0] * np.conj(p[3]) - p[1] * np.conj(p[2])))
    assert abs(ov[1]) ** 2 > 0.97 * power


def test_bloch_phases_match_oracle():
    """Complex fields with Bloch phases on x and y (bloch.py:61-97): the real and imaginary parts'
    sweeps couple across the periodic faces; every detector kind reads the complex fields,
    stencils crossing the faces included."""
    from scenarios import bloch_spec

    spec = bloch_spec()
    objects, arrays = _run(spec)
    assert arrays.fields.E.is_complex()
    sc = S.build_scene(spec)
    assert sc.complex_fields and sc.periodic == (True, True, False)
    st, ds = O.run(sc, spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H), flush=True)
    assert np.abs(st.E.imag).max() > 1e-2 * np.abs(st.E).max()
    assert rel(E, st.E) < 1e-5 and rel(H, st.H) < 1e-5
    keys = [("flux", "poynting_flux"), ("ph", "phasor"), ("en", "energy"), ("phr", "phasor"), ("fld", "fields"),
            ("fl3", "poynting_flux")]
    for (name, key), ref in zip(keys, ds):
        got = arrays.detector_states[name][key].cpu().numpy()
        print(name, rel(got, ref), flush=True)
        assert got.shape == ref.shape and rel(got, ref) < 1e-5, name


@pytest.mark.parametrize("magnetic", [False, True])
def test_full_conductivity_tensor_matches_oracle(magnetic):
    """Full electric (and magnetic) conductivity tensors on a full-tensor cube and on a diagonal-
    permittivity box (whose inverse permittivity the engine expands to nine components): per-cell
    A and B matrices, A avg(F) +- B avg(K) (update.py:356-492, 752-822)."""
    from scenarios import tensor_loss_spec

    spec = tensor_loss_spec(magnetic=magnetic)
    objects, arrays = _run(spec)
    assert arrays.electric_conductivity.shape[0] == 9
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H), flush=True)
    assert rel(E, st.E) < 1e-5 and rel(H, st.H) < 1e-5
    for q, name in enumerate(("ez", "flux")):
        got = next(iter(arrays.detector_states[name].values())).cpu().numpy()
        print(name, rel(got, ds[q]), flush=True)
        assert rel(got, ds[q]) < 1e-5


@pytest.mark.parametrize("sigma", [False, True])
def test_oriented_and_tensor_ade_match_oracle(sigma):
    """ADE poles in the full-tensor update (update.py:361-423): an oriented Lorentz pole (nine-
    component coupling with the symmetrized four-point means) in a diagonal-permittivity box that
    the engine expands to a full tensor, touching a full-tensor box with an isotropic Lorentz pole
    and an oriented Drude pole; optionally a full conductivity tensor on the second box."""
    from scenarios import oriented_ade_spec

    spec = oriented_ade_spec(sigma=sigma)
    objects, arrays = _run(spec)
    assert arrays.inv_permittivities.shape[0] == 9
    st, ds = O.run(S.build_scene(spec), spec["T"])
    E = arrays.fields.E.cpu().numpy()
    H = arrays.fields.H.cpu().numpy()
    print("E", rel(E, st.E), "H", rel(H, st.H), flush=True)
    assert rel(E, st.E) < 1e-5 and rel(H, st.H) < 1e-5
    for q, name in enumerate(("ez", "flux")):
        got = next(iter(arrays.detector_states[name].values())).cpu().numpy()
        print(name, rel(got, ds[q]), flush=True)
        assert rel(got, ds[q]) < 1e-5
