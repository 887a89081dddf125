"""Host-side checks that need no GPU: the C ABI library loads and exports every declared entry
point, descriptor validation errors, placement, and the lowering of placed objects into engine
tables (compared with the oracle's independently derived tables)."""

import ctypes
import dataclasses
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import fdtdx_amd as fx
from fdtdx_amd import engine
from fdtdx_amd import fdtd as F
from fdtdx_amd import scenarios as W
from oracle import fdtd_oracle as O
from oracle import scenes as S

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    txt = (ROOT / "include" / "fdtdx_hip.h").read_text()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(fdtdx_\w+)\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    if not engine.LIB_PATH.exists():
        pytest.skip("libfdtdx_hip.so not built")
    lib = ctypes.CDLL(str(engine.LIB_PATH))
    syms = _declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(engine.EXPORTED)


def test_plan_create_validates_before_touching_the_gpu():
    if not engine.LIB_PATH.exists():
        pytest.skip("libfdtdx_hip.so not built")
    lib = engine.load_library()
    h = ctypes.c_void_p()
    assert lib.fdtdx_plan_create(None, ctypes.byref(h)) == 1
    assert b"null" in lib.fdtdx_last_error()
    d = engine.PlanDesc()
    d.nx, d.ny, d.nz, d.num_steps, d.eps_comps = 0, 4, 4, 1, 1
    assert lib.fdtdx_plan_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d.nx, d.eps_comps = 4, 5
    assert lib.fdtdx_plan_create(ctypes.byref(d), ctypes.byref(h)) == 3
    assert b"eps_comps" in lib.fdtdx_last_error()


def _compare_lowering(spec):
    objects, arrays, params, config = W.build(spec)
    plan = F.lower(objects, config, arrays)
    sc = S.build_scene(spec)
    assert np.array_equal(arrays.inv_permittivities.cpu().numpy(), sc.inv_eps)
    for s, sl in zip(plan["cpml"], sc.slabs):
        assert (s["axis"], s["lo"], s["hi"]) == (sl.axis, sl.lo, sl.hi)
        for k, i in (("b_E", 0), ("a_E", 1), ("b_H", 3), ("a_H", 4)):
            assert np.array_equal(s[k], sl.coeffs[i]), k
    srcs = [s for s in sc.sources]
    assert len(plan["sources"]) == len(srcs)
    for a, b in zip(plan["sources"], srcs):
        if b["kind"] == "dipole":
            assert a["kind"] == "point" and tuple(a["lo"]) == b["cell"] and a["component"] == b["component"]
            assert np.array_equal(a["table"], b["table"])
        else:
            assert a["field"] == (0 if b["field"] == "E" else 1)
            assert np.array_equal(a["table"], b["amp"])
            assert np.allclose(a["profile"].reshape(3, -1), b["profile"].reshape(3, -1), rtol=2e-7, atol=0)
            for t, (tc, sc_, sg) in enumerate(b["terms"]):
                assert (a["target"][t], a["src_comp"][t], a["sign"][t]) == (tc, sc_, sg * b["sign"])
    for d_plan, d_or in zip(plan["detectors"], sc.detectors):
        assert np.array_equal(d_plan["time_to_idx"], d_or["time_to_idx"])
        assert d_plan["reduce"] == int(d_or.get("reduce_volume", d_plan["kind"] == 2))
        if d_plan["kind"] == 3:
            assert np.array_equal(d_plan["cos"], d_or["cos"]) and np.array_equal(d_plan["sin"], d_or["sin"])
            assert np.array_equal(d_plan["window"], d_or["window"])
            assert np.float32(d_plan["scale"]) == np.float32(d_or["scale"])
    return plan, sc


def test_lowering_dipole_cube_matches_oracle_tables():
    _compare_lowering(W.scenario_c1(T=80, n=28, pml=6, cube=6))


def test_lowering_detector_options_matches_oracle():
    """Apodization windows, reduce_volume and keep_all_components lower to the oracle's tables,
    and the detector states take the reference's shapes."""
    from scenarios import detector_zoo_spec

    spec = detector_zoo_spec()
    plan, sc = _compare_lowering(spec)
    objects, arrays, params, config = W.build(spec)
    for d_or in sc.detectors:
        st = arrays.detector_states[d_or["name"]]
        assert tuple(next(iter(st.values())).shape) == O.init_detector_state(d_or, sc).shape, d_or["name"]
    fl3 = next(d for d in plan["detectors"] if d["kind"] == 2 and d["components"])
    assert tuple(fl3["components"]) == (0, 1, 2)


def test_lowering_inverse_detectors():
    """Backward-time detectors lower with the inverse flag (recorded only by the time-reversed
    run, update.py:1073), an inverse phasor with sign -1 (phasor.py:226-229), and their tables
    match the oracle's."""
    from scenarios import backward_spec

    spec = backward_spec()
    plan, sc = _compare_lowering(spec)
    inv = {d["name"]: d["inverse"] for d in plan["detectors"]}
    assert inv == {"fwd": 0, "ifld": 1, "iph": 1, "ien": 1}
    assert next(d for d in plan["detectors"] if d["name"] == "iph")["sign"] == -1.0


def test_lowering_surface_detectors():
    """A closed surface lowers to one reduced flux detector per face (max, then min, per active
    axis; face boxes one cell thick; signs +-1, flipped for "inward"; the later faces
    accumulate into the first one's state); a closed phasor surface to one phasor detector per face
    plane with its own state entry and unit step weights."""
    from scenarios import surface_zoo_spec

    spec = surface_zoo_spec()
    objects, arrays, params, config = W.build(spec)
    plan = F.lower(objects, config, arrays)
    cf = [d for d in plan["detectors"] if d["name"] == "cf"]
    assert [(d["axis"], d["sign"], d["accumulate"]) for d in cf] == [
        (0, 1.0, 0), (0, -1.0, 1), (1, 1.0, 1), (1, -1.0, 1), (2, 1.0, 1), (2, -1.0, 1)]
    assert [(d["lo"], d["hi"]) for d in cf[:2]] == [((10, 8, 7), (11, 15, 14)), ((6, 8, 7), (7, 15, 14))]
    assert all(d["reduce"] == 1 and d["key"] == "poynting_flux" for d in cf)
    cp = [d for d in plan["detectors"] if d["name"] == "cp"]
    assert [d["key"] for d in cp] == ["phasor_axis0_min", "phasor_axis0_max", "phasor_axis2_min", "phasor_axis2_max"]
    assert all(np.array_equal(d["window"], (d["time_to_idx"] >= 0).astype(np.float32)) for d in cp)
    assert tuple(arrays.detector_states["cp"]["phasor_axis2_max"].shape) == (1, 1, 6, 9, 8, 1)


@pytest.mark.parametrize("periodic", [False, True])
def test_lowering_full_tensor_matches_oracle(periodic):
    """Off-diagonal permittivity and permeability lower to 9-component inverse tensors (the float32
    3x3 inverse, row-major), and a dipole in a tensor cell to one point source per component of
    column p of the inverse (dipole.py:15-37, 219-234)."""
    from scenarios import tensor_spec

    spec = tensor_spec(periodic=periodic)
    objects, arrays, params, config = W.build(spec)
    plan = F.lower(objects, config, arrays)
    sc = S.build_scene(spec)
    assert plan["eps_comps"] == 9 and plan["mu_comps"] == 9
    assert np.allclose(arrays.inv_permittivities.cpu().numpy(), sc.inv_eps, rtol=1e-6, atol=1e-7)
    assert np.allclose(arrays.inv_permeabilities.cpu().numpy(), sc.inv_mu, rtol=1e-6, atol=1e-7)
    pts = [s for s in plan["sources"] if s["kind"] == "point"]
    assert [s["component"] for s in pts] == [s["component"] for s in sc.sources]
    for a, b in zip(pts, sc.sources):
        assert tuple(a["lo"]) == b["cell"]
        assert np.allclose(a["table"], b["table"], rtol=1e-5, atol=1e-30)


def test_full_tensor_rejections():
    """Full tensors run without conductivity, poles, plane sources and energy detectors: each
    raises NotImplementedError before the engine runs."""
    from scenarios import TENSOR_EPS

    spec = W.scenario_c2(T=20, shape=(30, 20, 16), pml=4, slab=(8, 8, 4))
    spec["boxes"] = [dict(slc=((12, 16), (6, 12), (6, 10)), eps=TENSOR_EPS)]
    with pytest.raises(NotImplementedError, match="full-tensor"):
        W.build(spec)


def test_lowering_symmetry_reduction_matches_oracle():
    """fdtd/symmetry.py:84-264 on the host: the kept quarter, the PEC face on the electric plane,
    CPML only on the far faces, the mirror flag, clipped detectors and design, and the plane
    source normalised over the full plane (linear_polarization.py:295-306) equal the oracle's
    restatement of the reduction."""
    from scenarios import symmetry_spec

    for sym, pol in (((0, -1, 1), (0, 1, 0)), ((0, 1, -1), (0, 0, 1))):
        spec = symmetry_spec(symmetry=sym, e_pol=pol)
        objects, arrays, params, config = W.build(spec)
        plan = F.lower(objects, config, arrays)
        sc = S.build_scene(spec)
        assert tuple(plan["shape"]) == sc.shape == (36, 12, 10)
        assert tuple(bool(m) for m in plan["mirror"]) == sc.mirror
        assert [(w["kind"], w["axis"], w["index"]) for w in plan["walls"]] == [(0, a, i) for _, a, i in sc.walls]
        assert np.array_equal(arrays.inv_permittivities.cpu().numpy(), sc.inv_eps)
        assert [(c["axis"], c["lo"], c["hi"]) for c in plan["cpml"]] == [(x.axis, x.lo, x.hi) for x in sc.slabs]
        assert [tuple(d["lo"]) for d in plan["detectors"]] == [tuple(x.start for x in d["slc"]) for d in sc.detectors]
        src = next(x for x in plan["sources"] if x["kind"] == "plane")
        assert np.allclose(src["profile"].reshape(3, -1), sc.sources[0]["profile"].reshape(3, -1), rtol=2e-7)
        d = next(o for o in objects.detectors if o.name == "flux")
        assert d.straddled_symmetry_axes == tuple(a for a in range(3) if sym[a])


def test_symmetry_rejections():
    """A dipole in the first kept cell sits on the plane (dipole.py:110-128); an odd extent cannot
    be halved; periodic faces on a symmetric axis exclude the mirror."""
    spec = W.scenario_c1(T=10, n=20, pml=4, cube=4)
    spec["symmetry"] = (0, 0, -1)
    with pytest.raises(ValueError, match="symmetry"):
        W.build(spec)  # the dipole is at the centre cell, the first kept one
    spec = W.scenario_c1(T=10, n=21, pml=4, cube=4)
    spec["symmetry"] = (-1, 0, 0)
    with pytest.raises(ValueError, match="even"):
        W.build(spec)


def test_unfold_fields_electric_plane_known_answer():
    """unfold_fields (fdtd/symmetry.py:319-366) across an electric y plane at full index m: a field
    whose components are smooth even or odd functions of y - m, sampled at their Yee positions
    (tangential E and normal H at integer y, the rest half a cell off), is rebuilt from the kept
    half, except the outermost on-plane row, which repeats its neighbour."""
    import torch

    from fdtdx_amd.symmetry import component_sits_on_plane, field_component_parity

    ny, m = 12, 6
    for ft in ("E", "H"):
        full = np.zeros((3, 2, ny, 3))
        for c in range(3):
            off = 0.0 if component_sits_on_plane(ft, c, 1) else 0.5
            y = np.arange(ny) + off - m
            par = field_component_parity(ft, c, 1, -1)
            g = np.cos(0.3 * y) + 0.1 * y**2
            full[c] = (g if par > 0 else np.sin(0.7 * y) * g)[None, :, None]
        got = fx.unfold_fields(torch.tensor(full[:, :, m:]), (0, -1, 0), ft).numpy()
        for c in range(3):
            rows = slice(1, None) if component_sits_on_plane(ft, c, 1) else slice(None)
            assert np.allclose(got[c][:, rows], full[c][:, rows], rtol=0, atol=1e-12), (ft, c)


def test_unfold_matches_the_full_run():
    """unfold_fields and unfold_detector_states on the oracle's reduced run against its full run,
    away from the CPML (where the mirrored far-face layer is not the discarded near-face one), for
    one electric or magnetic plane: the driven field component and the snapshot within 3 %, the
    phasors within 2 %, the flux within 5 %. That is the discretisation-level agreement of the
    reduction: per-cell rasterisation breaks the node mirror symmetry at material interfaces, where
    the small odd components arise (their index maps and parities are pinned exactly by the
    known-answer test above)."""
    import torch

    from scenarios import full_spec, symmetry_spec

    rel = lambda a, b: float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))
    keys = {"flux": "poynting_flux", "snap": "fields", "ph": "phasor"}
    inner = (Ellipsis, slice(5, -5), slice(5, -5))
    for sym, pol in (((0, -1, 0), (0, 1, 0)), ((0, 0, -1), (0, 0, 1)), ((0, 1, 0), (0, 0, 1))):
        main = pol.index(1)
        spec = symmetry_spec(symmetry=sym, e_pol=pol, T=70)
        sr, dr = O.run(S.build_scene(spec, np.float64), spec["T"])
        sf, df = O.run(S.build_scene(full_spec(spec), np.float64), spec["T"])
        E = fx.unfold_fields(torch.tensor(sr.E), sym, "E").numpy()
        assert E.shape == sf.E.shape and rel(E[main][inner], sf.E[main][inner]) < 3e-2
        objects, arrays, params, config = W.build(spec)
        arrays.detector_states = {n: {keys[n]: torch.tensor(v)} for n, v in zip(("flux", "snap", "ph"), dr)}
        out = fx.unfold_detector_states(arrays, objects, config).detector_states
        assert rel(out["flux"]["poynting_flux"].numpy(), df[0]) < 5e-2
        assert rel(out["snap"]["fields"].numpy()[:, main][inner], df[1][:, main][inner]) < 3e-2
        assert rel(out["ph"]["phasor"].numpy()[inner], df[2][inner]) < 2e-2


def test_phasor_stride_and_window_checks():
    """phasor.py:92-103: dft_subsample="auto" keeps 12 samples per period of the highest
    frequency; a window that sums to zero over the recorded steps is an error (phasor.py:148-156)."""
    import fdtdx_amd as fx

    spec = W.scenario_c2(T=60, shape=(50, 30, 24), pml=5, slab=(16, 12, 4))
    objects, arrays, params, config = W.build(spec)
    ph = next(o for o in objects.detectors if isinstance(o, fx.PhasorDetector))
    f = max(w.get_frequency() for w in ph.wave_characters)
    dt = config.time_step_duration
    auto = fx.PhasorDetector(wave_characters=ph.wave_characters, dft_subsample="auto")
    assert auto.dft_stride(config) == max(1, int(1.0 / (12 * f * dt)))
    with pytest.raises(Exception, match="Invalid dft_subsample"):
        fx.PhasorDetector(wave_characters=ph.wave_characters, dft_subsample="fast").dft_stride(config)
    spec["detectors"][1]["apodization"] = dict(kind="tukey", start_time=2 * 60 * dt, end_time=3 * 60 * dt)
    objects, arrays, params, config = W.build(spec)
    with pytest.raises(Exception, match="sums to 0"):
        F.lower(objects, config, arrays)
    with pytest.raises(TypeError):
        dataclasses.replace(ph, apodization="hann").state_layout(config)


def test_lowering_plane_wave_matches_oracle_tables():
    _compare_lowering(W.scenario_c2(T=60, shape=(50, 30, 24), pml=5, slab=(16, 12, 4)))


def test_lowering_dispersive_matches_oracle_coefficients():
    spec = W.scenario_c4(T=40, n=36, pml=5, slab=(14, 14, 6), cube=6)
    plan, sc = _compare_lowering(spec)
    c1, c2, c3 = sc.ade
    for box in plan["ade"]:
        sl = (slice(None), slice(None), *[slice(l, h) for l, h in zip(box["lo"], box["hi"])])
        shp = c1[sl].shape
        assert np.array_equal(box["c1"].reshape(shp), c1[sl])
        assert np.array_equal(box["c3"].reshape(shp), c3[sl])
    # every dispersive cell of the oracle lies inside some box
    mask = np.zeros(spec["shape"], bool)
    for box in plan["ade"]:
        mask[tuple(slice(l, h) for l, h in zip(box["lo"], box["hi"]))] = True
    assert not np.any((c3[:, 0] != 0).any(axis=0) & ~mask)


def test_boundary_placement_and_interfaces():
    cfg = fx.SimulationConfig(time=1e-15, grid=fx.UniformGrid(spacing=50e-9))
    vol = fx.SimulationVolume(partial_grid_shape=(30, 20, 10), name="vol")
    bd, bc = fx.boundary_objects_from_config(fx.BoundaryConfig.from_uniform_bound(thickness=4), vol)
    cube = fx.UniformMaterialObject(partial_grid_shape=(6, 6, 4), material=fx.Material(permittivity=2.0), name="c")
    cons = bc + [cube.place_at_center(vol)]
    objects, arrays, params, config, _ = fx.place_objects([vol, *bd.values(), cube], cfg, cons)
    assert objects["min_x"].grid_slice_tuple == ((0, 4), (0, 20), (0, 10))
    assert objects["max_z"].grid_slice_tuple == ((0, 30), (0, 20), (6, 10))
    assert objects["c"].grid_slice_tuple == ((12, 18), (7, 13), (3, 7))
    assert objects["min_y"].interface_slice_tuple() == ((0, 30), (3, 4), (0, 10))
    assert objects["max_y"].interface_slice_tuple() == ((0, 30), (16, 17), (0, 10))
    assert [p.name for p in objects.pml_objects] == ["min_x", "max_x", "min_y", "max_y", "min_z", "max_z"]


def test_error_paths_match_reference_behaviour():
    with pytest.raises(Exception, match="Need Recorder"):
        fx.GradientConfig(method="reversible")
    with pytest.raises(Exception, match="Checkpoint Number"):
        fx.GradientConfig(method="checkpointed")
    spec = W.scenario_c4(T=10, n=30, pml=4, slab=(10, 10, 4), cube=4)
    with pytest.raises(NotImplementedError, match="Dispersive time-reversible"):
        W.build(spec, gradient=True)
    with pytest.raises(ValueError, match="omega_0"):
        fx.materials.compute_pole_coefficients_per_axis(
            [fx.LorentzPole(resonance_frequency=1e17, damping=0.0, delta_epsilon=1.0)], 1e-16)


def test_engine_has_no_cpu_fallback():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    spec = W.scenario_c1(T=5, n=12, pml=3, cube=2)
    objects, arrays, params, config = W.build(spec)
    with pytest.raises((engine.EngineError, RuntimeError)):
        fx.run_fdtd(arrays, objects, config)


def test_bench_rejects_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_reversible_schedule_slices_and_record_chunks():
    from fdtdx_amd.reversible import Schedule, slice_boundaries

    assert slice_boundaries(3000, 3) == [0, 750, 1500, 2250, 3000]
    s = Schedule(100, 0, 1000, None)  # no budget: the whole record is kept
    assert s.stored and s.keep == [] and s.chunks(0, 100) == [(0, 100)]
    s = Schedule(100, 2, 1000, 10 ** 9)  # fits: checkpoints only at the slice cuts
    assert s.stored and s.bounds == [0, 33, 67, 100] and s.keep == [33, 67]
    s = Schedule(100, 2, 1000, 20_000)  # 20 steps of record: chunks inside each slice
    assert not s.stored and s.chunk_steps == 20
    assert s.chunks(67, 100) == [(87, 100), (67, 87)]
    assert s.chunks(0, 33) == [(20, 33), (0, 20)]
    assert set(s.bounds[1:-1]) <= set(s.keep) and 0 not in s.keep
    covered = sorted(c for a, b in zip(s.bounds[:-1], s.bounds[1:]) for c in s.chunks(a, b))
    assert covered[0][0] == 0 and covered[-1][1] == 100
    assert all(x[1] == y[0] for x, y in zip(covered, covered[1:]))
    with pytest.raises(Exception, match="num_checkpoints_reversible"):
        Schedule(5, 5, 1000, None)


def test_lowering_conductor_faces_matches_oracle_walls():
    spec = W.scenario_c1(T=30, n=20, pml=4, cube=4)
    spec["pml"] = dict(spec["pml"], types={"max_x": "pec", "min_y": "pmc", "max_z": "pec"})
    objects, arrays, params, config = W.build(spec)
    plan = F.lower(objects, config, arrays)
    sc = S.build_scene(spec)
    assert [("pec" if w["kind"] == 0 else "pmc", w["axis"], w["index"]) for w in plan["walls"]] == sc.walls
    assert [(c["axis"], c["lo"], c["hi"]) for c in plan["cpml"]] == [(s.axis, s.lo, s.hi) for s in sc.slabs]
    assert [o.name for o in objects.boundary_objects] == ["min_x", "max_x", "min_y", "max_y", "min_z", "max_z"]
    bc = fx.BoundaryConfig.from_uniform_bound(thickness=4, override_types={"min_x": "bloch", "max_x": "bloch"},
                                              bloch_vector=(1e6, 0.0, 0.0))
    assert bc.bloch_vector == (1e6, 0.0, 0.0)  # a nonzero Bloch phase is accepted (complex fields)


def test_lowering_periodic_axes_matches_oracle():
    """Periodic faces become one-cell BlochBoundary layers; the lowered plan marks their axes
    periodic and drops the CPML there, as the oracle's scene does."""
    spec = W.scenario_c1(T=30, n=20, pml=4, cube=4)
    spec["pml"] = dict(spec["pml"], types={"min_x": "periodic", "max_x": "periodic", "min_z": "bloch",
                                           "max_z": "bloch"})
    objects, arrays, params, config = W.build(spec)
    plan = F.lower(objects, config, arrays)
    sc = S.build_scene(dict(spec, pml=dict(spec["pml"], types={"min_x": "periodic", "max_x": "periodic",
                                                                 "min_z": "periodic", "max_z": "periodic"})))
    assert plan["periodic"] == (1, 0, 1) and tuple(map(int, sc.periodic)) == (1, 0, 1)
    assert [(c["axis"], c["lo"], c["hi"]) for c in plan["cpml"]] == [(s.axis, s.lo, s.hi) for s in sc.slabs]
    assert all(isinstance(objects[f], fx.BlochBoundary) for f in ("min_x", "max_x", "min_z", "max_z"))
    assert objects["min_x"].grid_shape == (1, 20, 20)
    # a PML on a periodic axis has no counterpart in the engine
    spec["pml"] = dict(spec["pml"], types={"min_x": "periodic"})
    objects, arrays, params, config = W.build(spec)
    with pytest.raises(NotImplementedError, match="periodic axis"):
        F.lower(objects, config, arrays)


def test_lowering_conductivity_boxes_match_oracle():
    """Electric and magnetic conductivity painted in placement order (initialization.py:892-930),
    scaled by c0*dt/courant and lowered to the smallest box holding every conducting cell."""
    spec = W.scenario_c1(T=30, n=20, pml=4, cube=4)
    spec["boxes"] = [dict(slc=((3, 9), (4, 12), (5, 15)), eps=2.0, sigma_E=(4e4, 1e4, 2e4), sigma_H=3e9),
                     dict(slc=((5, 7), (6, 8), (7, 9)), eps=3.0),  # paints zero conductivity over the first
                     dict(slc=((10, 13), (2, 5), (11, 14)), eps=1.5, sigma_E=(5e3, 5e3, 6e3))]
    objects, arrays, params, config = W.build(spec)
    plan = F.lower(objects, config, arrays)
    sc = S.build_scene(spec)
    for key, ref in (("sigma_E", sc.sigma_E), ("sigma_H", sc.sigma_H)):
        box = plan[key]
        full = np.zeros_like(ref)
        full[(slice(None), *(slice(l, h) for l, h in zip(box["lo"], box["hi"])))] = box["sigma"]
        assert np.array_equal(full, ref), key
    assert plan["sigma_E"]["lo"] == (3, 2, 5) and plan["sigma_E"]["hi"] == (13, 12, 15)
    assert plan["sigma_H"]["lo"] == (3, 4, 5) and plan["sigma_H"]["hi"] == (9, 12, 15)
    assert np.array_equal(arrays.electric_conductivity.cpu().numpy(), sc.sigma_E)
    # no conductor: no arrays, no boxes
    objects, arrays, params, config = W.build(W.scenario_c1(T=30, n=20, pml=4, cube=4))
    plan = F.lower(objects, config, arrays)
    assert arrays.electric_conductivity is None and plan["sigma_E"] is None and plan["sigma_H"] is None


def test_recorder_lowering_and_slot_schedule():
    """Recorder modules lower to per-layer formats, the filter position and the time filter; the
    record schedule under a budget covers every reversed step with chunks whose slots fit the
    budget and whose regeneration runs start at kept steps (time_filter.py:163-257)."""
    import types

    from fdtdx_amd.recorder import DtypeConversion, LinearReconstructEveryK, Recorder
    from fdtdx_amd.reversible import Schedule
    from oracle import record_codec as R

    rec = Recorder(modules=[LinearReconstructEveryK(k=5), DtypeConversion(dtype="float8_e4m3fnuz",
                                                                         exclude_filter=["max_x"])])
    d = rec.lower(["min_x", "max_x"])
    assert d == dict(every_k=5, start=0, filter_first=1, format_E=[4, 0], format_H=[4, 0])
    import torch
    d = Recorder(modules=[DtypeConversion(dtype=torch.bfloat16), LinearReconstructEveryK(k=3, start_recording_after=2)]
                 ).lower(["min_z"])
    assert d == dict(every_k=3, start=2, filter_first=0, format_E=[1], format_H=[1])
    with pytest.raises(NotImplementedError):
        Recorder(modules=[DtypeConversion(dtype="float64")])
    with pytest.raises(NotImplementedError):
        Recorder(modules=[LinearReconstructEveryK(k=2), LinearReconstructEveryK(k=3)])
    for T, k, start, nck, budget_slots in [(100, 4, 0, 0, 7), (97, 6, 0, 2, 5), (60, 1, 0, 1, 8), (50, 7, 9, 0, 3)]:
        save, tta = R.time_filter_tables(T, k, start)
        slot_of = np.full(T, -1)
        slot_of[save] = np.arange(len(save))
        plan = types.SimpleNamespace(record_slot_bytes=100, slot_of=slot_of, slot_arr=tta.astype(np.int64),
                                     slot_step=np.asarray(save))
        s = Schedule(T, nck, plan, budget_slots * 100)
        assert not s.stored and s.chunk_steps == budget_slots
        covered = []
        for a, b in zip(s.bounds[:-1], s.bounds[1:]):
            for c, dd in s.chunks(a, b):
                r0, r1, first = s.run_range(c, dd)
                assert r0 in s.keep or r0 == 0
                assert r0 in save and first == tta[r0]  # runs start at the step of the first slot
                for n in range(c, dd):  # every slot step n reads is written by the run
                    need = [tta[n]] if slot_of[n] >= 0 else [tta[n], tta[n] + 1]
                    for sl in need:
                        assert r0 <= save[sl] < r1 and 0 <= sl - first < budget_slots
                covered.append((c, dd))
        covered.sort()
        assert covered[0][0] == 0 and covered[-1][1] == T
        assert all(x[1] == y[0] for x, y in zip(covered, covered[1:]))


def test_lowering_plane_source_in_dispersive_medium_matches_oracle():
    """A plane source inside a Lorentz medium: effective permittivity at the carrier for the
    amplitude, impedance and Yee offsets, and the broadband-filtered E-side table
    (linear_polarization.py:191-210, 330-351), against the oracle's own derivation."""
    spec = W.scenario_c2(T=60, shape=(50, 30, 24), pml=5, slab=(16, 12, 4))
    w = 2 * np.pi * 299792458.0 / 1.55e-6
    spec["boxes"] = [dict(slc=((0, 30), (0, 30), (0, 24)), eps=2.1, poles=[("lorentz", 1.4 * w, 0.1 * w, 0.8)])]
    objects, arrays, params, config = W.build(spec)
    plan = F.lower(objects, config, arrays)
    sc = S.build_scene(spec)
    assert sc.ade is not None
    ref = [s for s in sc.sources if s["kind"] == "plane"]
    got = [s for s in plan["sources"] if s["kind"] == "plane"]
    for a, b in zip(got, ref):
        assert np.allclose(a["table"], b["amp"], rtol=1e-5, atol=1e-6 * np.abs(b["amp"]).max())
        assert np.allclose(a["profile"].reshape(3, -1), b["profile"].reshape(3, -1), rtol=1e-6, atol=0)
    # the filtered E-side table differs from the raw pulse
    objects2, arrays2, _, config2 = W.build(dict(spec, boxes=[dict(spec["boxes"][0], poles=[])]))
    raw = [s for s in F.lower(objects2, config2, arrays2)["sources"] if s["kind"] == "plane"]
    assert not np.allclose(got[0]["table"], raw[0]["table"], rtol=1e-3)


def test_gaussian_mode_detector_reference_modes_match_oracle():
    """GaussianModeOverlapDetector's reference modes (evaluated by apply before the run) equal the
    oracle's unit-flux Gaussian cross-sections on the same plane (cell centres about the domain
    centre, the plane's mean index, a diverging wavefront)."""
    from scenarios import mode_spec, oracle_modes

    spec = mode_spec(T=20)
    objects, arrays, params, config = W.build(spec)
    det = next(o for o in objects.detectors if o.name == "mode")
    ref = oracle_modes(spec)
    assert det._mode_E.shape == (3, 3, 1, 40, 32)
    for f, (mE, mH) in enumerate(ref):
        assert np.allclose(det._mode_E[f], mE, rtol=0, atol=1e-12 * np.abs(mE).max())
        assert np.allclose(det._mode_H[f], mH, rtol=0, atol=1e-12 * np.abs(mH).max())
    with pytest.raises(ValueError, match="mode_radius"):
        spec["detectors"][-1]["mode_radius"] = 0.0
        W.build(spec)
