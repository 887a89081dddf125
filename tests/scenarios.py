"""Build a scenario spec (oracle/scenes.py format) through the public fdtdx_amd API, the way a user
of the reference would write it (examples/simulate_gaussian_source.py style)."""

from __future__ import annotations

import numpy as np

import fdtdx_amd as fx


def build(spec, gradient=False):
    shape = tuple(spec["shape"])
    dx = spec["dx"]
    T = spec["T"]
    dt = (spec.get("courant_factor", 0.99) / float(np.sqrt(3.0))) * dx / fx.constants.c
    gc = fx.GradientConfig(method="reversible", recorder=fx.Recorder(modules=[])) if gradient else None
    config = fx.SimulationConfig(time=T * dt, grid=fx.UniformGrid(spacing=dx), gradient_config=gc,
                                 courant_factor=spec.get("courant_factor", 0.99))
    assert config.time_steps_total == T
    objs, cons = [], []
    bg = spec.get("background_eps", 1.0)
    volume = fx.SimulationVolume(partial_grid_shape=shape, material=fx.Material(permittivity=bg), name="volume")
    objs.append(volume)
    th = spec.get("pml", {}).get("thickness", 0)
    if th:
        bcfg = fx.BoundaryConfig.from_uniform_bound(thickness=th, **spec.get("pml", {}).get("params", {}))
        bdict, bcons = fx.boundary_objects_from_config(bcfg, volume)
        objs.extend(bdict.values())
        cons.extend(bcons)
    for i, b in enumerate(spec.get("boxes", [])):
        disp = None
        if b.get("poles"):
            poles = []
            for p in b["poles"]:
                if p[0] == "lorentz":
                    poles.append(fx.LorentzPole(resonance_frequency=p[1], damping=p[2], delta_epsilon=p[3]))
                else:
                    poles.append(fx.DrudePole(plasma_frequency=p[1], damping=p[2]))
            disp = fx.DispersionModel(poles=poles)
        o = fx.UniformMaterialObject(partial_grid_shape=tuple(h - l for l, h in b["slc"]),
                                     material=fx.Material(permittivity=b["eps"], dispersion=disp),
                                     name=f"box{i}", placement_order=i)
        cons.append(o.set_grid_coordinates(axes=(0, 1, 2), sides=("-", "-", "-"),
                                           coordinates=tuple(l for l, _ in b["slc"])))
        objs.append(o)
    d = spec.get("design")
    if d is not None:
        dev = fx.Device(partial_grid_shape=tuple(h - l for l, h in d["slc"]), name="design",
                        materials={"a": fx.Material(permittivity=d["eps_a"]), "b": fx.Material(permittivity=d["eps_b"])})
        cons.append(dev.set_grid_coordinates(axes=(0, 1, 2), sides=("-", "-", "-"),
                                             coordinates=tuple(l for l, _ in d["slc"])))
        objs.append(dev)
    for i, s in enumerate(spec.get("sources", [])):
        prof = s["profile"]
        tp = fx.GaussianPulseProfile(spectral_width=fx.WaveCharacter(frequency=prof["spectral_width"]),
                                     center_wave=fx.WaveCharacter(frequency=prof["center_frequency"]))
        wc = fx.WaveCharacter(frequency=prof["center_frequency"])
        if s["kind"] == "dipole":
            src = fx.PointDipoleSource(partial_grid_shape=(1, 1, 1), polarization=s["polarization"],
                                       wave_character=wc, temporal_profile=tp, name=f"src{i}",
                                       amplitude=s.get("amplitude", 1.0))
            cons.append(src.set_grid_coordinates(axes=(0, 1, 2), sides=("-", "-", "-"), coordinates=tuple(s["cell"])))
        else:
            pg = [None, None, None]
            pg[s["axis"]] = 1
            src = fx.UniformPlaneSource(partial_grid_shape=tuple(pg), direction=s.get("direction", "+"),
                                        fixed_E_polarization_vector=tuple(s["e_pol"]), wave_character=wc,
                                        temporal_profile=tp, name=f"src{i}")
            cons.append(src.set_grid_coordinates(axes=(s["axis"],), sides=("-",), coordinates=(s["index"],)))
        objs.append(src)
    for dsp in spec.get("detectors", []):
        kw = dict(name=dsp["name"], partial_grid_shape=tuple(h - l for l, h in dsp["slc"]))
        sw = dsp.get("switch")
        if sw:
            kw["switch"] = fx.OnOffSwitch(**sw)
        if dsp["kind"] == "energy":
            det = fx.EnergyDetector(reduce_volume=dsp.get("reduce_volume", False), **kw)
        elif dsp["kind"] == "field":
            det = fx.FieldDetector(components=dsp["components"], **kw)
        elif dsp["kind"] == "flux":
            det = fx.PoyntingFluxDetector(direction=dsp.get("direction", "+"), reduce_volume=dsp.get("reduce_volume", True),
                                          **kw)
        else:
            det = fx.PhasorDetector(wave_characters=[fx.WaveCharacter(frequency=f) for f in dsp["frequencies"]],
                                    components=dsp["components"], **kw)
        cons.append(det.set_grid_coordinates(axes=(0, 1, 2), sides=("-", "-", "-"),
                                             coordinates=tuple(l for l, _ in dsp["slc"])))
        objs.append(det)
    objects, arrays, params, config, _ = fx.place_objects(object_list=objs, config=config, constraints=cons)
    if d is not None:
        params = {"design": np.asarray(d["rho"], np.float32)}
        arrays, objects, _ = fx.apply_params(arrays, objects, params)
    return objects, arrays, params, config
