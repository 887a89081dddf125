"""Containers, placement and the time-loop entry points (reference: src/fdtdx/fdtd/container.py,
initialization.py:100-521, fdtd.py:39-584, wrapper.py:14-63).

The reference threads immutable JAX pytrees through a jitted while-loop. Here the arrays are torch
tensors on the ROCm device; run_fdtd advances them in place through the HIP engine and returns
the container (the reference's donate-buffer behaviour). Gradients flow through
``reversible_fdtd`` as a torch.autograd.Function whose backward runs the engine's time-reversed
adjoint.
"""

from __future__ import annotations

import copy
from dataclasses import dataclass, field, replace
from typing import Any, Callable, Sequence

import numpy as np

from . import constants, engine
from .config import SimulationConfig
from .detectors import Detector
from .objects import (BlochBoundary, Device, PerfectElectricConductor, PerfectlyMatchedLayer, SimulationObject,
                      SimulationVolume, UniformMaterialObject, _ConductorFace, resolve_object_constraints)
from .sources import Source


# ---------------------------------------------------------------- containers (container.py:33-491)
class ObjectContainer:
    def __init__(self, object_list: Sequence[SimulationObject], volume_idx: int | None = None):
        self.object_list = list(object_list)
        if volume_idx is None:
            volume_idx = next(i for i, o in enumerate(self.object_list) if isinstance(o, SimulationVolume))
        self.volume_idx = volume_idx
        self._plan_cache = {}

    def filter_objects(self, cls, predicate=None):
        return [o for o in self.object_list if isinstance(o, cls) and (predicate is None or predicate(o))]

    @property
    def volume(self):
        return self.object_list[self.volume_idx]

    @property
    def sources(self):
        return self.filter_objects(Source)

    @property
    def devices(self):
        return self.filter_objects(Device)

    @property
    def detectors(self):
        return self.filter_objects(Detector)

    @property
    def forward_detectors(self):
        return self.filter_objects(Detector, lambda o: not o.inverse)

    @property
    def backward_detectors(self):
        return self.filter_objects(Detector, lambda o: o.inverse)

    @property
    def pml_objects(self):
        return self.filter_objects(PerfectlyMatchedLayer)

    @property
    def conductor_objects(self):
        return self.filter_objects(_ConductorFace)

    @property
    def periodic_objects(self):
        return self.filter_objects(BlochBoundary)

    @property
    def boundary_objects(self):
        return self.filter_objects((PerfectlyMatchedLayer, _ConductorFace, BlochBoundary))

    @property
    def static_material_objects(self):
        return self.filter_objects(UniformMaterialObject)

    def __getitem__(self, name):
        for o in self.object_list:
            if o.name == name:
                return o
        raise KeyError(name)


@dataclass
class FieldState:
    E: Any
    H: Any
    psi_E: dict
    psi_H: dict
    dispersive_P_curr: Any = None
    dispersive_P_prev: Any = None


@dataclass
class ArrayContainer:
    fields: FieldState
    inv_permittivities: Any
    inv_permeabilities: Any
    detector_states: dict
    recording_state: Any = None
    electric_conductivity: Any = None
    magnetic_conductivity: Any = None
    initial_inv_permittivities: Any = None
    # packed engine state (views above alias these)
    psi: Any = None
    ade_state: Any = None
    ade_step: int = 0
    # complex fields: the engine's (2, 3, nx, ny, nz) float32 storage of E and H (real part, then
    # imaginary part); fields.E / fields.H are complex64 copies refreshed after every engine call
    E_parts: Any = None
    H_parts: Any = None

    @property
    def complex_fields(self) -> bool:
        return self.E_parts is not None

    def sync_complex_fields(self):
        """Refresh fields.E / fields.H (complex64) from the engine's real and imaginary parts."""
        import torch

        if self.E_parts is not None:
            self.fields.E = torch.complex(self.E_parts[0], self.E_parts[1])
            self.fields.H = torch.complex(self.H_parts[0], self.H_parts[1])

    def reset(self) -> "ArrayContainer":
        """container.py:451-491: zero fields, psi, ADE state and detector states (in place)."""
        if self.E_parts is not None:
            self.E_parts.zero_()
            self.H_parts.zero_()
        self.fields.E.zero_()
        self.fields.H.zero_()
        if self.psi is not None:
            self.psi.zero_()
        if self.ade_state is not None:
            self.ade_state.zero_()
        self.ade_step = 0
        for st in self.detector_states.values():
            for v in st.values():
                v.zero_()
        return self


ParameterContainer = dict


# ---------------------------------------------------------------- host material rasterisation
def _material_arrays(objects: ObjectContainer, config: SimulationConfig):
    """initialization.py:714-870: inverse permittivity/permeability painted in placement order."""
    vol = objects.volume
    shape = vol.grid_shape
    painters = [vol] + sorted(objects.static_material_objects, key=lambda o: o.placement_order)
    mats = [o.material for o in painters] + [m for d in objects.devices for m in d.materials.values()]
    off = (1, 2, 3, 5, 6, 7)
    # components: 1 isotropic, 3 diagonal, 9 full tensor (initialization.py:714-740)
    # oriented poles take the full-tensor update (initialization.py:708-711)
    oriented = any(getattr(p, "is_oriented", False) for m in mats if m.dispersion is not None
                   for p in m.dispersion.poles)
    full_eps = oriented or any(any(m.permittivity[i] for i in off) for m in mats)
    ec = 9 if full_eps else (1 if all(m.is_isotropic_permittivity for m in mats) else 3)
    magnetic = any(m.is_magnetic for m in mats)
    mc = 0
    if magnetic:
        if any(any(m.permeability[i] for i in off) for m in mats):
            mc = 9
        else:
            mc = 1 if all(m.permeability[0] == m.permeability[4] == m.permeability[8] for m in mats) else 3

    def inverse(t, comps):
        if comps == 9:  # the float32 inverse of the 3x3 tensor, flattened row-major (initialization.py:866-873)
            return np.linalg.inv(np.asarray(t, np.float32).reshape(3, 3)).astype(np.float32).reshape(9)
        vals = (t[0],) if comps == 1 else (t[0], t[4], t[8])
        return np.float32(1) / np.asarray(vals, np.float32)

    inv_eps = np.zeros((ec, *shape), np.float32)
    inv_mu = np.zeros((mc, *shape), np.float32) if mc else 1.0
    for o in painters:
        inv_eps[(slice(None), *o.grid_slice)] = inverse(o.material.permittivity, ec)[:, None, None, None]
        if mc:
            inv_mu[(slice(None), *o.grid_slice)] = inverse(o.material.permeability, mc)[:, None, None, None]
    return inv_eps, inv_mu, ec, mc


def _conductivity_array(objects: ObjectContainer, config: SimulationConfig, attr: str):
    """initialization.py:757-783, 892-930: the electric or magnetic conductivity array, allocated
    only when some material conducts; one component when every material's tensor is isotropic,
    else three; painted in placement order, scaled by c0*dt/courant in float32."""
    vol = objects.volume
    painters = [vol] + sorted(objects.static_material_objects, key=lambda o: o.placement_order)
    mats = [o.material for o in painters] + [m for d in objects.devices for m in d.materials.values()]
    if not any(any(getattr(m, attr)) for m in mats):
        return None
    comps = 1 if all(getattr(m, attr)[0] == getattr(m, attr)[4] == getattr(m, attr)[8] for m in mats) else 3
    if any(any(getattr(m, attr)[t] for t in (1, 2, 3, 5, 6, 7)) for m in mats):
        comps = 9  # some conductivity tensor has off-diagonal entries
    spacing = constants.c * config.time_step_duration / config.courant_number
    sig = np.zeros((comps, *vol.grid_shape), np.float32)
    for o in painters:
        t = getattr(o.material, attr)
        v = (t[0],) if comps == 1 else ((t[0], t[4], t[8]) if comps == 3 else tuple(t))
        sig[(slice(None), *o.grid_slice)] = (np.asarray(v, np.float32) * np.float32(spacing))[:, None, None, None]
    return sig


def _record_desc(config, pml_names):
    """The interface-record compression of the gradient config's Recorder (None: exact)."""
    gc = config.gradient_config
    rec = None if gc is None else gc.recorder
    if rec is None or not getattr(rec, "modules", None):
        return None
    return rec.lower(pml_names)


def _host(a):
    """A host NumPy view of an array that may live on the device (None stays None)."""
    if a is None:
        return None
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _loss_box(sig):
    """The smallest box holding every nonzero conductivity cell, as an engine loss descriptor
    (None when the array is None or zero)."""
    if sig is None:
        return None
    nz = np.nonzero(np.any(sig != 0, axis=0))
    if nz[0].size == 0:
        return None
    lo = tuple(int(a.min()) for a in nz)
    hi = tuple(int(a.max()) + 1 for a in nz)
    return dict(lo=lo, hi=hi, sigma=np.ascontiguousarray(sig[(slice(None), *(slice(l, h) for l, h in zip(lo, hi)))]))


def _dispersive_slice(boxes, region):
    """The ADE coefficients (c1, c2, c3), each (poles, comps, *region), over the cells of `region`
    ((lo, hi) per axis), zero outside the dispersive boxes; None when no pole acts there."""
    if not boxes:
        return None
    shp = tuple(h - l for l, h in region)
    P, C = boxes[0]["num_poles"], boxes[0]["coef_comps"]
    C3 = boxes[0].get("coupling_comps") or C
    out = [np.zeros((P, C, *shp), np.float32), np.zeros((P, C, *shp), np.float32), np.zeros((P, C3, *shp), np.float32)]
    hit = False
    for b in boxes:
        inter = [(max(region[a][0], b["lo"][a]), min(region[a][1], b["hi"][a])) for a in range(3)]
        if any(l >= h for l, h in inter):
            continue
        src = tuple(slice(l - b["lo"][a], h - b["lo"][a]) for a, (l, h) in enumerate(inter))
        dst = tuple(slice(l - region[a][0], h - region[a][0]) for a, (l, h) in enumerate(inter))
        for t, key in enumerate(("c1", "c2", "c3")):
            arr = np.asarray(b[key], np.float32).reshape(P, C3 if key == "c3" else C, *b["shape"])
            out[t][(slice(None), slice(None), *dst)] = arr[(slice(None), slice(None), *src)]
        hit = True
    if not hit or not (np.any(out[0]) or np.any(out[2])):
        return None
    return tuple(out)


def _ade_boxes(objects: ObjectContainer, config: SimulationConfig):
    """Dispersive cells as boxes with per-cell ADE coefficients (initialization.py:948-999): every
    static object paints its (zero-padded) pole coefficients in placement order; boxes of
    overlapping dispersive objects are merged so each cell is updated once."""
    from .materials import compute_pole_coefficients_per_axis, compute_pole_coefficients_tensor

    vol = objects.volume
    painters = [vol] + sorted(objects.static_material_objects, key=lambda o: o.placement_order)
    disp = [o for o in painters if o.material.num_poles > 0]
    if any(m.num_poles for d in objects.devices for m in d.materials.values()):
        raise NotImplementedError("dispersive device materials are not supported by this engine")
    if not disp:
        return [], 0
    P = max(o.material.num_poles for o in disp)
    poles_all = [p for o in disp for p in o.material.dispersion.poles]
    oriented = any(getattr(p, "is_oriented", False) for p in poles_all)
    iso = all(p.is_isotropic or getattr(p, "is_oriented", False) for p in poles_all)
    C = 1 if iso else 3
    C3 = 9 if oriented else C  # coupling tensors of oriented poles (initialization.py:686-690)
    boxes = [list(map(list, o.grid_slice_tuple)) for o in disp]
    # boxes are merged where they overlap, and with coupling tensors also where they touch: the
    # off-diagonal means read the coefficients one cell beyond a cell
    touch = 1 if oriented else 0
    merged = True
    while merged:
        merged = False
        for i in range(len(boxes)):
            for j in range(i + 1, len(boxes)):
                a, b = boxes[i], boxes[j]
                if all(a[x][0] < b[x][1] + touch and b[x][0] < a[x][1] + touch for x in range(3)):
                    boxes[i] = [[min(a[x][0], b[x][0]), max(a[x][1], b[x][1])] for x in range(3)]
                    del boxes[j]
                    merged = True
                    break
            if merged:
                break
    dt = config.time_step_duration
    out = []
    for bx in boxes:
        bshape = tuple(h - l for l, h in bx)
        c1 = np.zeros((P, C, *bshape), np.float32)
        c2 = np.zeros_like(c1)
        c3 = np.zeros((P, C3, *bshape), np.float32)
        for o in painters:
            inter = [(max(bx[x][0], o.grid_slice_tuple[x][0]), min(bx[x][1], o.grid_slice_tuple[x][1])) for x in range(3)]
            if any(l >= h for l, h in inter):
                continue
            sl = (slice(None), slice(None), *[slice(l - bx[x][0], h - bx[x][0]) for x, (l, h) in enumerate(inter)])
            poles = o.material.dispersion.poles if o.material.dispersion is not None else ()
            if C3 == 9:
                k1, k2, k3 = compute_pole_coefficients_tensor(poles, dt)
            else:
                k1, k2, k3 = compute_pole_coefficients_per_axis(poles, dt)
                k3 = k3[:, :C]
            v = [np.zeros((P, C), np.float32), np.zeros((P, C), np.float32), np.zeros((P, C3), np.float32)]
            for t, kk in enumerate((k1[:, :C], k2[:, :C], k3)):
                v[t][: len(poles)] = kk.astype(np.float32)
            c1[sl] = v[0][:, :, None, None, None]
            c2[sl] = v[1][:, :, None, None, None]
            c3[sl] = v[2][:, :, None, None, None]
        out.append(dict(lo=tuple(l for l, _ in bx), hi=tuple(h for _, h in bx), num_poles=P, coef_comps=C,
                        coupling_comps=C3 if C3 == 9 else 0,
                        c1=c1.reshape(-1), c2=c2.reshape(-1), c3=c3.reshape(-1), shape=bshape))
    if len(out) > engine.FDTDX_MAX_ADE:
        raise NotImplementedError(f"more than {engine.FDTDX_MAX_ADE} separate dispersive regions")
    return out, P


def lower(objects: ObjectContainer, config: SimulationConfig, arrays: ArrayContainer) -> dict:
    """Lower the placed objects to an engine plan spec (the fdtdx_plan_desc fields)."""
    shape = objects.volume.grid_shape
    T = config.time_steps_total
    cpml = []
    for p in objects.pml_objects:
        co = p._coefficients
        cpml.append(dict(axis=p.axis, side=0 if p.direction == "-" else 1, lo=p.grid_slice_tuple[p.axis][0],
                         hi=p.grid_slice_tuple[p.axis][1], kappa_is_one=p.kappa_is_one,
                         b_E=co["b_E"], a_E=co["a_E"], ikm1_E=(co["inv_kappa_E"] - np.float32(1)).astype(np.float32),
                         b_H=co["b_H"], a_H=co["a_H"], ikm1_H=(co["inv_kappa_H"] - np.float32(1)).astype(np.float32),
                         name=p.name, shape=p.grid_shape))
        if p.grid_shape[(p.axis + 1) % 3] != shape[(p.axis + 1) % 3] or p.grid_shape[(p.axis + 2) % 3] != shape[(p.axis + 2) % 3]:
            raise NotImplementedError("a PML slab must span the full domain on its two transverse axes")
    sources = []
    for s in objects.sources:
        sources.extend(s.lower(config, None, None))
    if len(sources) > engine.FDTDX_MAX_SOURCES:
        raise NotImplementedError(f"more than {engine.FDTDX_MAX_SOURCES} source hooks")
    dets = []
    for d in objects.forward_detectors + objects.backward_detectors:
        # a detector lowers to one engine detector per state entry, or to a group of them that
        # accumulate into one entry (closed surfaces)
        low = d.lower(config)
        first_key = next(iter(d.state_layout(config)))
        for dd in (low if isinstance(low, list) else [low]):
            dd["name"] = d.name
            dd.setdefault("key", first_key)
            dets.append(dd)
    if len(dets) > engine.FDTDX_MAX_DETECTORS:
        raise NotImplementedError(f"more than {engine.FDTDX_MAX_DETECTORS} detectors")
    ade, _ = _ade_boxes(objects, config)
    ie = arrays.inv_permittivities
    im = arrays.inv_permeabilities
    walls = []
    for w in objects.conductor_objects:
        if w.grid_shape[w.axis] != 1 or any(w.grid_shape[a] != shape[a] for a in range(3) if a != w.axis):
            raise NotImplementedError("a PEC / PMC boundary must be one cell thick and span its face")
        walls.append(dict(kind=0 if isinstance(w, PerfectElectricConductor) else 1, axis=w.axis,
                          index=w.layer_index, name=w.name))
    if len(walls) > engine.FDTDX_MAX_WALLS:
        raise NotImplementedError(f"more than {engine.FDTDX_MAX_WALLS} PEC / PMC faces")
    # update.py:29-45: an axis wraps when any of its boundaries uses wrap padding
    periodic = [0, 0, 0]
    for b in objects.periodic_objects:
        periodic[b.axis] = 1
    for b in objects.pml_objects:
        if periodic[b.axis]:
            raise NotImplementedError(f"{b.name}: a PML on a periodic axis is not supported by this engine")
    for b in objects.conductor_objects:
        if periodic[b.axis]:
            raise NotImplementedError(f"{b.name}: a PEC / PMC face on a periodic axis is not supported by this engine")
    mirror = tuple(int(any(getattr(w, "_is_symmetry_wall", False) and w.axis == a for w in objects.conductor_objects))
                   for a in range(3))
    # Bloch phases k L of each periodic face (get_bloch_phase, bloch.py:140-153)
    bloch = [[0.0, 0.0], [0.0, 0.0], [0.0, 0.0]]
    for b in objects.periodic_objects:
        bloch[b.axis][0 if b.direction == "-" else 1] = b.get_bloch_phase(shape, config.uniform_spacing())
    cplx = use_complex_fields(objects, config)
    ec, mc = engine_material_comps(arrays)
    return dict(complex_fields=int(cplx), bloch_phase=bloch if cplx else [[0.0, 0.0]] * 3,
                mirror=mirror, shape=tuple(shape), T=T, courant=config.courant_number, eps_comps=ec,
                mu_comps=mc, cpml=cpml, sources=sources, ade=ade,
                detectors=dets, walls=walls, periodic=tuple(periodic),
                sigma_E=_loss_box(_host(arrays.electric_conductivity)),
                sigma_H=_loss_box(_host(arrays.magnetic_conductivity)),
                record=_record_desc(config, [p["name"] for p in cpml]))


def engine_material_comps(arrays: ArrayContainer):
    """The (eps_comps, mu_comps) the engine runs: the arrays' own, except that a conductivity
    tensor with off-diagonal entries takes the full-tensor update of its field, whose inverse
    material is then expanded to nine components (expand_to_3x3, update.py:356-359, 755-756)."""
    ie, im = arrays.inv_permittivities, arrays.inv_permeabilities
    ec = int(ie.shape[0])
    mc = 0 if np.isscalar(im) else int(im.shape[0])
    se, sh = arrays.electric_conductivity, arrays.magnetic_conductivity
    if se is not None and se.shape[0] == 9:
        ec = 9
    if sh is not None and sh.shape[0] == 9:
        mc = 9
    return ec, mc


def expand_inverse_material(t, comps: int):
    """An inverse material array (a scalar, or 1, 3 or 9 components) as `comps` components; only
    1 or 3 -> 9 (the diagonal of a row-major tensor) and scalar -> 9 widen."""
    import torch

    if comps != 9:
        return t
    if not torch.is_tensor(t):
        raise ValueError("expand_inverse_material needs a tensor or a scalar with its shape")
    if t.shape[0] == 9:
        return t
    out = torch.zeros((9, *t.shape[1:]), dtype=t.dtype, device=t.device)
    for i in range(3):
        out[4 * i] = t[0 if t.shape[0] == 1 else i]
    return out


def use_complex_fields(objects: ObjectContainer, config: SimulationConfig) -> bool:
    """initialization.py:580-594: complex fields when config.use_complex_fields says so, or (left
    None) when a Bloch boundary has a nonzero Bloch vector along its axis."""
    needs = any(b.needs_complex_fields for b in objects.periodic_objects)
    if config.use_complex_fields is None:
        return needs
    if needs and not config.use_complex_fields:
        raise ValueError("use_complex_fields=False but Bloch boundaries with non-zero wave vector are present. "
                         "These require complex-valued fields.")
    return bool(config.use_complex_fields)


# ---------------------------------------------------------------- placement (initialization.py:100-314)
def _device():
    import torch

    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def _psi_layout(objects):
    out, off = [], 0
    for p in objects.pml_objects:
        cells = int(np.prod(p.grid_shape))
        out.append((p.name, off, tuple(p.grid_shape)))
        off += 4 * cells
    return out, off


def _make_arrays(objects: ObjectContainer, config: SimulationConfig, inv_eps, inv_mu, device):
    import torch

    shape = objects.volume.grid_shape
    cplx = use_complex_fields(objects, config)
    parts = 2 if cplx else 1  # complex fields: every engine array holds the real, then the imaginary part
    E_parts = H_parts = None
    if cplx:
        E_parts = torch.zeros((2, 3, *shape), dtype=torch.float32, device=device)
        H_parts = torch.zeros_like(E_parts)
        E = torch.zeros((3, *shape), dtype=torch.complex64, device=device)
        H = torch.zeros_like(E)
    else:
        E = torch.zeros((3, *shape), dtype=torch.float32, device=device)
        H = torch.zeros_like(E)
    layout, psi_floats = _psi_layout(objects)
    # the psi views below show the real part; the imaginary part follows at psi_floats
    psi = torch.zeros(max(parts * psi_floats, 1), dtype=torch.float32, device=device)
    psi_E, psi_H = {}, {}
    for name, off, shp in layout:
        n = int(np.prod(shp))
        psi_E[name] = (psi[off:off + n].view(shp), psi[off + n:off + 2 * n].view(shp))
        psi_H[name] = (psi[off + 2 * n:off + 3 * n].view(shp), psi[off + 3 * n:off + 4 * n].view(shp))
    ade, P = _ade_boxes(objects, config)
    ade_floats = parts * sum(3 * b["num_poles"] * 3 * int(np.prod(b["shape"])) for b in ade)
    ade_state = torch.zeros(max(ade_floats, 1), dtype=torch.float32, device=device) if ade else None
    det_states = {}
    for d in objects.detectors:
        st = {}
        for k, (shp, dt) in d.state_layout(config).items():
            tdt = torch.complex64 if np.dtype(dt) == np.complex64 else torch.float32
            st[k] = torch.zeros(shp, dtype=tdt, device=device)
        det_states[d.name] = st
    ie = torch.as_tensor(inv_eps, device=device)
    im = inv_mu if np.isscalar(inv_mu) else torch.as_tensor(inv_mu, device=device)
    fields = FieldState(E=E, H=H, psi_E=psi_E, psi_H=psi_H)
    sig_e = _conductivity_array(objects, config, "electric_conductivity")
    sig_h = _conductivity_array(objects, config, "magnetic_conductivity")
    return ArrayContainer(fields=fields, inv_permittivities=ie, inv_permeabilities=im, detector_states=det_states,
                          psi=psi, ade_state=ade_state, initial_inv_permittivities=ie.clone(),
                          E_parts=E_parts, H_parts=H_parts,
                          electric_conductivity=None if sig_e is None else torch.as_tensor(sig_e, device=device),
                          magnetic_conductivity=None if sig_h is None else torch.as_tensor(sig_h, device=device))


def _reduce_for_symmetry(object_list, slices, config):
    """fdtd/symmetry.py:84-195: clip every resolved slice onto the kept upper half of each symmetric
    axis (the plane at the centre of the volume's even extent becomes index 0), drop objects
    wholly in the discarded half, and keep each object's unclipped slice in the reduced frame.
    Periodic boundaries and mirror symmetry exclude each other on one axis."""
    sym = config.symmetry
    vol = next(o for o in object_list if isinstance(o, SimulationVolume))
    vs = slices[vol.name]
    mid, new_vol = [], []
    for a in range(3):
        v0, v1 = vs[a]
        n = v1 - v0
        if sym[a]:
            if n % 2:
                raise ValueError(f"the simulation volume needs an even number of cells along the symmetric "
                                 f"{'xyz'[a]} axis, got {n}")
            mid.append(v0 + n // 2)
            new_vol.append((0, v1 - mid[a]))
        else:
            mid.append(v0)
            new_vol.append((v0, v1))
    kept, new_slices, unreduced = [], {}, {}
    for o in object_list:
        sl = slices[o.name]
        if o is vol:
            new_slices[o.name] = tuple(new_vol)
            unreduced[o.name] = tuple((sl[a][0] - mid[a], sl[a][1] - mid[a]) if sym[a] else sl[a] for a in range(3))
            kept.append(o)
            continue
        clipped, unclipped, drop = [], [], False
        for a in range(3):
            s0, s1 = sl[a]
            if sym[a]:
                n0, n1 = max(s0, mid[a]) - mid[a], min(s1, vs[a][1]) - mid[a]
                drop = drop or n1 <= n0
                clipped.append((n0, n1))
                unclipped.append((s0 - mid[a], s1 - mid[a]))
            else:
                clipped.append((s0, s1))
                unclipped.append((s0, s1))
        if drop:
            continue
        if isinstance(o, BlochBoundary) and sym[o.axis]:
            raise NotImplementedError(f"periodic boundary '{o.name}' on the {'xyz'[o.axis]}-symmetry axis: periodic "
                                      "boundaries and mirror symmetry exclude each other on one axis")
        new_slices[o.name] = tuple(clipped)
        unreduced[o.name] = tuple(unclipped)
        kept.append(o)
    return kept, new_slices, unreduced, tuple(b - a for a, b in new_vol)


def _symmetry_walls(config, vol_shape, names, key):
    """fdtd/symmetry.py:198-264: a PEC face on every electric (-1) symmetry plane, at the reduced
    min face; a magnetic plane gets none (the zero halo half a cell below is its mirror)."""
    walls, used = [], set(names)
    for a in range(3):
        if config.symmetry[a] != -1:
            continue
        name = f"_sym_wall_{'xyz'[a]}"
        while name in used:
            name += "_"
        used.add(name)
        pgs = tuple(1 if b == a else None for b in range(3))
        w = PerfectElectricConductor(name=name, axis=a, direction="-", partial_grid_shape=pgs)
        gs = [(0, vol_shape[b]) for b in range(3)]
        gs[a] = (0, 1)
        w = w.place_on_grid(tuple(gs), config, key)
        w._is_symmetry_wall = True
        walls.append(w)
    return walls


def place_objects(object_list: Sequence[SimulationObject], config: SimulationConfig, constraints, key=None):
    """Resolve constraints, place objects, rasterise materials and allocate device arrays.

    Returns (ObjectContainer, ArrayContainer, ParameterContainer, SimulationConfig, info)."""
    slices = resolve_object_constraints(object_list, constraints, config)
    unreduced = {}
    if config.has_symmetry:
        object_list, slices, unreduced, vol_shape = _reduce_for_symmetry(object_list, slices, config)
    placed = [o.place_on_grid(slices[o.name], config, key) for o in object_list]
    for o in placed:
        if o.name in unreduced:
            o._unreduced_grid_slice_tuple = unreduced[o.name]
    if config.has_symmetry:
        placed = placed + _symmetry_walls(config, vol_shape, [o.name for o in placed], key)
    objects = ObjectContainer(placed)
    if config.gradient_config is not None and config.gradient_config.method == "reversible":
        if any(o.material.num_poles for o in objects.static_material_objects):
            raise NotImplementedError(
                "Dispersive time-reversible gradient computation under active development. "
                "Use GradientConfig(method='checkpointed') instead.")
    inv_eps, inv_mu, ec, mc = _material_arrays(objects, config)
    params = {d.name: d.init_params(key if isinstance(key, (int, np.random.Generator)) else None)
              for d in objects.devices}
    new_list = []
    ade_boxes, _ = _ade_boxes(objects, config)
    from .detectors import BaseModeOverlapDetector

    conducting = [m for m in [objects.volume, *objects.static_material_objects]
                  if any(m.material.electric_conductivity)]
    for o in objects.object_list:
        if isinstance(o, (Source, BaseModeOverlapDetector)):
            o = o.apply(key, inv_eps, inv_mu, dispersive=_dispersive_slice(ade_boxes, o.grid_slice_tuple),
                        lossy_plane=any(m.check_overlap(o) for m in conducting))
        new_list.append(o)
    objects = ObjectContainer(new_list, objects.volume_idx)
    arrays = _make_arrays(objects, config, inv_eps, inv_mu, _device())
    return objects, arrays, params, config, {}


def apply_params(arrays: ArrayContainer, objects: ObjectContainer, params: dict, key=None, **kwargs):
    """initialization.py:317-521, continuous devices: eps = eps_a + rho (eps_b - eps_a) written as
    1/eps into the device slice (differentiable in rho through torch); sources overlapping a device
    are re-applied with the new (detached) materials."""
    import torch

    ie = arrays.initial_inv_permittivities
    if ie is None:
        ie = arrays.inv_permittivities
    out = ie.detach().clone()
    for d in objects.devices:
        (_, ma), (_, mb) = d.ordered_materials
        rho = params[d.name]
        if not isinstance(rho, torch.Tensor):
            rho = torch.as_tensor(np.asarray(rho, np.float32), device=out.device)
        ea = torch.tensor(np.float32(ma.permittivity[0]), device=out.device)
        eb = torch.tensor(np.float32(mb.permittivity[0]), device=out.device)
        perm = ea + rho * (eb - ea)
        idx = (slice(None), *d.grid_slice)
        if out.shape[0] == 9:  # an isotropic design in a tensor array: the diagonal, zero off-diagonals
            val = torch.zeros((9, *perm.shape), dtype=out.dtype, device=out.device)
            val = val.index_copy(0, torch.tensor([0, 4, 8], device=out.device),
                                 (1.0 / perm)[None].expand(3, *perm.shape))
            out[idx] = val
        else:
            out[idx] = (1.0 / perm)[None].expand(out.shape[0], *perm.shape)
    new = copy.copy(arrays)
    new.inv_permittivities = out
    new_list = []
    devs = objects.devices
    host_ie = None
    for o in objects.object_list:
        if isinstance(o, Source) and any(d.check_overlap(o) for d in devs):
            if host_ie is None:
                host_ie = out.detach().cpu().numpy()
            im = arrays.inv_permeabilities
            boxes, _ = _ade_boxes(objects, o._config)
            o = o.apply(key, host_ie, im if np.isscalar(im) else im.detach().cpu().numpy(),
                        dispersive=_dispersive_slice(boxes, o.grid_slice_tuple))
        new_list.append(o)
    return new, ObjectContainer(new_list, objects.volume_idx), {}


# ---------------------------------------------------------------- running the engine
class Solve:
    """One solve's device buffers bound to a plan (workspace, record, ADE slot bookkeeping)."""

    def __init__(self, objects: ObjectContainer, config: SimulationConfig, arrays: ArrayContainer, record_steps=0):
        import torch

        spec = lower(objects, config, arrays)
        self.plan = engine.Plan(spec)
        self.spec = spec
        self.objects, self.config, self.arrays = objects, config, arrays
        dev = arrays.fields.E.device
        if dev.type != "cuda":
            raise engine.EngineError("the HIP engine needs the arrays on a ROCm device; no CPU fallback exists")
        self.workspace = torch.empty(max(self.plan.workspace_bytes, 256), dtype=torch.uint8, device=dev)
        self.record = None
        self.det_keys = [(d["name"], d["key"]) for d in spec["detectors"]]  # per engine detector
        self.state_keys = list(dict.fromkeys(self.det_keys))  # the state entries they write
        self.ensure_record(record_steps)

    def record_bytes(self, steps: int) -> int:
        """Device bytes of the interface record of steps [0, steps): the slots of the kept steps
        up to the one the last step reads (its own, or the next kept one)."""
        p = self.plan
        if not steps or not p.record_floats_per_step:
            return 0
        last = p.slot_after(steps - 1)
        return (last + 1) * p.record_slot_bytes

    def ensure_record(self, steps: int):
        """Hold an interface record of `steps` steps (none for 0), as raw bytes in the plan's slot
        layout."""
        import torch

        need = self.record_bytes(steps)
        if not need:
            self.record = None
            return
        if self.record is None or self.record.numel() < need:  # zeroed: the alignment padding stays 0
            self.record = torch.zeros(need, dtype=torch.uint8, device=self.arrays.fields.E.device)

    def det_tensors(self):
        """The state buffer of each engine detector, in plan order (a group accumulating into one
        entry repeats its buffer)."""
        st = self.arrays.detector_states
        return [st[n][k] for n, k in self.det_keys]

    def state_tensors(self):
        """The distinct state entries the engine writes, in plan order."""
        st = self.arrays.detector_states
        return [st[n][k] for n, k in self.state_keys]

    def buffers(self, inv_eps=None):
        import torch

        a = self.arrays
        ie = (a.inv_permittivities if inv_eps is None else inv_eps).detach()
        ie = expand_inverse_material(ie, self.spec["eps_comps"]).contiguous()
        im = a.inv_permeabilities
        if self.spec["mu_comps"] == 9 and np.isscalar(im):
            im = torch.full((1, *ie.shape[1:]), float(im), dtype=torch.float32, device=ie.device)
        im = None if np.isscalar(im) else expand_inverse_material(im.detach(), self.spec["mu_comps"]).contiguous()
        E, H = (a.E_parts, a.H_parts) if a.complex_fields else (a.fields.E, a.fields.H)
        b, keep = engine.make_buffers(E, H, ie, im, a.psi, a.ade_state, self.det_tensors(), self.record,
                                      self.workspace)
        self._keep = (keep, ie, im)
        return b

    def forward(self, t0, t1, record=False, inv_eps=None):
        b = self.buffers(inv_eps)
        engine.forward(self.plan, b, t0, t1, record, engine.current_stream_handle())
        self.arrays.sync_complex_fields()


def _solve_for(objects, config, arrays, record_steps=0):
    cache = objects._plan_cache
    key = (id(arrays), tuple(arrays.inv_permittivities.shape))
    s = cache.get(key)
    if s is None or s.arrays is not arrays:
        s = Solve(objects, config, arrays, record_steps)
        cache.clear()
        cache[key] = s
    elif record_steps:
        s.ensure_record(record_steps)
    return s


def custom_fdtd_forward(arrays, objects, config, key=None, reset_container=True, record_detectors=True,
                        start_time=0, end_time=None, show_progress=False, progress_callback=None):
    """fdtd.py:499-584: advance steps [start_time, end_time) (optionally after a reset)."""
    if end_time is None:
        end_time = config.time_steps_total
    if not record_detectors:
        raise NotImplementedError("record_detectors=False is not supported by this engine")
    if reset_container:
        arrays.reset()
    s = _solve_for(objects, config, arrays)
    chunk = max(1, (int(end_time) - int(start_time)) // 20) if progress_callback else None
    t = int(start_time)
    while t < int(end_time):
        t1 = int(end_time) if chunk is None else min(int(end_time), t + chunk)
        s.forward(t, t1)
        t = t1
        if progress_callback:
            progress_callback(t, int(end_time))
    return int(end_time), arrays


def full_backward(state, objects, config, key=None, record_detectors=True, reset_fields=True, start_time_step=0):
    """backward.py:18-59: the time-reversed run from the state's step back to start_time_step
    (fdtdx_reverse): each step restores the recorded interface layers, undoes the H and E
    half-steps with the CPML psi frozen, zeroes the CPML regions when reset_fields, and records
    the inverse detectors when record_detectors. Needs the interface record of the forward run
    (reversible_fdtd keeps it whenever it fits the record budget)."""
    from .reversible import _stream_handle

    t, arrays = state
    t, start = int(t), int(start_time_step)
    if start >= t:
        return start, arrays
    solve = _solve_for(objects, config, arrays)
    if solve.plan.record_floats_per_step and (solve.record is None or solve.record.numel() < solve.record_bytes(t)):
        raise Exception("Need recorder to add interfaces: run the forward with reversible_fdtd (whose interface "
                        "record full_backward replays) before the time-reversed run")
    engine.reverse(solve.plan, solve.buffers(), t, start, reset_fields, record_detectors, _stream_handle())
    return start, arrays


def backward(state, config, objects, key=None, record_detectors=True, reset_fields=True):
    """backward.py:62-135: one time-reversed step."""
    t, _ = state
    return full_backward(state, objects, config, key, record_detectors, reset_fields, int(t) - 1)


def checkpointed_fdtd(arrays, objects, config, key=None, stopping_condition=None, show_progress=False,
                      progress_callback=None):
    """fdtd.py:421-496: reset, then T forward steps (the forward-only path of run_fdtd)."""
    if stopping_condition is not None:
        raise NotImplementedError("custom stopping conditions are not supported by this engine")
    if config.gradient_config is not None and config.gradient_config.method == "checkpointed":
        if getattr(arrays.inv_permittivities, "requires_grad", False):
            raise NotImplementedError("checkpointed (autodiff) gradients are not supported; use method='reversible'")
    return custom_fdtd_forward(arrays, objects, config, key, True, True, 0, config.time_steps_total,
                               show_progress, progress_callback)


def run_fdtd(arrays, objects, config, key=None, stopping_condition=None, show_progress=False, progress_callback=None):
    """wrapper.py:14-63."""
    if stopping_condition is not None and config.gradient_config is not None:
        raise NotImplementedError(
            "Custom stopping conditions are not yet compatible with gradient computation. "
            "Set config.gradient_config to None or use default time-based stopping by setting stopping_condition=None.")
    if config.gradient_config is None:
        return checkpointed_fdtd(arrays, objects, config, key, stopping_condition, show_progress, progress_callback)
    if config.gradient_config.method == "reversible":
        return reversible_fdtd(arrays, objects, config, key, show_progress, progress_callback)
    if config.gradient_config.method == "checkpointed":
        return checkpointed_fdtd(arrays, objects, config, key, stopping_condition, show_progress, progress_callback)
    raise Exception(f"Unknown gradient computation method: {config.gradient_config.method}")


def reversible_fdtd(arrays, objects, config, key=None, show_progress=False, progress_callback=None):
    """fdtd.py:39-418: forward with interface recording; when inv_permittivities requires grad, the
    detector states returned are differentiable and their backward runs the HIP time-reversed
    adjoint (fdtdx_reverse_adjoint)."""
    from .reversible import reversible_run

    if arrays.complex_fields:
        raise NotImplementedError("reversible gradients of complex-field (Bloch) simulations are not supported by "
                                  "this engine")
    if any(o.material.num_poles for o in objects.static_material_objects) or arrays.ade_state is not None:
        raise NotImplementedError(
            "Dispersive time-reversible gradient computation under active development. "
            "Use GradientConfig(method='checkpointed') instead.")
    return reversible_run(arrays, objects, config, progress_callback)
