"""Simulation objects and placement constraints (uniform-grid subset of the reference's
src/fdtdx/objects/object.py:62-919 and the constraint resolver of fdtd/initialization.py:1229-2143).

Objects are plain dataclasses. Placement returns placed copies that carry ``grid_slice_tuple``.
The constraint helpers keep the reference's names and argument meaning; positions are anchors in
[-1, 1] (lower side, centre, upper side) and margins are metres (``grid_margins`` in cells).
"""

from __future__ import annotations

import copy
import itertools
from dataclasses import dataclass, field
from typing import Literal, Sequence

import numpy as np

from . import constants
from .materials import Material

_COUNTER = itertools.count()


# ---------------------------------------------------------------- constraints (object.py:62-200)
@dataclass(frozen=True)
class PositionConstraint:
    object: str
    other_object: str
    axes: tuple
    object_positions: tuple
    other_object_positions: tuple
    margins: tuple
    grid_margins: tuple


@dataclass(frozen=True)
class SizeConstraint:
    object: str
    other_object: str
    axes: tuple
    other_axes: tuple
    proportions: tuple
    offsets: tuple
    grid_offsets: tuple


@dataclass(frozen=True)
class SizeExtensionConstraint:
    object: str
    other_object: str | None
    axis: int
    direction: str
    other_position: float
    offset: float
    grid_offset: int


@dataclass(frozen=True)
class GridCoordinateConstraint:
    object: str
    axes: tuple
    sides: tuple
    coordinates: tuple


def _tup(v, n):
    if v is None:
        return tuple([0] * n)
    if np.isscalar(v):
        return (v,)
    return tuple(v)


@dataclass(kw_only=True)
class SimulationObject:
    partial_real_shape: tuple = (None, None, None)
    partial_grid_shape: tuple = (None, None, None)
    partial_real_position: tuple = (None, None, None)
    name: str | None = None
    color: tuple | None = None

    def __post_init__(self):
        if self.name is None:
            self.name = f"Object_{next(_COUNTER)}"
        self._grid_slice_tuple = None
        self._unreduced_grid_slice_tuple = None
        self._config = None

    # ------------------------------------------------------------ placement state
    @property
    def grid_slice_tuple(self):
        if self._grid_slice_tuple is None:
            raise Exception(f"Object is not yet initialized: {self.name}")
        return self._grid_slice_tuple

    @property
    def grid_slice(self):
        return tuple(slice(a, b) for a, b in self.grid_slice_tuple)

    # ------------------------------------------------------------ mirror symmetry (object.py:261-332)
    @property
    def unreduced_grid_slice_tuple(self):
        """The full-domain slice in the reduced grid's frame: a negative start where a symmetry
        plane clipped the object."""
        return self._unreduced_grid_slice_tuple or self.grid_slice_tuple

    def straddles_symmetry_plane(self, axis: int) -> bool:
        """The object extended past the symmetry plane of `axis` into the discarded half."""
        return self.unreduced_grid_slice_tuple[axis][0] < 0

    @property
    def straddled_symmetry_axes(self) -> tuple:
        return tuple(a for a in range(3) if self.straddles_symmetry_plane(a))

    def touches_symmetry_plane(self, axis: int) -> bool:
        """The object starts on the symmetry plane of `axis` (clipped there or not)."""
        return (self._config is not None and self._config.symmetry[axis] != 0
                and self.grid_slice_tuple[axis][0] == 0)

    @property
    def grid_shape(self):
        return tuple(b - a for a, b in self.grid_slice_tuple)

    @property
    def real_shape(self):
        dx = self._config.uniform_spacing()
        return tuple(s * dx for s in self.grid_shape)

    def place_on_grid(self, grid_slice_tuple, config, key=None):
        obj = copy.copy(self)
        obj._grid_slice_tuple = tuple((int(a), int(b)) for a, b in grid_slice_tuple)
        obj._config = config
        return obj

    def apply(self, key, inv_permittivities, inv_permeabilities, **kwargs):
        return self

    def check_overlap(self, other: "SimulationObject") -> bool:
        for (a0, a1), (b0, b1) in zip(self.grid_slice_tuple, other.grid_slice_tuple):
            if a1 <= b0 or b1 <= a0:
                return False
        return True

    # ------------------------------------------------------------ constraint helpers (object.py:431-890)
    def place_relative_to(self, other, axes, own_positions, other_positions, margins=None, grid_margins=None):
        axes = _tup(axes, 1)
        n = len(axes)
        own_positions, other_positions = _tup(own_positions, n), _tup(other_positions, n)
        margins = _tup(margins, n) if margins is not None else tuple([0.0] * n)
        grid_margins = _tup(grid_margins, n) if grid_margins is not None else tuple([0] * n)
        if not (len(own_positions) == len(other_positions) == len(margins) == len(grid_margins) == n):
            raise Exception("All inputs should have same lengths")
        return PositionConstraint(object=self.name, other_object=other.name, axes=axes,
                                  object_positions=own_positions, other_object_positions=other_positions,
                                  margins=margins, grid_margins=grid_margins)

    def size_relative_to(self, other, axes, other_axes=None, proportions=None, offsets=None, grid_offsets=None):
        axes = _tup(axes, 1)
        n = len(axes)
        other_axes = axes if other_axes is None else _tup(other_axes, n)
        proportions = tuple([1.0] * n) if proportions is None else _tup(proportions, n)
        offsets = tuple([0.0] * n) if offsets is None else _tup(offsets, n)
        grid_offsets = tuple([0] * n) if grid_offsets is None else _tup(grid_offsets, n)
        return SizeConstraint(object=self.name, other_object=other.name, axes=axes, other_axes=other_axes,
                              proportions=proportions, offsets=offsets, grid_offsets=grid_offsets)

    def same_size(self, other, axes=(0, 1, 2), offsets=None, grid_offsets=None):
        return self.size_relative_to(other, axes=axes, offsets=offsets, grid_offsets=grid_offsets)

    def place_at_center(self, other, axes=(0, 1, 2), own_positions=None, other_positions=None, margins=None,
                        grid_margins=None):
        axes = _tup(axes, 1)
        n = len(axes)
        own_positions = tuple([0.0] * n) if own_positions is None else own_positions
        other_positions = tuple([0.0] * n) if other_positions is None else other_positions
        return self.place_relative_to(other, axes, own_positions, other_positions, margins, grid_margins)

    def same_position(self, other, axes=(0, 1, 2)):
        return self.place_at_center(other, axes=axes)

    def same_position_and_size(self, other, axes=(0, 1, 2)):
        return [self.same_position(other, axes), self.same_size(other, axes)]

    def face_to_face_positive_direction(self, other, axes, margins=None, grid_margins=None):
        axes = _tup(axes, 1)
        n = len(axes)
        return self.place_relative_to(other, axes, tuple([-1] * n), tuple([1] * n), margins, grid_margins)

    def face_to_face_negative_direction(self, other, axes, margins=None, grid_margins=None):
        axes = _tup(axes, 1)
        n = len(axes)
        return self.place_relative_to(other, axes, tuple([1] * n), tuple([-1] * n), margins, grid_margins)

    def place_above(self, other, margins=None, grid_margins=None):
        return self.face_to_face_positive_direction(other, (2,), margins, grid_margins)

    def place_below(self, other, margins=None, grid_margins=None):
        return self.face_to_face_negative_direction(other, (2,), margins, grid_margins)

    def set_grid_coordinates(self, axes, sides, coordinates):
        axes = _tup(axes, 1)
        return GridCoordinateConstraint(object=self.name, axes=axes, sides=_tup(sides, len(axes)),
                                        coordinates=_tup(coordinates, len(axes)))

    def extend_to(self, other, axis, direction, other_position=None, offset=0.0, grid_offset=0):
        if other_position is None:
            other_position = -1.0 if direction == "+" else 1.0
        return SizeExtensionConstraint(object=self.name, other_object=None if other is None else other.name,
                                       axis=axis, direction=direction, other_position=other_position,
                                       offset=offset, grid_offset=grid_offset)


# ---------------------------------------------------------------- static material objects
@dataclass(kw_only=True)
class SimulationVolume(SimulationObject):
    """static_material/static.py:123-139: the simulation domain and its background material."""

    material: Material = field(default_factory=lambda: Material(permittivity=constants.relative_permittivity_air))
    placement_order: int = -1000


@dataclass(kw_only=True)
class UniformMaterialObject(SimulationObject):
    material: Material = field(default_factory=Material)
    placement_order: int = 0


# ---------------------------------------------------------------- design region
@dataclass(kw_only=True)
class Device(SimulationObject):
    """objects/device/device.py: a design region whose latent parameters rho in [0, 1] interpolate
    between two materials (continuous output, fdtd/initialization.py:404-417).

    Only ``param_transforms=[]`` with two materials is supported; the materials are ordered by
    permittivity like the reference (the lower one is index 0)."""

    materials: dict = field(default_factory=dict)
    param_transforms: list = field(default_factory=list)
    partial_voxel_grid_shape: tuple = (1, 1, 1)
    partial_voxel_real_shape: tuple = (None, None, None)

    def __post_init__(self):
        super().__post_init__()
        if self.param_transforms:
            raise NotImplementedError("latent parameter transforms are outside this engine's scope")
        if len(self.materials) != 2:
            raise NotImplementedError("continuous devices with exactly two materials are supported")

    @property
    def ordered_materials(self):
        return sorted(self.materials.items(), key=lambda kv: kv[1].permittivity[0])

    def init_params(self, key=None):
        """Uniform latent parameters in [0, 1] (device.py:317-340); ``key`` is a numpy Generator or
        seed. The reference draws with jax.random; seeded numpy draws are used here."""
        rng = key if isinstance(key, np.random.Generator) else np.random.default_rng(0 if key is None else key)
        return rng.uniform(0.0, 1.0, size=self.grid_shape).astype(np.float32)


# ---------------------------------------------------------------- boundaries
@dataclass(kw_only=True)
class PerfectlyMatchedLayer(SimulationObject):
    """objects/boundaries/perfectly_matched_layer.py:12-190 (uniform grid)."""

    axis: int = 0
    direction: str = "-"
    alpha_start: float | None = None
    alpha_end: float | None = None
    alpha_order: float | None = None
    kappa_start: float | None = None
    kappa_end: float | None = None
    kappa_order: float | None = None
    sigma_start: float | None = None
    sigma_end: float | None = None
    sigma_order: float | None = None

    def __post_init__(self):
        super().__post_init__()
        if self.alpha_start is None:
            self.alpha_start = 0.01 * 2 * np.pi * constants.c / 1.55e-6 * constants.eps0
        if self.alpha_end is None:
            self.alpha_end = 0.0
        if self.alpha_order is None:
            self.alpha_order = 1.0
        if self.kappa_start is None:
            self.kappa_start = 1.0
        if self.kappa_end is None:
            self.kappa_end = 1.0
        if self.kappa_order is None:
            self.kappa_order = 3.0
        if self.sigma_start is None:
            self.sigma_start = 0.0
        if self.sigma_order is None:
            self.sigma_order = 3.0

    @property
    def thickness(self) -> int:
        return self.grid_shape[self.axis]

    @property
    def descriptive_name(self) -> str:
        return f"{'min' if self.direction == '-' else 'max'}_{'xyz'[self.axis]}"

    def interface_slice_tuple(self):
        s = list(self.grid_slice_tuple)
        lo, hi = s[self.axis]
        s[self.axis] = (lo, lo + 1) if self.direction == "+" else (hi - 1, hi)
        return tuple(s)

    def place_on_grid(self, grid_slice_tuple, config, key=None):
        obj = super().place_on_grid(grid_slice_tuple, config, key)
        if obj.sigma_end is None:
            # PML.py:109 evaluated in float32 (jnp.log of a python float is weak float32)
            thickness_m = obj.thickness * config.uniform_spacing()
            val = (np.float32(-(obj.sigma_order + 1)) * np.float32(np.log(np.float32(1e-6)))
                   / np.float32(2 * (constants.eta0 / 1.0) * thickness_m))
            obj.sigma_end = float(np.float32(val))
        obj._coefficients = obj._compute_coefficients(config)
        return obj

    def _compute_profile(self, v0, v1, order):
        """PML.py:231-275, float32: value_start + (value_end - value_start) * (d / L) ** order."""
        f = np.float32
        L = self.thickness
        if self.direction == "-":
            dE = np.arange(L - 1, -1, -1, dtype=f)
            dH = np.append(np.arange(L - 1.5, -0.5, -1, dtype=f), f(0)).astype(f)
        else:
            dE = np.insert(np.arange(0.5, L - 0.5, 1, dtype=f), 0, f(0)).astype(f)
            dH = np.arange(0, L, 1, dtype=f)
        norm = f(L)
        pe = (f(v0) + f(v1 - v0) * np.power(dE / norm, f(order))).astype(f)
        ph = (f(v0) + f(v1 - v0) * np.power(dH / norm, f(order))).astype(f)
        return pe, ph

    def _compute_coefficients(self, config):
        """PML.py:97-136: b = expm1(-dt/eps0 (sigma/kappa + alpha)) + 1,
        a = nan_to_num((b - 1) sigma / (sigma + alpha kappa) / kappa)."""
        f = np.float32
        sE, sH = self._compute_profile(self.sigma_start, self.sigma_end, self.sigma_order)
        kE, kH = self._compute_profile(self.kappa_start, self.kappa_end, self.kappa_order)
        aE, aH = self._compute_profile(self.alpha_start, self.alpha_end, self.alpha_order)
        s = f(-config.time_step_duration / constants.eps0)
        out = {}
        for tag, sig, kap, alp in (("E", sE, kE, aE), ("H", sH, kH, aH)):
            b = (np.expm1(s * (sig / kap + alp)) + f(1)).astype(f)
            with np.errstate(invalid="ignore", divide="ignore"):
                a = (b - f(1.0)) * sig / (sig + alp * kap) / kap
            out["b_" + tag] = b
            out["a_" + tag] = np.nan_to_num(a, nan=0.0).astype(f)
            out["inv_kappa_" + tag] = (f(1.0) / kap).astype(f)
        return out

    @property
    def kappa_is_one(self) -> bool:
        return self.kappa_start == 1.0 and self.kappa_end == 1.0


@dataclass(kw_only=True)
class _ConductorFace(SimulationObject):
    """A one-cell boundary layer whose post-update zeroing the engine applies."""

    axis: int = 0
    direction: str = "-"

    @property
    def thickness(self) -> int:
        return 1

    @property
    def descriptive_name(self) -> str:
        return f"{'min' if self.direction == '-' else 'max'}_{'xyz'[self.axis]}"

    @property
    def tangential_components(self) -> tuple[int, int]:
        return tuple(c for c in range(3) if c != self.axis)

    @property
    def layer_index(self) -> int:
        lo, hi = self.grid_slice_tuple[self.axis]
        return lo


@dataclass(kw_only=True)
class PerfectElectricConductor(_ConductorFace):
    """objects/boundaries/pec.py: tangential E is set to zero on the face layer after every E
    update (pec.py:70-77)."""


@dataclass(kw_only=True)
class PerfectMagneticConductor(_ConductorFace):
    """objects/boundaries/pmc.py: tangential H is set to zero on the face layer after every H
    update (pmc.py:62-75)."""


@dataclass(kw_only=True)
class BlochBoundary(SimulationObject):
    """objects/boundaries/bloch.py: a one-cell face layer that makes its axis periodic. The curls
    read the far face's field as the near face's neighbour (wrap padding, update.py:92-137), times
    the Bloch phase exp(i k L) of this boundary (its conjugate across the min face,
    apply_pad_correction, bloch.py:61-97). A nonzero component of ``bloch_vector`` along the axis
    needs complex fields (bloch.py:32-38), which the engine then keeps as a real and an imaginary
    part. The backward-pass field reset of a one-cell layer copies the layer onto itself
    (bloch.py:99-133), so it is a no-op."""

    axis: int = 0
    direction: str = "-"
    bloch_vector: tuple = (0.0, 0.0, 0.0)

    @property
    def needs_complex_fields(self) -> bool:
        """bloch.py:32-38: only a nonzero Bloch vector component along this boundary's axis."""
        return self.bloch_vector[self.axis] != 0.0

    def get_bloch_phase(self, volume_shape, resolution: float) -> float:
        """The phase angle k_axis * L, L = volume_shape[axis] * resolution (get_bloch_phase,
        bloch.py:140-153, uniform grid); the reference returns exp(i k L)."""
        return float(self.bloch_vector[self.axis]) * float(volume_shape[self.axis]) * float(resolution)

    @property
    def thickness(self) -> int:
        return 1

    @property
    def uses_wrap_padding(self) -> bool:
        return True

    @property
    def descriptive_name(self) -> str:
        return f"{'min' if self.direction == '-' else 'max'}_{'xyz'[self.axis]}"


_FACES = ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z")


@dataclass
class BoundaryConfig:
    """objects/boundaries/initialization.py:17-551. Per-face types "pml", "pec", "pmc",
    "periodic" or "bloch" (the last with the phase of ``bloch_vector``'s component along the face's axis)."""

    thickness_grid_minx: int = 10
    thickness_grid_maxx: int = 10
    thickness_grid_miny: int = 10
    thickness_grid_maxy: int = 10
    thickness_grid_minz: int = 10
    thickness_grid_maxz: int = 10
    boundary_type_minx: str = "pml"
    boundary_type_maxx: str = "pml"
    boundary_type_miny: str = "pml"
    boundary_type_maxy: str = "pml"
    boundary_type_minz: str = "pml"
    boundary_type_maxz: str = "pml"
    pml_params: dict = field(default_factory=dict)
    bloch_vector: tuple = (0.0, 0.0, 0.0)

    def __post_init__(self):
        for face, t in self.get_type_dict().items():
            if t not in ("pml", "pec", "pmc", "periodic", "bloch"):
                raise Exception(f"Unknown boundary type {t!r} for {face}")

    @classmethod
    def from_uniform_bound(cls, thickness: int = 10, boundary_type: str = "pml", override_types=None, **pml_params):
        """initialization.py:444-551."""
        override_types = override_types or {}
        types = {f: override_types.get(f, boundary_type) for f in _FACES}
        params = {k: v for k, v in pml_params.items() if v is not None and k != "bloch_vector"}
        bloch = tuple(pml_params.get("bloch_vector") or (0.0, 0.0, 0.0))
        return cls(thickness, thickness, thickness, thickness, thickness, thickness,
                   types["min_x"], types["max_x"], types["min_y"], types["max_y"], types["min_z"], types["max_z"],
                   params, bloch)

    def get_dict(self) -> dict:
        return {"min_x": self.thickness_grid_minx, "max_x": self.thickness_grid_maxx,
                "min_y": self.thickness_grid_miny, "max_y": self.thickness_grid_maxy,
                "min_z": self.thickness_grid_minz, "max_z": self.thickness_grid_maxz}

    def get_type_dict(self) -> dict:
        return {"min_x": self.boundary_type_minx, "max_x": self.boundary_type_maxx,
                "min_y": self.boundary_type_miny, "max_y": self.boundary_type_maxy,
                "min_z": self.boundary_type_minz, "max_z": self.boundary_type_maxz}


def boundary_objects_from_config(config: BoundaryConfig, volume: SimulationVolume):
    """initialization.py:554-663: one boundary object per face, in order min_x, max_x, min_y, max_y,
    min_z, max_z, each placed flush with its volume face and spanning the volume on the other axes:
    a PML of the face's thickness, or a one-cell PEC / PMC / periodic layer."""
    boundaries, constraints = {}, []
    types = config.get_type_dict()
    for kind, thickness in config.get_dict().items():
        side, ax = kind.split("_")
        axis = "xyz".index(ax)
        direction = "-" if side == "min" else "+"
        shape = [None, None, None]
        btype = types[kind]
        shape[axis] = thickness if btype == "pml" else 1
        if btype == "pml":
            pml = PerfectlyMatchedLayer(axis=axis, direction=direction, partial_grid_shape=tuple(shape), name=kind,
                                        **config.pml_params)
        elif btype in ("periodic", "bloch"):
            pml = BlochBoundary(axis=axis, direction=direction, partial_grid_shape=tuple(shape), name=kind,
                                bloch_vector=(0.0, 0.0, 0.0) if btype == "periodic" else tuple(config.bloch_vector))
        elif btype == "pec":
            pml = PerfectElectricConductor(axis=axis, direction=direction, partial_grid_shape=tuple(shape), name=kind)
        else:
            pml = PerfectMagneticConductor(axis=axis, direction=direction, partial_grid_shape=tuple(shape), name=kind)
        pos = -1.0 if direction == "-" else 1.0
        constraints.append(pml.place_relative_to(volume, axes=(axis,), own_positions=(pos,), other_positions=(pos,)))
        other = tuple(a for a in range(3) if a != axis)
        constraints.append(pml.size_relative_to(volume, axes=other))
        constraints.append(pml.place_at_center(volume, axes=other))
        boundaries[kind] = pml
    return boundaries, constraints


# ---------------------------------------------------------------- constraint resolution
def resolve_object_constraints(objects: Sequence[SimulationObject], constraints, config, max_iter: int = 1000):
    """Uniform-grid, index-space version of initialization.py:1229-1300, 1872-2060."""
    names = [o.name for o in objects]
    dup = {n for n in names if names.count(n) > 1}
    if dup:
        raise Exception(f"Duplicate object names detected: {', '.join(sorted(dup))}.")
    omap = {o.name: o for o in objects}
    for c in constraints:
        for nm in (c.object, getattr(c, "other_object", None)):
            if nm is not None and nm not in omap:
                raise Exception(f"Unknown object name in constraint: {nm}")
    vols = [o for o in objects if isinstance(o, SimulationVolume)]
    if len(vols) != 1:
        raise Exception("Need exactly one SimulationVolume")
    vol = vols[0]
    dx = config.uniform_spacing()

    def shape_of(o):
        out = []
        for a in range(3):
            g, r = o.partial_grid_shape[a], o.partial_real_shape[a]
            if g is not None and r is not None:
                raise Exception(f"{o.name}: both grid and real shape given on axis {a}")
            out.append(int(g) if g is not None else (int(round(r / dx)) if r is not None else None))
        return out

    shape = {o.name: shape_of(o) for o in objects}
    sl = {o.name: [[None, None] for _ in range(3)] for o in objects}
    if any(s is None for s in shape[vol.name]):
        raise Exception("SimulationVolume needs a fully specified shape")
    for a in range(3):
        sl[vol.name][a] = [0, shape[vol.name][a]]
    for o in objects:
        for a in range(3):
            rp = o.partial_real_position[a]
            if rp is not None and shape[o.name][a] is not None:
                gc = round(rp / dx + shape[vol.name][a] / 2)
                lo = round(gc - shape[o.name][a] / 2)
                sl[o.name][a] = [lo, lo + shape[o.name][a]]

    def settle():
        ch = False
        for nm in sl:
            for a in range(3):
                lo, hi = sl[nm][a]
                s = shape[nm][a]
                if lo is not None and hi is not None and s is None:
                    shape[nm][a] = hi - lo; ch = True
                elif lo is not None and hi is None and s is not None:
                    sl[nm][a][1] = lo + s; ch = True
                elif lo is None and hi is not None and s is not None:
                    sl[nm][a][0] = hi - s; ch = True
        return ch

    def setv(nm, a, idx, v):
        old = sl[nm][a][idx]
        if old is None:
            sl[nm][a][idx] = int(v)
            return True
        if old != v:
            raise Exception(f"Inconsistent placement of {nm} on axis {a}: {old} != {v}")
        return False

    for _ in range(max_iter):
        changed = settle()
        for c in constraints:
            if isinstance(c, GridCoordinateConstraint):
                for a, side, coord in zip(c.axes, c.sides, c.coordinates):
                    changed |= setv(c.object, a, 0 if side == "-" else 1, coord)
            elif isinstance(c, PositionConstraint):
                for idx, a in enumerate(c.axes):
                    ob0, ob1 = sl[c.other_object][a]
                    s = shape[c.object][a]
                    if ob0 is None or ob1 is None or s is None:
                        continue
                    anchor = ob0 + 0.5 * (c.other_object_positions[idx] + 1.0) * (ob1 - ob0)
                    anchor += (c.margins[idx] or 0.0) / dx + (c.grid_margins[idx] or 0)
                    cell = round(anchor)
                    lo = cell - round(0.5 * (c.object_positions[idx] + 1.0) * s)
                    changed |= setv(c.object, a, 0, lo)
                    changed |= setv(c.object, a, 1, lo + s)
            elif isinstance(c, SizeConstraint):
                for idx, a in enumerate(c.axes):
                    oa = c.other_axes[idx]
                    ob0, ob1 = sl[c.other_object][oa]
                    if ob0 is None or ob1 is None:
                        continue
                    length = (ob1 - ob0) * c.proportions[idx] + (c.offsets[idx] or 0.0) / dx + (c.grid_offsets[idx] or 0)
                    s = int(round(length))
                    if shape[c.object][a] is None:
                        shape[c.object][a] = s; changed = True
                    elif shape[c.object][a] != s:
                        raise Exception(f"Inconsistent size of {c.object} on axis {a}")
            elif isinstance(c, SizeExtensionConstraint):
                di = 0 if c.direction == "-" else 1
                if c.other_object is None:
                    target = sl[vol.name][c.axis][di]
                else:
                    ob0, ob1 = sl[c.other_object][c.axis]
                    if ob0 is None or ob1 is None:
                        continue
                    target = round(ob0 + 0.5 * (c.other_position + 1.0) * (ob1 - ob0) + (c.offset or 0.0) / dx
                                   + (c.grid_offset or 0))
                changed |= setv(c.object, c.axis, di, target)
        changed |= settle()
        if not changed:
            # an axis with neither size nor position extends over the whole volume
            for nm in sl:
                for a in range(3):
                    if sl[nm][a] == [None, None] and shape[nm][a] is None:
                        sl[nm][a] = [0, shape[vol.name][a]]
                        changed = True
            if not changed:
                break
    out = {}
    for nm in sl:
        st = []
        for a in range(3):
            lo, hi = sl[nm][a]
            if lo is None or hi is None:
                raise Exception(f"Object '{nm}' has unresolved bounds on axis {a}")
            if lo < 0 or hi > shape[vol.name][a] or hi <= lo:
                raise Exception(f"Object '{nm}' out of bounds on axis {a}: ({lo}, {hi})")
            st.append((lo, hi))
        out[nm] = tuple(st)
    return out
