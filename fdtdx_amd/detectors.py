"""Detectors (reference: src/fdtdx/objects/detectors/{detector,energy,field,poynting_flux,phasor}.py).

Each detector lowers to one engine descriptor; its state dict has the reference's keys, shapes
and dtypes (detector.py:231-244) and lives on the device.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from .core import OnOffSwitch, TemporalWindow
from .objects import SimulationObject

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
KIND_ENERGY, KIND_FIELD, KIND_FLUX, KIND_PHASOR = 0, 1, 2, 3


@dataclass(kw_only=True)
class Detector(SimulationObject):
    dtype: object = np.float32
    exact_interpolation: bool = True
    inverse: bool = False
    switch: OnOffSwitch = field(default_factory=OnOffSwitch)
    plot: bool = False

    def _on_list(self, config):
        return self.switch.calculate_on_list(config.time_steps_total, config.time_step_duration)

    def time_to_idx(self, config):
        on = self._on_list(config)
        idx = np.full(len(on), -1, np.int32)
        c = 0
        for t, o in enumerate(on):
            if o:
                idx[t] = c
                c += 1
        return idx

    @property
    def num_time_steps_recorded(self) -> int:
        return int(np.sum(self.time_to_idx(self._config) >= 0))

    def _check(self):
        pass

    def state_layout(self, config) -> dict:  # name -> (shape, dtype)
        raise NotImplementedError

    def lower(self, config) -> dict:
        raise NotImplementedError

    def _base_desc(self, config, kind):
        # an inverse detector is recorded on the time-reversed run (backward.py:124-135)
        return dict(kind=kind, lo=tuple(a for a, _ in self.grid_slice_tuple),
                    hi=tuple(b for _, b in self.grid_slice_tuple), exact=int(self.exact_interpolation),
                    time_to_idx=self.time_to_idx(config), reduce=0, components=(), axis=0, sign=1.0,
                    weight=0.0, scale=0.0, cos=None, sin=None, window=None, inverse=int(bool(self.inverse)))


@dataclass(kw_only=True)
class EnergyDetector(Detector):
    """energy.py:90-143: u = 1/2 (sum |E_i|^2 / inv_eps_i + sum |H_i|^2 / inv_mu_i)."""

    as_slices: bool = False
    reduce_volume: bool = False
    x_slice: float | None = None
    y_slice: float | None = None
    z_slice: float | None = None
    aggregate: str | None = None

    # as_slices planes: state key, the two kept axes, the collapsed axis (energy.py:76-88)
    _PLANES = (("XY Plane", (0, 1), 2), ("XZ Plane", (0, 2), 1), ("YZ Plane", (1, 2), 0))

    def state_layout(self, config):
        if self.as_slices and self.reduce_volume:
            raise Exception("Cannot both reduce volume and save slices!")
        n = self.num_time_steps_recorded
        gs = self.grid_shape
        if self.as_slices:
            return {key: ((n, gs[a], gs[b]), np.float32) for key, (a, b), _ in self._PLANES}
        return {"energy": ((n, 1) if self.reduce_volume else (n, *gs), np.float32)}

    def slice_index(self, axis, config):
        """energy.py:56-71 (uniform grid): the cell of the box holding the real position
        `{x,y,z}_slice`, clamped to the box; the middle cell when it is None."""
        pos = (self.x_slice, self.y_slice, self.z_slice)[axis]
        n = self.grid_shape[axis]
        if pos is None:
            return n // 2
        spacing = config.uniform_spacing()
        idx = int((pos - self.grid_slice_tuple[axis][0] * spacing) / spacing)
        return max(0, min(idx, n - 1))

    def lower(self, config):
        self._check()
        d = self._base_desc(config, KIND_ENERGY)
        d.update(reduce=int(self.reduce_volume), weight=config.uniform_spacing() ** 3)
        if not self.as_slices:
            return d
        # energy.py:107-132: the mean along each collapsed axis when aggregate == "mean" or a slice
        # position is missing, else the plane at each slice position
        use_mean = self.aggregate == "mean" or any(v is None for v in (self.x_slice, self.y_slice, self.z_slice))
        out = []
        for key, _, ax in self._PLANES:
            e = dict(d, key=key)
            if use_mean:
                e.update(reduce=2, axis=ax)
            else:
                lo, hi = list(d["lo"]), list(d["hi"])
                lo[ax] = d["lo"][ax] + self.slice_index(ax, config)
                hi[ax] = lo[ax] + 1
                e.update(reduce=0, lo=tuple(lo), hi=tuple(hi))
            out.append(e)
        return out


@dataclass(kw_only=True)
class FieldDetector(Detector):
    """field.py:31-63: records the selected (interpolated) field components, or with reduce_volume
    their volume-weighted mean over the box (detector.py:95-109)."""

    reduce_volume: bool = False
    components: Sequence[str] = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")

    def state_layout(self, config):
        n = self.num_time_steps_recorded
        nc = len(self.components)
        return {"fields": ((n, nc) if self.reduce_volume else (n, nc, *self.grid_shape), np.float32)}

    def lower(self, config):
        self._check()
        d = self._base_desc(config, KIND_FIELD)
        d.update(reduce=int(self.reduce_volume),
                 components=tuple(COMPONENTS.index(c) for c in COMPONENTS if c in self.components))
        return d


@dataclass(kw_only=True)
class PoyntingFluxDetector(Detector):
    """poynting_flux.py:171-195: Re(E x H*)[axis] (sign by direction), summed over the plane."""

    direction: str = "+"
    reduce_volume: bool = True
    fixed_propagation_axis: int | None = None
    keep_all_components: bool = False

    @property
    def propagation_axis(self) -> int:
        if self.fixed_propagation_axis is not None:
            if self.fixed_propagation_axis not in (0, 1, 2):
                raise Exception(f"Invalid: {self.fixed_propagation_axis=}")
            return self.fixed_propagation_axis
        if sum(a == 1 for a in self.grid_shape) != 1:
            raise Exception(f"Invalid poynting flux detector shape: {self.grid_shape}")
        return self.grid_shape.index(1)

    def state_layout(self, config):
        n = self.num_time_steps_recorded
        if self.keep_all_components:  # poynting_flux.py:161-162: the x, y and z components
            return {"poynting_flux": ((n, 3) if self.reduce_volume else (n, 3, *self.grid_shape), np.float32)}
        return {"poynting_flux": ((n, 1) if self.reduce_volume else (n, *self.grid_shape), np.float32)}

    def lower(self, config):
        self._check()
        d = self._base_desc(config, KIND_FLUX)
        d.update(reduce=int(self.reduce_volume), axis=self.propagation_axis,
                 sign=1.0 if self.direction == "+" else -1.0, weight=config.uniform_spacing() ** 2,
                 components=(0, 1, 2) if self.keep_all_components else ())
        return d


def _face_box(lo, hi, axis, side):
    """The one-cell face of box [lo, hi) normal to `axis` on `side` ("min" or "max")."""
    lo, hi = list(lo), list(hi)
    if side == "max":
        lo[axis] = hi[axis] - 1
    else:
        hi[axis] = lo[axis] + 1
    return tuple(lo), tuple(hi)


def _check_surface(orientation, axes):
    if orientation not in ("outward", "inward"):
        raise ValueError(f"orientation must be 'outward' or 'inward', got {orientation!r}")
    if axes is not None and any(a not in (0, 1, 2) for a in axes):
        raise ValueError(f"axes entries must be in (0, 1, 2), got {axes}")


@dataclass(kw_only=True)
class ClosedSurfacePoyntingFluxDetector(Detector):
    """poynting_flux.py:196-286 with net_poynting_flux_through_box (core/physics/metrics.py:120-157):
    the net outward flux through the faces of the box, sum over active axes a of (flux through
    the max face) - (flux through the min face), negated for orientation "inward". It lowers to
    one reduced flux detector per face, the first storing the state and the others adding to it
    in the reference's order (max face, then min face, per axis)."""

    orientation: str = "outward"
    axes: tuple | None = None

    def active_axes(self):
        if self.axes is not None:
            return tuple(self.axes)
        return tuple(a for a in range(3) if self.grid_shape[a] > 1)

    def state_layout(self, config):
        _check_surface(self.orientation, self.axes)
        return {"poynting_flux": ((self.num_time_steps_recorded, 1), np.float32)}

    def lower(self, config):
        self._check()
        _check_surface(self.orientation, self.axes)
        base = self._base_desc(config, KIND_FLUX)
        flip = -1.0 if self.orientation == "inward" else 1.0
        out = []
        for a in self.active_axes():
            for side, sgn in (("max", 1.0), ("min", -1.0)):
                d = dict(base)
                d["lo"], d["hi"] = _face_box(base["lo"], base["hi"], a, side)
                d.update(reduce=1, axis=a, sign=sgn * flip, weight=config.uniform_spacing() ** 2,
                         accumulate=int(bool(out)))
                out.append(d)
        return out


@dataclass(kw_only=True)
class PhasorDetector(Detector):
    """phasor.py:16-235: on-the-fly DFT, phasor += EH * exp(i w n dt) * scale * window, the window
    being the on-mask times the optional apodization (phasor.py:139-157); with reduce_volume the
    box mean of each component's phasor."""

    wave_characters: Sequence = ()
    reduce_volume: bool = False
    components: Sequence[str] = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    dtype: object = np.complex64
    scaling_mode: str = "continuous"
    dft_subsample: int | str = 1
    apodization: object = None

    def dft_stride(self, config) -> int:
        """phasor.py:92-103: the DFT sampling stride; "auto" keeps 12 samples per period of the
        highest recorded frequency."""
        sub = self.dft_subsample
        if isinstance(sub, str):
            if sub != "auto":
                raise Exception(f"Invalid dft_subsample: {sub!r}")
            dt = float(config.time_step_duration)
            f_max = max((abs(float(w.get_frequency())) for w in self.wave_characters), default=0.0)
            if f_max <= 0.0 or dt <= 0.0:
                return 1
            return max(1, math.floor(1.0 / (12 * f_max * dt)))
        return max(1, int(sub))

    def _on_list(self, config):
        on = super()._on_list(config)
        stride = self.dft_stride(config)
        if stride <= 1:
            return on
        active = [t for t, o in enumerate(on) if o]
        kept = [False] * len(on)
        for t in active[::stride]:
            kept[t] = True
        return kept

    def state_layout(self, config):
        if self.apodization is not None and not isinstance(self.apodization, TemporalWindow):
            raise TypeError(f"apodization must be a TemporalWindow, got {type(self.apodization).__name__}")
        nc = len([c for c in COMPONENTS if c in self.components])
        box = () if self.reduce_volume else tuple(self.grid_shape)
        return {"phasor": ((1, len(self.wave_characters), nc, *box), np.complex64)}

    def window(self, config) -> np.ndarray:
        """phasor.py:139-157: per-step weights, the on-mask times the apodization at t = n dt."""
        on = np.asarray(self._on_list(config), bool)
        if self.apodization is None:
            return on.astype(np.float32)
        t = (np.arange(config.time_steps_total, dtype=np.float32) * np.float32(config.time_step_duration))
        return (self.apodization.get_window(t.astype(np.float32)) * on.astype(np.float32)).astype(np.float32)

    def lower(self, config):
        self._check()
        T = config.time_steps_total
        dt = config.time_step_duration
        d = self._base_desc(config, KIND_PHASOR)
        on = np.asarray(self._on_list(config), bool)
        window = self.window(config)
        wsum = float(np.sum(window, dtype=np.float32))
        if not math.isfinite(wsum) or wsum <= 0:
            raise Exception(f"Detector '{self.name}': the apodization window sums to {wsum} over the recorded "
                            "time steps; it must be finite and positive. Check that the window's support "
                            "overlaps the detector's active interval.")
        if self.scaling_mode not in ("continuous", "pulse"):
            raise Exception(f"Invalid scaling mode: {self.scaling_mode=}")
        scale = 2 / wsum if self.scaling_mode == "continuous" else self.dft_stride(config)
        # phase = omega(float32) * (n * dt)(float32) (phasor.py:195-214)
        om = (2 * np.pi * np.asarray([w.get_frequency() for w in self.wave_characters])).astype(np.float32)
        t = (np.arange(T, dtype=np.float32) * np.float32(dt)).astype(np.float32)
        ph = (om[:, None] * t[None, :]).astype(np.float32)
        d.update(components=tuple(COMPONENTS.index(c) for c in COMPONENTS if c in self.components),
                 sign=-1.0 if self.inverse else 1.0,  # phasor.py:226-229: an inverse detector subtracts
                 reduce=int(self.reduce_volume), scale=float(np.float32(scale)), cos=np.cos(ph).astype(np.float32), sin=np.sin(ph).astype(np.float32),
                 window=window, time_to_idx=np.where(on, 0, -1).astype(np.int32))
        return d


def _phasor_poynting_vector(ph):
    """poynting_flux.py:58-71: Re(E x conj(H)) of phasors (num_freqs, 6, ...) -> (num_freqs, 3, ...)."""
    import torch

    ph = torch.as_tensor(ph)
    E, Hc = ph[:, :3], torch.conj(ph[:, 3:])
    return torch.stack([(E[:, (a + 1) % 3] * Hc[:, (a + 2) % 3] - E[:, (a + 2) % 3] * Hc[:, (a + 1) % 3]).real
                        for a in range(3)], dim=1)


@dataclass(kw_only=True)
class PhasorPoyntingFluxDetector(PhasorDetector):
    """poynting_flux.py:289-387: the phasors of all six components over a plane, and the
    time-averaged flux 1/2 Re(E(w) x H*(w)) through it formed after the run."""

    direction: str = "+"
    fixed_propagation_axis: int | None = None
    keep_all_components: bool = False

    def __post_init__(self):
        super().__post_init__()
        self.components = COMPONENTS
        self.reduce_volume = False

    @property
    def propagation_axis(self) -> int:
        return PoyntingFluxDetector.propagation_axis.fget(self)

    def compute_poynting_flux(self, state):
        """Real flux per frequency (num_freqs,), or (num_freqs, 3) with keep_all_components."""
        pv = _phasor_poynting_vector(state["phasor"][0])
        if self.direction == "-":
            pv = -pv
        w = float(np.float32(self._config.uniform_spacing() ** 2))
        if self.keep_all_components:
            flux = (pv * w).sum(dim=(2, 3, 4))
        else:
            flux = (pv[:, self.propagation_axis] * w).sum(dim=(1, 2, 3))
        return 0.5 * flux if self.scaling_mode == "continuous" else flux


@dataclass(kw_only=True)
class ClosedSurfacePhasorPoyntingFluxDetector(PhasorDetector):
    """poynting_flux.py:390-533: the phasors of all six components on the two boundary planes of
    every active axis (state entries "phasor_axis{a}_{min|max}", (1, num_freqs, 6, plane)), and
    the net time-averaged outward flux formed after the run. Like the reference, the face
    phasors take the static scale but not the apodization weight of each step."""

    orientation: str = "outward"
    axes: tuple | None = None

    def __post_init__(self):
        super().__post_init__()
        self.components = COMPONENTS
        self.reduce_volume = False

    def active_axes(self):
        return ClosedSurfacePoyntingFluxDetector.active_axes(self)

    def state_layout(self, config):
        _check_surface(self.orientation, self.axes)
        nf = len(self.wave_characters)
        out = {}
        for a in self.active_axes():
            plane = tuple(1 if i == a else self.grid_shape[i] for i in range(3))
            for side in ("min", "max"):
                out[f"phasor_axis{a}_{side}"] = ((1, nf, 6, *plane), np.complex64)
        return out

    def lower(self, config):
        _check_surface(self.orientation, self.axes)
        base = super().lower(config)
        base["window"] = (base["time_to_idx"] >= 0).astype(np.float32)
        out = []
        for a in self.active_axes():
            for side in ("min", "max"):
                d = dict(base)
                d["lo"], d["hi"] = _face_box(base["lo"], base["hi"], a, side)
                d["key"] = f"phasor_axis{a}_{side}"
                out.append(d)
        return out

    def compute_net_flux(self, state):
        """Real net outward flux per frequency (num_freqs,)."""
        import torch

        w = float(np.float32(self._config.uniform_spacing() ** 2))
        net = None
        for a in self.active_axes():
            for side, sgn in (("max", 1.0), ("min", -1.0)):
                s_a = _phasor_poynting_vector(state[f"phasor_axis{a}_{side}"][0])[:, a]
                term = sgn * (s_a * w).sum(dim=(1, 2, 3))
                net = term if net is None else net + term
        if net is None:
            net = torch.zeros(len(self.wave_characters))
        if self.orientation == "inward":
            net = -net
        return 0.5 * net if self.scaling_mode == "continuous" else net


# ---------------------------------------------------------------- mode overlap (detectors/mode.py)
def plane_centers(config, grid_slice_tuple, volume_shape):
    """Cell-centre coordinates (X, Y, Z) of a placed box as (ij-indexed) meshgrids: the uniform
    grid's origin is the domain centre shifted by the grid's `center` (core/grid.py:13-47)."""
    dx = config.uniform_spacing()
    vol = volume_shape
    axes = []
    for a in range(3):
        lo, hi = grid_slice_tuple[a]
        origin = float(config.grid.center[a]) - vol[a] * dx / 2.0
        axes.append(origin + (np.arange(lo, hi) + 0.5) * dx)
    return np.meshgrid(*axes, indexing="ij")


def integrated_power(E, H, axis, area_weights=None):
    """|sum 1/2 Re((conj E x H)_axis) w| over a plane (metrics.py:160-189), the weights taken
    relative to their mean."""
    E = np.asarray(E, np.complex128)
    H = np.asarray(H, np.complex128)
    a1, a2 = (axis + 1) % 3, (axis + 2) % 3
    s = 0.5 * np.real(np.conj(E[a1]) * H[a2] - np.conj(E[a2]) * H[a1])
    if area_weights is not None:
        s = s * (np.asarray(area_weights) / np.mean(area_weights))
    return abs(float(np.sum(s)))


def normalize_by_poynting_flux(E, H, axis, area_weights=None):
    """E and H scaled so their integrated Poynting flux along `axis` is one (metrics.py:192-221)."""
    f = math.sqrt(integrated_power(E, H, axis, area_weights))
    return np.asarray(E) / f, np.asarray(H) / f


def gaussian_mode_fields(coordinates, propagation_axis, *, radius, direction, polarization_axis=None,
                         fixed_E_polarization_vector=None, fixed_H_polarization_vector=None, azimuth_angle=0.0,
                         elevation_angle=0.0, divergence_angle=0.0, center=(0.0, 0.0), wavelength,
                         refractive_index=1.0):
    """Cross-section of a Gaussian beam on a plane (detectors/mode.py:26-141): amplitude
    exp(-r^2 / radius^2) around `center` (the two transverse axes in ascending order), a paraxial
    wavefront phase k r^2 tan(divergence) / (2 radius) with k = 2 pi n / wavelength, E along the
    source polarisation and H = n (k_hat x E) (mu_r = 1). Normal incidence only (the engine's plane
    sources are not tilted either)."""
    if radius <= 0:
        raise ValueError(f"radius must be positive, got {radius}")
    if azimuth_angle or elevation_angle:
        raise NotImplementedError("tilted Gaussian reference modes are not supported by this engine")
    if polarization_axis is not None and (fixed_E_polarization_vector is not None
                                          or fixed_H_polarization_vector is not None):
        raise ValueError("Specify either polarization_axis or fixed_E/H_polarization_vector, not both")
    t0, t1 = [a for a in range(3) if a != propagation_axis]
    if fixed_E_polarization_vector is None and fixed_H_polarization_vector is None:
        pol = t0 if polarization_axis is None else polarization_axis
        if pol == propagation_axis:
            raise ValueError(f"polarization_axis ({pol}) must be transverse to the propagation axis "
                             f"({propagation_axis})")
        fixed_E_polarization_vector = tuple(1.0 if a == pol else 0.0 for a in range(3))
    k = np.zeros(3)
    k[propagation_axis] = 1.0 if direction == "+" else -1.0
    e = None if fixed_E_polarization_vector is None else np.asarray(fixed_E_polarization_vector, float)
    h = None if fixed_H_polarization_vector is None else np.asarray(fixed_H_polarization_vector, float)
    e = None if e is None else e / np.linalg.norm(e)
    h = None if h is None else h / np.linalg.norm(h)
    e = np.cross(h, k) if e is None else e
    h = np.cross(k, e) if h is None else h
    u = np.asarray(coordinates[t0], np.float64) - center[0]
    v = np.asarray(coordinates[t1], np.float64) - center[1]
    r2 = u * u + v * v
    amp = np.exp(-r2 / radius**2).astype(np.complex128)
    if divergence_angle:
        kn = 2.0 * math.pi * float(refractive_index) / float(wavelength)
        amp = amp * np.exp(1j * 0.5 * kn * (math.tan(math.radians(divergence_angle)) / radius) * r2)
    mode_E = amp[None] * e[:, None, None, None]
    mode_H = amp[None] * (float(refractive_index) * h)[:, None, None, None]
    return mode_E, mode_H


def gaussian_mode_function(*, radius, direction, polarization_axis=None, fixed_E_polarization_vector=None,
                           fixed_H_polarization_vector=None, azimuth_angle=0.0, elevation_angle=0.0,
                           divergence_angle=0.0, center=(0.0, 0.0)):
    """A mode_function for CustomModeOverlapDetector evaluating gaussian_mode_fields with the
    plane's mean refractive index (detectors/mode.py:144-190)."""

    def mode_function(*, coordinates, frequency, propagation_axis, inv_permittivity):
        n = math.sqrt(float(np.mean(1.0 / np.asarray(inv_permittivity, np.float64))))
        return gaussian_mode_fields(coordinates, propagation_axis, radius=radius, direction=direction,
                                    polarization_axis=polarization_axis,
                                    fixed_E_polarization_vector=fixed_E_polarization_vector,
                                    fixed_H_polarization_vector=fixed_H_polarization_vector,
                                    azimuth_angle=azimuth_angle, elevation_angle=elevation_angle,
                                    divergence_angle=divergence_angle, center=center,
                                    wavelength=_C0 / frequency, refractive_index=n)

    return mode_function


_C0 = 299792458.0


@dataclass(kw_only=True)
class BaseModeOverlapDetector(PhasorDetector):
    """detectors/mode.py:193-447: the phasors of all six components over a plane, and after the run
    the overlap of each frequency's phasors with a reference mode (mode_E, mode_H) of that
    frequency, sum over the plane of ((mode_E x conj H) + (conj E x mode_H))_axis times the face
    weights (relative to their mean), divided by 4 unless scaling_mode is "pulse". The reference
    modes are evaluated by `apply` (before the run) with the plane's inverse permittivity, the
    effective one at each carrier frequency in dispersive media."""

    def __post_init__(self):
        super().__post_init__()
        self.components = COMPONENTS
        self.reduce_volume = False

    @property
    def propagation_axis(self) -> int:
        if sum(1 for s in self.grid_shape if s == 1) != 1:
            raise Exception(f"Invalid ModeOverlapDetector shape: {self.grid_shape}")
        return self.grid_shape.index(1)

    def _face_area_weights(self):
        return np.ones(self.grid_shape)  # uniform grid: equal faces, relative to their mean

    def _compute_mode_fields(self, *, frequency, inv_permittivity):  # pragma: no cover
        raise NotImplementedError

    def apply(self, key, inv_permittivities, inv_permeabilities, dispersive=None, **kwargs):
        self._volume_shape = tuple(np.asarray(inv_permittivities).shape[1:])
        if self.symmetry_touched():
            raise NotImplementedError(f"mode-overlap detector '{self.name}' on a symmetry plane is not supported")
        ie = np.asarray(inv_permittivities)[(slice(None), *self.grid_slice)].astype(np.float64)
        if ie.shape[0] == 9:
            ie = ie[[0, 4, 8]]
        im = inv_permeabilities
        self._inv_mu_plane = (float(im) if np.isscalar(im) or np.ndim(im) == 0 else
                              np.asarray(im)[(slice(None), *self.grid_slice)].astype(np.float64))
        from .sources import effective_inv_permittivity

        Es, Hs = [], []
        for w in self.wave_characters:
            f = float(w.get_frequency())
            ie_f = ie
            if dispersive is not None:
                c1, c2, c3 = dispersive
                ie_f = np.asarray(effective_inv_permittivity(ie, c1, c2, c3, 2.0 * math.pi * f,
                                                             self._config.time_step_duration), np.float64)
            mE, mH = self._compute_mode_fields(frequency=f, inv_permittivity=ie_f)
            Es.append(np.asarray(mE, np.complex128))
            Hs.append(np.asarray(mH, np.complex128))
        self._mode_E = np.stack(Es)
        self._mode_H = np.stack(Hs)
        return self

    def symmetry_touched(self) -> bool:
        cfg = self._config
        return cfg is not None and any(cfg.symmetry[a] and self.touches_symmetry_plane(a) for a in range(3))

    def compute_overlap_to_mode(self, state, mode_E, mode_H, wave_character_index=0):
        """Complex overlap of one reference mode with the phasors of one frequency
        (detectors/mode.py:379-421)."""
        import torch

        ph = state["phasor"][0, wave_character_index].to(torch.complex128)
        dev = ph.device
        mE = torch.as_tensor(np.asarray(mode_E, np.complex128), device=dev)
        mH = torch.as_tensor(np.asarray(mode_H, np.complex128), device=dev)
        E, H = ph[:3], ph[3:]
        a = self.propagation_axis
        a1, a2 = (a + 1) % 3, (a + 2) % 3
        Hc, Ec = torch.conj(H), torch.conj(E)
        integrand = (mE[a1] * Hc[a2] - mE[a2] * Hc[a1]) + (Ec[a1] * mH[a2] - Ec[a2] * mH[a1])
        w = torch.as_tensor(self._face_area_weights(), dtype=torch.float64, device=dev)
        alpha = torch.sum(integrand * w)
        return alpha if self.scaling_mode == "pulse" else alpha / 4.0

    def compute_overlap(self, state):
        """Overlaps at every wave character, complex (num_freqs,) (detectors/mode.py:423-447)."""
        import torch

        if getattr(self, "_mode_E", None) is None:
            raise Exception("Need to call apply on the mode-overlap detector before calling compute_overlap!")
        return torch.stack([self.compute_overlap_to_mode(state, self._mode_E[i], self._mode_H[i], i)
                            for i in range(len(self.wave_characters))])


@dataclass(kw_only=True)
class CustomModeOverlapDetector(BaseModeOverlapDetector):
    """detectors/mode.py:598-665: the reference mode of each frequency comes from
    mode_function(coordinates=(X, Y, Z) cell centres, frequency, propagation_axis,
    inv_permittivity) -> (mode_E, mode_H), each (3, *plane), normalised to unit Poynting flux
    when `normalize`."""

    mode_function: object = None
    normalize: bool = True

    def _compute_mode_fields(self, *, frequency, inv_permittivity):
        if self.mode_function is None:
            raise ValueError(f"CustomModeOverlapDetector '{self.name}' needs a mode_function")
        mE, mH = self.mode_function(coordinates=plane_centers(self._config, self.grid_slice_tuple, self._volume_shape),
                                    frequency=frequency, propagation_axis=self.propagation_axis,
                                    inv_permittivity=inv_permittivity)
        if self.normalize:
            mE, mH = normalize_by_poynting_flux(mE, mH, self.propagation_axis, self._face_area_weights())
        return mE, mH


@dataclass(kw_only=True)
class GaussianModeOverlapDetector(BaseModeOverlapDetector):
    """detectors/mode.py:666-798: an analytic Gaussian cross-section (gaussian_mode_fields) with
    the plane's mean refractive index as the reference mode; mode_radius is the 1/e amplitude
    radius on the plane."""

    mode_radius: float = 0.0
    direction: str = "+"
    polarization_axis: int | None = None
    fixed_E_polarization_vector: tuple | None = None
    fixed_H_polarization_vector: tuple | None = None
    azimuth_angle: float = 0.0
    elevation_angle: float = 0.0
    divergence_angle: float = 0.0
    center: tuple = (0.0, 0.0)
    normalize: bool = True

    def place_on_grid(self, grid_slice_tuple, config, key=None):
        obj = super().place_on_grid(grid_slice_tuple, config, key)
        if obj.mode_radius <= 0:
            raise ValueError(f"mode_radius must be positive, got {obj.mode_radius}")
        return obj

    def _compute_mode_fields(self, *, frequency, inv_permittivity):
        n = math.sqrt(float(np.mean(1.0 / inv_permittivity)))
        mE, mH = gaussian_mode_fields(plane_centers(self._config, self.grid_slice_tuple, self._volume_shape), self.propagation_axis,
                                      radius=self.mode_radius, direction=self.direction,
                                      polarization_axis=self.polarization_axis,
                                      fixed_E_polarization_vector=self.fixed_E_polarization_vector,
                                      fixed_H_polarization_vector=self.fixed_H_polarization_vector,
                                      azimuth_angle=self.azimuth_angle, elevation_angle=self.elevation_angle,
                                      divergence_angle=self.divergence_angle, center=self.center,
                                      wavelength=_C0 / frequency, refractive_index=n)
        if self.normalize:
            mE, mH = normalize_by_poynting_flux(mE, mH, self.propagation_axis, self._face_area_weights())
        return mE, mH


@dataclass(kw_only=True)
class ModeOverlapDetector(BaseModeOverlapDetector):
    """detectors/mode.py:448-597: the reference mode of each frequency is mode `mode_index` (after
    the optional polarization filter) of the detector plane, from the engine's mode solver
    (`fdtdx_amd.modes.compute_mode`, in place of the reference's tidy3d call) with the plane's
    inverse permittivity and permeability; normalised to unit Poynting flux."""

    direction: str = "+"
    mode_index: int = 0
    filter_pol: str | None = None
    symmetry: tuple = (0, 0)
    bend_radius: float | None = None
    bend_axis: int | None = None

    def place_on_grid(self, grid_slice_tuple, config, key=None):
        obj = super().place_on_grid(grid_slice_tuple, config, key)
        if (obj.bend_radius is None) != (obj.bend_axis is None):
            raise ValueError("bend_radius and bend_axis must both be set or both be None")
        if obj.bend_radius is not None:
            raise NotImplementedError("bent-waveguide modes (a conformal transformation in the solver) are not "
                                      "supported by this engine")
        return obj

    def _compute_mode_fields(self, *, frequency, inv_permittivity):
        from .modes import compute_mode

        mE, mH, _ = compute_mode(frequency, inv_permittivity, self._inv_mu_plane, self._config.uniform_spacing(),
                                 direction=self.direction, mode_index=self.mode_index, filter_pol=self.filter_pol,
                                 symmetry=self.symmetry)
        return mE, mH
