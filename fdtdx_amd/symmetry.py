"""Full-domain views of a mirror-symmetry-reduced run.

A run reduced by ``SimulationConfig.symmetry`` keeps the upper half of every axis with a nonzero
entry (-1: an electric mirror plane, where a PEC face sits; +1: a magnetic one). After the run,
``unfold_fields`` and ``unfold_detector_states`` rebuild what the full domain would hold; the
reference does this in src/fdtdx/fdtd/symmetry.py:311-753 with the parity table of
src/fdtdx/core/physics/symmetry.py:18-239, and the results here equal it.

Model. On a uniform Yee grid every sample has a half-cell offset of 0 or 1/2 along each axis:
E_c is offset along c only, H_c along the two axes other than c. Mirroring about a plane normal
to axis a maps a component onto itself times a parity. The parity follows from one rule: a
sample with offset 0 along a (it lies in the plane: tangential E, normal H) is odd about an
electric plane and even about a magnetic one, and a sample with offset 1/2 the opposite, i.e.
parity = wall for offset 0 and -wall for offset 1/2.

Index map. The full axis of length 2n is one gather from the kept rows: output row u >= n is
kept row u - n; output row u < n is the mirror of kept row m + shift, m = n - 1 - u, times the
parity. shift is 1 for samples lying in an electric plane (their row 0 is its own mirror, so the
mirror of row m is row m + 1; the outermost output row has no partner and repeats its
neighbour, row n - 1), 0 otherwise (a magnetic plane pairs row r with row -1 - r for every
component). One `index_select` and one multiplication per axis do the unfolding, on the device.

Detector outputs follow the same map along their spatial axes. Exactly interpolated detectors
sample at the E_z position (offset 0 along x and y), so about an electric x or y plane they pair
like in-plane samples; point-sampled ones take the plain map. Reduced outputs are scaled: a box
sum of a component grows by (1 + parity) per unfolded axis and a box mean by (1 + parity) / 2,
so even parts double or stay and odd parts vanish; energy densities are even.
"""

from __future__ import annotations

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _offset(field_type: str, component: int, axis: int) -> float:
    """Half-cell offset (0 or 0.5) of field component `component` along `axis`."""
    if field_type == "E":
        return 0.5 if component == axis else 0.0
    if field_type == "H":
        return 0.0 if component == axis else 0.5
    raise ValueError(f"field type {field_type!r}: expected 'E' or 'H'")


def field_component_parity(field_type: str, component: int, axis: int, wall: int) -> int:
    """+1 when the component is even under the mirror normal to `axis`, -1 when it is odd; `wall`
    is -1 for an electric (PEC) plane and +1 for a magnetic (PMC) one
    (core/physics/symmetry.py:18-53)."""
    if wall not in (-1, 1):
        raise ValueError(f"mirror wall {wall!r}: expected -1 (electric) or +1 (magnetic)")
    return wall if _offset(field_type, component, axis) == 0.0 else -wall


def component_sits_on_plane(field_type: str, component: int, axis: int) -> bool:
    """True when the component is sampled in planes of integer coordinate along `axis`
    (core/physics/symmetry.py:56-78)."""
    return _offset(field_type, component, axis) == 0.0


def _mirror_index(n: int, shift: int, device):
    """Source rows of the 2n output rows (see the module docstring) and a mask of the mirror half."""
    import torch

    u = torch.arange(2 * n, device=device)
    m = (n - 1 - u).clamp(min=0)
    src = torch.where(u >= n, u - n, (m + shift).clamp(max=n - 1))
    return src, u < n


def _unfold_dim(t, dim: int, shift: int, sign):
    """Unfold tensor t along `dim`: the gather of `_mirror_index`, the mirror half multiplied by
    `sign` (a number, or a tensor broadcasting against t)."""
    import torch

    n = t.shape[dim]
    src, mirror = _mirror_index(n, shift, t.device)
    out = t.index_select(dim, src)
    shape = [1] * t.dim()
    shape[dim] = 2 * n
    mirror = mirror.reshape(shape)
    if not torch.is_tensor(sign):
        sign = torch.tensor(float(sign), dtype=out.real.dtype, device=t.device)
    one = torch.ones((), dtype=sign.dtype, device=t.device)
    return out * torch.where(mirror, sign, one).to(out.dtype)


def _planes(symmetry):
    planes = [(a, int(w)) for a, w in enumerate(symmetry) if w]
    if not planes:
        raise ValueError("unfolding needs a symmetry-reduced run: config.symmetry has no mirror plane "
                         f"({tuple(symmetry)})")
    return planes


def unfold_fields(field, symmetry, field_type: str):
    """The full-domain (3, Nx, Ny, Nz) E or H field of a symmetry-reduced run
    (fdtd/symmetry.py:319-366)."""
    import torch

    _offset(field_type, 0, 0)
    t = torch.as_tensor(field)
    for a, wall in _planes(symmetry):
        comps = []
        for c in range(3):
            shift = 1 if wall == -1 and component_sits_on_plane(field_type, c, a) else 0
            comps.append(_unfold_dim(t[c], a, shift, field_component_parity(field_type, c, a, wall)))
        t = torch.stack(comps)
    return t


def _component_spec(names):
    return [("E" if nm[0] == "E" else "H", "xyz".index(nm[1])) for nm in names]


def _parity_vector(spec, axis, wall, like, ndim, comp_dim):
    """Per-component parities about one plane as a tensor along comp_dim of an ndim array."""
    import torch

    shape = [1] * ndim
    shape[comp_dim] = len(spec)
    vals = [float(field_component_parity(ft, c, axis, wall)) for ft, c in spec]
    return torch.tensor(vals, dtype=like.real.dtype, device=like.device).reshape(shape)


def _flux_parity(component, axis, wall):
    """Parity of the Poynting component S_i = E_j H_k - E_k H_j about the plane normal to `axis`
    (both products share it)."""
    j, k = [x for x in range(3) if x != component]
    return field_component_parity("E", j, axis, wall) * field_component_parity("H", k, axis, wall)


def _box_scale(parities, mean):
    """Factor of a reduced value over the unfolded box: prod (1 + p) for a sum, prod (1 + p)/2 for a
    mean (fdtd/symmetry.py:522-551)."""
    f = 1.0
    for p in parities:
        f *= (1 + p) / 2 if mean else (1 + p)
    return f


class _DetectorUnfold:
    """Unfolding of one detector's state about the planes it straddles."""

    def __init__(self, detector, planes):
        self.det = detector
        self.planes = planes  # [(axis, wall)]
        self.paired = {a for a, w in planes if w == -1 and a in (0, 1) and getattr(detector, "exact_interpolation", False)}

    def spatial(self, arr, first_spatial_dim, signs_for):
        """Unfold the three spatial dims starting at first_spatial_dim; signs_for(axis, wall) gives the
        mirror-half sign (number or tensor)."""
        for a, wall in self.planes:
            arr = _unfold_dim(arr, first_spatial_dim + a, 1 if a in self.paired else 0, signs_for(a, wall))
        return arr

    def scaled(self, arr, comp_dim, parity_of, ncomp, mean):
        """A reduced value per component along comp_dim, scaled by _box_scale."""
        import torch

        f = [_box_scale([parity_of(ci, a, w) for a, w in self.planes], mean) for ci in range(ncomp)]
        shape = [1] * arr.dim()
        shape[comp_dim] = ncomp
        return arr * torch.tensor(f, dtype=arr.real.dtype, device=arr.device).reshape(shape).to(arr.dtype)


def _unfold_state(detector, state, planes):
    from .detectors import EnergyDetector, FieldDetector, PhasorDetector, PoyntingFluxDetector

    u = _DetectorUnfold(detector, planes)
    if isinstance(detector, EnergyDetector):
        arr = state["energy"]
        if detector.reduce_volume:
            return {"energy": arr * (2 ** len(planes))}
        return {"energy": u.spatial(arr, 1, lambda a, w: 1.0)}
    if isinstance(detector, PoyntingFluxDetector):
        arr = state["poynting_flux"]
        comps = (0, 1, 2) if detector.keep_all_components else (detector.propagation_axis,)
        par = lambda ci, a, w: _flux_parity(comps[ci], a, w)
        if detector.reduce_volume:
            return {"poynting_flux": u.scaled(arr, 1, par, len(comps), mean=False) if len(comps) == 3
                    else arr * _box_scale([par(0, a, w) for a, w in planes], False)}
        if len(comps) == 3:
            import torch

            def signs(a, w):
                v = [float(_flux_parity(i, a, w)) for i in comps]
                return torch.tensor(v, dtype=arr.dtype, device=arr.device).reshape(1, 3, 1, 1, 1)

            return {"poynting_flux": u.spatial(arr, 2, signs)}
        return {"poynting_flux": u.spatial(arr, 1, lambda a, w: float(par(0, a, w)))}
    if isinstance(detector, (PhasorDetector, FieldDetector)):
        key, comp_dim = ("phasor", 2) if isinstance(detector, PhasorDetector) else ("fields", 1)
        if key not in state or len(state) != 1:
            raise NotImplementedError(f"detector {detector.name!r}: no unfolding for a state with keys "
                                      f"{sorted(state)}")
        arr = state[key]
        spec = _component_spec([nm for nm in COMPONENTS if nm in detector.components])
        if detector.reduce_volume:
            return {key: u.scaled(arr, comp_dim, lambda ci, a, w: field_component_parity(*spec[ci], a, w),
                                  len(spec), mean=True)}
        return {key: u.spatial(arr, comp_dim + 1,
                               lambda a, w: _parity_vector(spec, a, w, arr, arr.dim(), comp_dim).to(arr.dtype))}
    raise NotImplementedError(f"detector {detector.name!r} ({type(detector).__name__}): its state has no "
                              "unfolding rule; unfold the fields with unfold_fields")


def unfold_detector_states(arrays, objects, config):
    """A copy of `arrays` whose detector states are the full-domain ones of a symmetry-reduced run
    (fdtd/symmetry.py:699-753). Only detectors that the reduction clipped (they extended into the
    discarded half) change; a detector lying in the kept half, even one starting on a plane, is
    returned as it is."""
    import copy

    _planes(config.symmetry)
    dets = {d.name: d for d in objects.detectors}
    out = {}
    for name, state in arrays.detector_states.items():
        det = dets.get(name)
        planes = [] if det is None else [(a, int(config.symmetry[a])) for a in range(3)
                                          if config.symmetry[a] and det.straddles_symmetry_plane(a)]
        out[name] = _unfold_state(det, state, planes) if planes else state
    new = copy.copy(arrays)
    new.detector_states = out
    return new
