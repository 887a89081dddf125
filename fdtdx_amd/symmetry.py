"""Mirror-symmetry unfolding (reference: src/fdtdx/fdtd/symmetry.py:311-753 and
src/fdtdx/core/physics/symmetry.py:18-240).

A run reduced by ``SimulationConfig.symmetry`` (-1 electric, +1 magnetic plane at the centre of
an axis; the upper half is kept) is turned back into full-domain arrays after the run: fields
and spatial detector outputs are mirrored with each component's parity and concatenated in
front of the kept half, reduced sums and means are rescaled (an even component doubles a sum
and keeps a mean, an odd one vanishes). Across an electric plane a component sampled on the
plane (tangential E, normal H) has its first sample as its own mirror; across a magnetic plane
every component mirrors one-to-one. These are device-side torch expressions over the detector
states, outside the time loop.
"""

from __future__ import annotations

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
_SPEC = (("E", 0), ("E", 1), ("E", 2), ("H", 0), ("H", 1), ("H", 2))


def field_component_parity(field_type: str, component: int, axis: int, wall: int) -> int:
    """core/physics/symmetry.py:18-53: +1 (even) or -1 (odd) under the mirror normal to `axis`."""
    normal = component == axis
    if wall == -1:
        if field_type == "E":
            return 1 if normal else -1
        return -1 if normal else 1
    if wall == 1:
        if field_type == "E":
            return -1 if normal else 1
        return 1 if normal else -1
    raise ValueError(f"wall must be -1 (PEC) or +1 (PMC), got {wall}")


def component_sits_on_plane(field_type: str, component: int, axis: int) -> bool:
    """core/physics/symmetry.py:56-78."""
    if field_type == "E":
        return component != axis
    if field_type == "H":
        return component == axis
    raise ValueError(f"field_type must be 'E' or 'H', got {field_type!r}")


def mirror_pairs_on_plane(field_type: str, component: int, axis: int, wall: int) -> bool:
    """core/physics/symmetry.py:81-108: samples pair about a shared row only across an electric
    plane, for components sampled on it."""
    return wall == -1 and component_sits_on_plane(field_type, component, axis)


def _slice_axis(t, axis, start, stop=None):
    idx = [slice(None)] * t.dim()
    idx[axis] = slice(start, stop)
    return t[tuple(idx)]


def mirror_extend_low_side(arr, axis: int, parity: float, on_plane: bool):
    """core/physics/symmetry.py:217-239: the mirrored (discarded) half in ascending order; on-plane
    samples skip their own row and repeat the outermost neighbour."""
    import torch

    if not on_plane:
        return parity * torch.flip(arr, dims=(axis,))
    mirrored = parity * torch.flip(_slice_axis(arr, axis, 1), dims=(axis,))
    return torch.cat([_slice_axis(mirrored, axis, 0, 1), mirrored], dim=axis)


def _check_has_symmetry(symmetry):
    if not any(s != 0 for s in symmetry):
        raise ValueError("Nothing to unfold: this simulation has no symmetry (config.symmetry is (0, 0, 0)). "
                         "The unfold helpers are only meaningful for symmetry-reduced simulations.")


def unfold_fields(field, symmetry, field_type: str):
    """fdtd/symmetry.py:319-366: the full-domain (3, Nx, Ny, Nz) field of a reduced one."""
    import torch

    if field_type not in ("E", "H"):
        raise ValueError(f"field_type must be 'E' or 'H', got {field_type!r}")
    _check_has_symmetry(symmetry)
    arr = torch.as_tensor(field)
    for a in range(3):
        if symmetry[a] == 0:
            continue
        comps = []
        for c in range(3):
            single = arr[c:c + 1]
            low = mirror_extend_low_side(single, a + 1, field_component_parity(field_type, c, a, symmetry[a]),
                                         mirror_pairs_on_plane(field_type, c, a, symmetry[a]))
            comps.append(torch.cat([low, single], dim=a + 1))
        arr = torch.cat(comps, dim=0)
    return arr


def unfold_array(arr, symmetry, spatial_axes, signs=None, on_plane_axes=()):
    """fdtd/symmetry.py:434-474: mirror and concatenate along each symmetric axis (array axis
    spatial_axes[a]), the mirror image times signs[a]."""
    import torch

    _check_has_symmetry(symmetry)
    for a in range(3):
        if symmetry[a] == 0:
            continue
        ax = spatial_axes[a]
        sign = 1.0 if signs is None else signs.get(a, 1.0)
        if a in on_plane_axes:
            mirror = mirror_extend_low_side(arr, ax, 1.0, True) * sign
        else:
            mirror = torch.flip(arr, dims=(ax,)) * sign
        arr = torch.cat([mirror, arr], dim=ax)
    return arr


def _colocated_on_plane_axes(detector, touched):
    """fdtd/symmetry.py:477-496: exactly interpolated samples sit at the E_z point, integer along x
    and y, so they lie on an electric x or y plane; point-sampled ones fall back to the flip."""
    if not getattr(detector, "exact_interpolation", False):
        return ()
    return tuple(a for a in (0, 1) if touched[a] == -1)


def _component_signs(spec, touched, ndim, component_axis, like):
    import torch

    out = {}
    for a in range(3):
        if touched[a] == 0:
            continue
        shape = [1] * ndim
        shape[component_axis] = len(spec)
        out[a] = torch.tensor([float(field_component_parity(ft, ca, a, touched[a])) for ft, ca in spec],
                              dtype=like.real.dtype, device=like.device).reshape(shape)
    return out


def _reduce_factor(parities, ndim, component_axis, mean, like):
    """fdtd/symmetry.py:522-551: prod (1 + p) for a sum, prod (1 + p)/2 for a mean, per component."""
    import torch

    factors = []
    for per_axis in parities:
        f = 1.0
        for p in per_axis:
            f *= (1 + p) / 2 if mean else (1 + p)
        factors.append(f)
    shape = [1] * ndim
    shape[component_axis] = len(factors)
    return torch.tensor(factors, dtype=like.real.dtype, device=like.device).reshape(shape)


def _poynting_parity(component, axis, wall):
    """fdtd/symmetry.py:554-561: parity of (E x H)_component."""
    j, k = (x for x in range(3) if x != component)
    return field_component_parity("E", j, axis, wall) * field_component_parity("H", k, axis, wall)


def _unfold_poynting(detector, state, touched):
    """fdtd/symmetry.py:564-591."""
    import torch

    arr = state["poynting_flux"]
    on_plane = _colocated_on_plane_axes(detector, touched)
    comps = (0, 1, 2) if detector.keep_all_components else (detector.propagation_axis,)
    axes = [a for a in range(3) if touched[a] != 0]
    if detector.reduce_volume:
        parities = [[_poynting_parity(i, a, touched[a]) for a in axes] for i in comps]
        if detector.keep_all_components:
            return {"poynting_flux": arr * _reduce_factor(parities, arr.dim(), 1, False, arr)}
        scalar = 1.0
        for p in parities[0]:
            scalar *= 1 + p
        return {"poynting_flux": arr * scalar}
    if detector.keep_all_components:
        signs = {a: torch.tensor([float(_poynting_parity(i, a, touched[a])) for i in comps], dtype=arr.dtype,
                                 device=arr.device).reshape(1, 3, 1, 1, 1) for a in axes}
        return {"poynting_flux": unfold_array(arr, touched, (2, 3, 4), signs, on_plane)}
    p = detector.propagation_axis
    signs = {a: float(_poynting_parity(p, a, touched[a])) for a in axes}
    return {"poynting_flux": unfold_array(arr, touched, (1, 2, 3), signs, on_plane)}


def _unfold_one(detector, state, touched, count):
    """fdtd/symmetry.py:628-696."""
    from .detectors import EnergyDetector, FieldDetector, PhasorDetector, PoyntingFluxDetector

    if isinstance(detector, PhasorDetector) and len(state) == 1:
        spec = [_SPEC[i] for i, n in enumerate(COMPONENTS) if n in detector.components]
        arr = state["phasor"]
        if detector.reduce_volume:  # (1, freqs, comps): box means
            par = [[field_component_parity(ft, ca, a, touched[a]) for a in range(3) if touched[a]] for ft, ca in spec]
            return {"phasor": arr * _reduce_factor(par, arr.dim(), 2, True, arr).to(arr.dtype)}
        signs = _component_signs(spec, touched, arr.dim(), 2, arr)
        return {"phasor": unfold_array(arr, touched, (3, 4, 5), {a: v.to(arr.dtype) for a, v in signs.items()},
                                       _colocated_on_plane_axes(detector, touched))}
    if isinstance(detector, FieldDetector):
        spec = [_SPEC[i] for i, n in enumerate(COMPONENTS) if n in detector.components]
        arr = state["fields"]
        if detector.reduce_volume:  # (T, comps): box means
            par = [[field_component_parity(ft, ca, a, touched[a]) for a in range(3) if touched[a]] for ft, ca in spec]
            return {"fields": arr * _reduce_factor(par, arr.dim(), 1, True, arr)}
        signs = _component_signs(spec, touched, arr.dim(), 1, arr)
        return {"fields": unfold_array(arr, touched, (2, 3, 4), signs, _colocated_on_plane_axes(detector, touched))}
    if isinstance(detector, EnergyDetector):
        if detector.reduce_volume:  # energy density is even: x2 per plane
            return {"energy": state["energy"] * (2 ** count)}
        return {"energy": unfold_array(state["energy"], touched, (1, 2, 3),
                                       on_plane_axes=_colocated_on_plane_axes(detector, touched))}
    if isinstance(detector, PoyntingFluxDetector):
        return _unfold_poynting(detector, state, touched)
    raise NotImplementedError(f"unfold_detector_states does not know how to unfold detector type "
                              f"{type(detector).__name__!r}. Unfold the fields with unfold_fields instead.")


def unfold_detector_states(arrays, objects, config):
    """fdtd/symmetry.py:699-753: full-domain detector states of a reduced run. A detector that the
    reduction did not clip (it lies wholly in the kept half, even if it starts on a plane) is
    returned unchanged."""
    import copy

    _check_has_symmetry(config.symmetry)
    by_name = {d.name: d for d in objects.detectors}
    new = {}
    for name, state in arrays.detector_states.items():
        det = by_name.get(name)
        if det is None:
            new[name] = state
            continue
        touched = tuple(config.symmetry[a] if det.straddles_symmetry_plane(a) else 0 for a in range(3))
        count = sum(1 for t in touched if t)
        new[name] = state if count == 0 else _unfold_one(det, state, touched, count)
    out = copy.copy(arrays)
    out.detector_states = new
    return out
