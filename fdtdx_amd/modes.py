"""Guided modes of a plane cross-section, solved on the FDTD grid's own Yee lattice.

The reference hands the cross-section to tidy3d's mode solver (core/physics/modes.py:101-385,
579-730; tidy3d >= 2.8 in pyproject.toml, not installed here). This module solves the same
eigenproblem without it: the time-harmonic Maxwell equations (e^{+j w t}, fields e^{-j beta s}
along the propagation axis s) discretised with exactly the differences the FDTD sweeps use on
the plane, i.e. forward differences for curl E (zero beyond the high face) and backward
differences for curl H (zero beyond the low face). A mode found here is therefore an exact
eigenmode of the engine's spatial operator restricted to a waveguide that is uniform along s:
repeated plane by plane with the phase e^{-j beta s}, with 2 sin(beta / 2) in place of the
continuous beta (cell units) on the s differences, it satisfies every curl equation of the grid.

Transverse axes (u, v) = ((s + 1) % 3, (s + 2) % 3), a right-handed (u, v, s). Fields are in the
engine's units: E, and H scaled by the free-space impedance (what the sweeps advance). With
k0 = w dx / c0 and the operators U (forward) and V = -U^T (backward):

    beta Hu = -k0 eps_v Ev - (1/k0) Vu mu_s^-1 (Uu Ev - Uv Eu)
    beta Hv =  k0 eps_u Eu - (1/k0) Vv mu_s^-1 (Uu Ev - Uv Eu)
    beta Eu =  k0 mu_v Hv + (1/k0) Uu eps_s^-1 (Vu Hv - Vv Hu)
    beta Ev = -k0 mu_u Hu + (1/k0) Uv eps_s^-1 (Vu Hv - Vv Hu)

so beta^2 is an eigenvalue of Q R acting on (Eu, Ev); Es = (Vu Hv - Vv Hu) / (j k0 eps_s) and
Hs = j (Uu Ev - Uv Eu) / (k0 mu_s). Material values are the per-cell inverse permittivity and
permeability of each component, as the sweeps read them. The largest eigenvalues come from a
shift-invert Arnoldi iteration (scipy.sparse.linalg.eigs) around (k0 n_max)^2.

A transverse axis of exactly two cells whose material does not vary along it is a 2-D cross-
section (modes.py:177-195): derivatives along it are dropped, the mode is solved on one row and
repeated.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import constants


@dataclass
class Mode:
    """One mode: E, H as (3, n1, n2) complex arrays over the cross-section in physical component
    order, the effective index, and the transverse axes (u, v) of the (n1, n2) layout."""

    E: np.ndarray
    H: np.ndarray
    neff: complex
    axes: tuple


def _fwd(n: int):
    import scipy.sparse as sp

    return sp.diags([-np.ones(n), np.ones(n - 1)], [0, 1], shape=(n, n), format="csr")


def _component_values(arr, n1n2, s, uv, default):
    """Per-component material values (1 | 3 components, or a scalar) as flat (3, n) arrays in
    physical component order over the (u, v) plane."""
    if np.isscalar(arr) or np.ndim(arr) == 0:
        return np.full((3, n1n2), float(arr))
    a = np.asarray(arr, np.complex128)
    if a.shape[0] not in (1, 3):
        raise NotImplementedError("mode solving in full-tensor (off-diagonal) media is not supported")
    plane = np.take(a, 0, axis=s + 1)  # (C, *two transverse axes in physical order)
    if uv[0] > uv[1]:
        plane = np.swapaxes(plane, 1, 2)
    plane = plane.reshape(a.shape[0], -1)
    return np.broadcast_to(plane, (3, n1n2)).copy() if a.shape[0] == 1 else plane


def solve_modes(inv_permittivity, inv_permeability, frequency: float, resolution: float, num_modes: int):
    """Modes of the cross-section given by `inv_permittivity` ((1 | 3, *plane) with one singleton
    axis, the propagation axis) and `inv_permeability` (same layout, or a scalar), sorted by the
    real part of the effective index, largest first."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    ie = np.asarray(inv_permittivity)
    if ie.ndim != 4 or sum(1 for d in ie.shape[1:] if d == 1) != 1:
        raise Exception(f"Invalid shape of inv_permittivities: {ie.shape}")
    s = ie.shape[1:].index(1)
    uv = ((s + 1) % 3, (s + 2) % 3)
    n1, n2 = ie.shape[1 + uv[0]], ie.shape[1 + uv[1]]
    n = n1 * n2
    eps = 1.0 / _component_values(ie, n, s, uv, 1.0)
    mu = 1.0 / _component_values(inv_permeability, n, s, uv, 1.0)
    # 2-D cross-sections: an axis of two cells with uniform material along it
    live = []
    for a, na in ((0, n1), (1, n2)):
        flat = na == 2 and all(np.allclose(e.reshape(n1, n2).take(0, axis=a), e.reshape(n1, n2).take(1, axis=a))
                               for e in (*eps, *mu))
        live.append(not flat)
    keep = [slice(None), slice(None)]
    for a, ok in enumerate(live):
        if not ok:
            keep[a] = slice(0, 1)
    m1 = n1 if live[0] else 1
    m2 = n2 if live[1] else 1
    if not all(live):
        eps = np.stack([e.reshape(n1, n2)[tuple(keep)].ravel() for e in eps])
        mu = np.stack([m.reshape(n1, n2)[tuple(keep)].ravel() for m in mu])
    m = m1 * m2
    I1, I2 = sp.identity(m1, format="csr"), sp.identity(m2, format="csr")
    zero = sp.csr_matrix((m, m))
    Uu = sp.kron(_fwd(m1), I2, format="csr") if live[0] else zero
    Uv = sp.kron(I1, _fwd(m2), format="csr") if live[1] else zero
    Vu, Vv = -Uu.T.tocsr(), -Uv.T.tocsr()
    k0 = 2.0 * math.pi * frequency * resolution / constants.c
    D = sp.diags
    cu, cv, cs = uv[0], uv[1], s
    ims = D(1.0 / mu[cs])
    ies = D(1.0 / eps[cs])
    # R: (Eu, Ev) -> beta (Hu, Hv)
    R = sp.bmat([[(1.0 / k0) * (Vu @ ims @ Uv), -k0 * D(eps[cv]) - (1.0 / k0) * (Vu @ ims @ Uu)],
                 [k0 * D(eps[cu]) + (1.0 / k0) * (Vv @ ims @ Uv), -(1.0 / k0) * (Vv @ ims @ Uu)]], format="csc")
    # Q: (Hu, Hv) -> beta (Eu, Ev)
    Q = sp.bmat([[-(1.0 / k0) * (Uu @ ies @ Vv), k0 * D(mu[cv]) + (1.0 / k0) * (Uu @ ies @ Vu)],
                 [-k0 * D(mu[cu]) - (1.0 / k0) * (Uv @ ies @ Vv), (1.0 / k0) * (Uv @ ies @ Vu)]], format="csc")
    A = (Q @ R).tocsc()
    nmax2 = float(np.max(np.real(eps * mu)))
    k = int(min(num_modes, 2 * m - 2))
    if k < 1:
        raise ValueError("cross-section too small for a mode solve")
    sigma = (k0 * k0) * nmax2 * (1.0 + 1e-6)
    if 2 * m <= 400:  # small systems: dense, every eigenpair
        w, vecs = np.linalg.eig(A.toarray())
        order = np.argsort(-np.real(w))[:k]
        w, vecs = w[order], vecs[:, order]
    else:
        w, vecs = sla.eigs(A, k=k, sigma=sigma, which="LM", v0=np.ones(2 * m))  # fixed start: reproducible
    modes = []
    for idx in range(len(w)):
        beta = np.sqrt(complex(w[idx]))
        if beta.real < 0:
            beta = -beta
        Et = vecs[:, idx]
        # a real operator has real eigenvectors for real eigenvalues: remove the arbitrary phase
        j = int(np.argmax(np.abs(Et)))
        Et = Et * (abs(Et[j]) / Et[j])
        if abs(np.imag(w[idx])) < 1e-12 * abs(w[idx]):
            Et = np.real(Et).astype(np.complex128)
        Eu, Ev = Et[:m], Et[m:]
        BH = R @ Et
        Hu, Hv = BH[:m] / beta, BH[m:] / beta
        Es = (Vu @ Hv - Vv @ Hu) / (1j * k0 * eps[cs])
        Hs = 1j * (Uu @ Ev - Uv @ Eu) / (k0 * mu[cs])
        E = np.zeros((3, m), np.complex128)
        H = np.zeros((3, m), np.complex128)
        E[cu], E[cv], E[cs] = Eu, Ev, Es
        H[cu], H[cv], H[cs] = Hu, Hv, Hs
        E = E.reshape(3, m1, m2)
        H = H.reshape(3, m1, m2)
        if not live[0]:
            E, H = np.repeat(E, n1, axis=1), np.repeat(H, n1, axis=1)
        if not live[1]:
            E, H = np.repeat(E, n2, axis=2), np.repeat(H, n2, axis=2)
        modes.append(Mode(E=E, H=H, neff=complex(beta / k0), axes=uv))
    # by effective index; evanescent modes (imaginary neff, all tied at zero) by their beta^2
    modes.sort(key=lambda md: (0, -md.neff.real) if md.neff.real > 1e-6 else (1, -(md.neff**2).real))
    return modes


def polarization_fraction(mode: Mode, pol: str) -> float:
    """Share of |E|^2 along the first (lower-index) transverse axis for "te", along the second
    for "tm" (modes.py:38-66)."""
    a, b = sorted(mode.axes)
    e1 = np.sum(np.abs(mode.E[a]) ** 2)
    e2 = np.sum(np.abs(mode.E[b]) ** 2)
    if pol == "te":
        num = e1
    elif pol == "tm":
        num = e2
    else:
        raise ValueError(f"pol must be 'te' or 'tm', but got {pol}")
    return float(num / (e1 + e2 + 1e-18))


def order_modes(modes, filter_pol):
    """Modes of the requested polarization first (share >= 1/2), each group by decreasing
    effective index (modes.py:69-98)."""
    if filter_pol is None:
        return sorted(modes, key=lambda md: -md.neff.real)
    match = [md for md in modes if polarization_fraction(md, filter_pol) >= 0.5]
    rest = [md for md in modes if polarization_fraction(md, filter_pol) < 0.5]
    return sorted(match, key=lambda md: -md.neff.real) + sorted(rest, key=lambda md: -md.neff.real)


def integrated_power(E, H, axis):
    """|sum 1/2 Re((conj E x H)_axis)| over the plane (metrics.py:160-189)."""
    a1, a2 = (axis + 1) % 3, (axis + 2) % 3
    E = np.asarray(E, np.complex128)
    H = np.asarray(H, np.complex128)
    return abs(float(np.sum(0.5 * np.real(np.conj(E[a1]) * H[a2] - np.conj(E[a2]) * H[a1]))))


def compute_mode(frequency: float, inv_permittivities, inv_permeabilities, resolution: float, direction: str = "+",
                 mode_index: int = 0, filter_pol=None, symmetry=(0, 0)):
    """Mode `mode_index` of the cross-section as (E, H, neff) with E, H complex (3, *plane) in the
    layout of `inv_permittivities` (one singleton axis), normalised to unit Poynting flux along the
    propagation axis (modes.py:101-385). A "-" mode is the forward mode under the reciprocity
    transformation: the longitudinal E and the transverse H change sign (modes.py:683-691)."""
    if direction not in ("+", "-"):
        raise ValueError(f"direction must be '+' or '-', got {direction!r}")
    if tuple(symmetry) != (0, 0):
        raise NotImplementedError("mode solving with a magnetic (PMC) wall at a transverse min edge is not supported")
    ie = np.asarray(inv_permittivities)
    s = ie.shape[1:].index(1)
    modes = solve_modes(ie, inv_permeabilities, frequency, resolution, 2 * (mode_index + 1) + 10)
    md = order_modes(modes, filter_pol)[mode_index]
    E, H = md.E.copy(), md.H.copy()
    if direction == "-":
        E[s] = -E[s]
        for c in md.axes:
            H[c] = -H[c]
    E, H = _to_physical(md, np.expand_dims(E, axis=s + 1), np.expand_dims(H, axis=s + 1), s)
    p = integrated_power(E, H, s)
    f = math.sqrt(p)
    return E / f, H / f, md.neff


def _to_physical(md: Mode, E, H, s):
    """(3, n1, n2) with the singleton inserted at s + 1 -> (3, nx, ny, nz) in physical axis order."""
    u, v = md.axes
    if u < v:
        return E, H
    # the plane was laid out (u, v) with u > v: swap the two transverse axes back
    ax = [1 + a for a in range(3) if a != s]
    return np.swapaxes(E, ax[0], ax[1]), np.swapaxes(H, ax[0], ax[1])
