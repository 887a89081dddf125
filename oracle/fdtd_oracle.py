"""CPU oracle: a NumPy restatement of fdtdx's FDTD time step, used ONLY as a checker.

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module. The product path (``fdtdx_amd``) never
routes through it and fails loudly when its HIP library is missing.

Every function restates one piece of the reference (ymahlau/fdtdx 0.6.2) and cites the
file:line it follows (paths relative to the reference's ``src/fdtdx``). The float32 variant
evaluates every expression in the same order and with the same float32 roundings that JAX's
weak-typed float32 evaluation performs, so that a kernel compiled without FMA contraction can
be compared to it almost bit for bit. Passing ``dtype=np.float64`` gives the FP64 reference run.

Parity pinning (what checks each function; see tests/test_oracle.py):
  * curl_H / curl_E: reference KATs (linear fields give constant curl), integer-exact fields,
    and the adjoint identity <curlH a, b> = <a, curlE b> (zero-padded Yee duality).
  * cpml_profiles: closed forms at the interface (sigma=0 -> a=0), monotonicity, the
    sigma_end formula of PML.py:109, and a physics check (plane-wave reflection << 1 %).
  * step (E/H update, CPML, dipole): closed form of one step from a delta field, energy
    decay in an all-PML box, and the FP64 run agreeing with the FP32 run.
  * detectors: brute-force loops on tiny regions and the pure-tone phasor identity.
  * ade_coefficients: closed forms of dispersion.py:789-857 and the Lorentz susceptibility
    chi(omega) recovered from the discrete recurrence.
  * reverse_step: round trip forward->reverse on the interior (time reversal).
  * adjoint_step: dot-product test <J d, l> = <d, J^T l> against the oracle's own linearised
    step, and central finite differences of the full loss in FP64.
  * complex fields with Bloch phases (Scene.bloch): a discrete Bloch plane wave on a periodic
    axis is an eigenmode whose per-step phase follows the Yee dispersion relation (pins the
    phase's sign on each face), a phase of 2 pi equals the plain periodic run, the run with the
    conjugate phase is the conjugate run, and an interior detector's complex energy and flux equal
    the real-field formulas applied to the real and imaginary parts.
  * loss_factors (conductivity): a uniform field in a periodic box decays by exactly
    ((1-x)/(1+x))^n; a plane wave's attenuation in a conductor matches the continuum constant; a
    matched absorber (sigma_H = eta0^2 sigma_E) barely reflects; the dot-product test and finite
    differences with lossy cells.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------- constants (constants.py:13-22)
C0 = 299792458.0
MU0 = 4e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0**2)
ETA0 = MU0 * C0


def courant_number(courant_factor: float = 0.99) -> float:
    """config.py:165-177: courant_factor / sqrt(3)."""
    return courant_factor / math.sqrt(3)


def time_step_duration(dx: float, courant_factor: float = 0.99) -> float:
    """core/grid.py:495-511 (uniform branch)."""
    return (courant_factor / float(np.sqrt(3.0))) * dx / C0


# ---------------------------------------------------------------- CPML (perfectly_matched_layer.py)
def cpml_profiles(
    L: int,
    side: str,
    dt: float,
    dx: float,
    dtype=np.float32,
    sigma_start: float = 0.0,
    sigma_end: float | None = None,
    sigma_order: float = 3.0,
    kappa_start: float = 1.0,
    kappa_end: float = 1.0,
    kappa_order: float = 3.0,
    alpha_start: float | None = None,
    alpha_end: float = 0.0,
    alpha_order: float = 1.0,
):
    """Per-slab 1-D CPML coefficients (PML.py:69-136, profile 231-275).

    Returns b_E, a_E, inv_kappa_E, b_H, a_H, inv_kappa_H as 1-D arrays of length L, indexed
    along the PML axis in grid order (slab-local index 0 = lowest grid index).
    """
    if alpha_start is None:
        alpha_start = 0.01 * 2 * math.pi * C0 / 1.55e-6 * EPS0  # PML.py:74
    f = dtype
    if sigma_end is None:
        # PML.py:109 evaluated as JAX float32 (jnp.log of a python float is weak float32)
        thickness_m = L * dx
        if f == np.float32:
            log_term = np.float32(np.log(np.float32(1e-6)))
            sigma_end = np.float32(-(sigma_order + 1)) * log_term / np.float32(2 * (ETA0 / 1.0) * thickness_m)
            sigma_end = float(np.float32(sigma_end))
        else:
            sigma_end = -(sigma_order + 1) * math.log(1e-6) / (2 * (ETA0 / 1.0) * thickness_m)
    if side == "-":  # PML.py:255-259
        dE = np.arange(L - 1, -1, -1, dtype=f)
        dH = np.append(np.arange(L - 1.5, -0.5, -1, dtype=f), f(0)).astype(f)
    else:  # PML.py:260-264
        dE = np.insert(np.arange(0.5, L - 0.5, 1, dtype=f), 0, f(0)).astype(f)
        dH = np.arange(0, L, 1, dtype=f)
    norm = f(L)

    def prof(v0, v1, order, d):
        # value_start + (value_end - value_start) * (d / L) ** order   (PML.py:267-268)
        return (f(v0) + f(v1 - v0) * np.power(d / norm, f(order))).astype(f)

    sE, sH = prof(sigma_start, sigma_end, sigma_order, dE), prof(sigma_start, sigma_end, sigma_order, dH)
    kE, kH = prof(kappa_start, kappa_end, kappa_order, dE), prof(kappa_start, kappa_end, kappa_order, dH)
    aE_, aH_ = prof(alpha_start, alpha_end, alpha_order, dE), prof(alpha_start, alpha_end, alpha_order, dH)
    s = f(-dt / EPS0)  # python-float product, then weak-typed to float32 (PML.py:123)

    def coeffs(sig, kap, alp):
        b = (np.expm1(s * (sig / kap + alp)) + f(1)).astype(f)
        with np.errstate(invalid="ignore", divide="ignore"):
            a = (b - f(1.0)) * sig / (sig + alp * kap) / kap
        a = np.nan_to_num(a, nan=0.0).astype(f)
        return b, a, (f(1.0) / kap).astype(f)

    bE, aE, ikE = coeffs(sE, kE, aE_)
    bH, aH, ikH = coeffs(sH, kH, aH_)
    return bE, aE, ikE, bH, aH, ikH


# ---------------------------------------------------------------- temporal profiles
def gaussian_pulse(time, center_frequency: float, spectral_width: float, phase_shift: float = 0.0, dtype=np.float32):
    """GaussianPulseProfile.get_amplitude (profile.py:322-345) with gaussian_envelope (window.py:16-22).

    ``time`` must already be in ``dtype``; every python scalar is cast to ``dtype`` before use,
    as JAX's weak typing does.
    """
    f = dtype
    sigma_t = 1.0 / (2 * math.pi * spectral_width)
    t0 = 6 * sigma_t
    d = (time - f(t0)).astype(f)
    env = np.exp((-(d * d)) / f(2.0 * sigma_t**2)).astype(f)
    phase = (f(2 * math.pi * center_frequency) * time + f(phase_shift)).astype(f)
    return (env * np.cos(phase).astype(f)).astype(f)


def single_frequency(time, period: float, phase_shift: float = 0.0, profile_phase: float = math.pi,
                     num_startup_periods: int = 4, dtype=np.float32):
    """SingleFrequencyProfile.get_amplitude (profile.py:263-275)."""
    f = dtype
    tp = (f(2 * math.pi) * time / f(period) + f(phase_shift) + f(profile_phase)).astype(f)
    raw = np.cos(tp).astype(f)
    factor = np.clip(time / f(num_startup_periods * period), f(0.0), f(1.0)).astype(f)
    return (factor * raw).astype(f)


# ---------------------------------------------------------------- scene description
@dataclass
class Slab:
    """One CPML object: axis, side ('-' min / '+' max), thickness L and its coefficients."""

    axis: int
    side: str
    L: int
    lo: int = 0
    hi: int = 0
    coeffs: tuple = ()

    def slc(self, shape):
        s = [slice(None)] * 3
        s[self.axis] = slice(self.lo, self.hi)
        return tuple(s)

    def interface_slc(self, shape):
        """boundary.py:134-144: innermost layer of the slab."""
        s = [slice(None)] * 3
        if self.side == "+":
            s[self.axis] = slice(self.lo, self.lo + 1)
        else:
            s[self.axis] = slice(self.hi - 1, self.hi)
        return tuple(s)

    def bcast(self, v):
        shp = [1, 1, 1]
        shp[self.axis] = v.shape[0]
        return v.reshape(shp)


@dataclass
class Scene:
    shape: tuple
    dx: float
    dtype: type = np.float32
    courant_factor: float = 0.99
    inv_eps: np.ndarray | None = None  # (C, Nx, Ny, Nz)
    inv_mu: float | np.ndarray = 1.0
    slabs: list = field(default_factory=list)
    # sources: dicts with kind 'dipole' or 'plane'
    sources: list = field(default_factory=list)
    detectors: list = field(default_factory=list)
    ade: tuple | None = None  # (c1, c2, c3) each (P, C, Nx, Ny, Nz)
    # PEC / PMC faces: ("pec" | "pmc", axis, index of the one-cell layer) in object order
    walls: list = field(default_factory=list)
    # axes with periodic boundaries (wrap padding, update.py:92-137)
    periodic: tuple = (False, False, False)
    # axes with an electric symmetry plane at the min face: detector stencils read the mirror of
    # the interior below it instead of the zero halo (pad_fields_with_symmetry_mirror,
    # update.py:139-190)
    mirror: tuple = (False, False, False)
    # conductivities (C, Nx, Ny, Nz), already scaled by c0*dt/courant (initialization.py:786,
    # 892-930); None where no object conducts
    sigma_E: np.ndarray | None = None
    sigma_H: np.ndarray | None = None
    # complex-valued fields (initialization.py:580-594) and, per periodic axis, the Bloch phase
    # angles (theta_min, theta_max) = (k L) of its min and max boundary (bloch.py:140-153), or None
    complex_fields: bool = False
    bloch: tuple = (None, None, None)

    @property
    def c(self):
        return courant_number(self.courant_factor)

    @property
    def field_dtype(self):
        if not self.complex_fields:
            return self.dtype
        return np.complex64 if self.dtype == np.float32 else np.complex128

    @property
    def wrap(self):
        """Per axis: False (zero halo), True (wrap), or for a Bloch axis of a complex-field scene the
        pair (exp(i theta_min), exp(i theta_max)): the halo below the min face is the far face times
        the conjugate of the first, the halo above the max face the near face times the second
        (apply_pad_correction, bloch.py:61-97)."""
        out = []
        for a in range(3):
            b = self.bloch[a]
            if self.periodic[a] and self.complex_fields and b is not None:
                ct = self.field_dtype
                out.append((ct(np.exp(1j * b[0])), ct(np.exp(1j * b[1]))))
            else:
                out.append(bool(self.periodic[a]))
        return tuple(out)

    @property
    def dt(self):
        return time_step_duration(self.dx, self.courant_factor)


def make_slabs(shape, thickness, dt, dx, dtype=np.float32, **pml_params):
    """boundary_objects_from_config order (objects/boundaries/initialization.py:554-663):
    min_x, max_x, min_y, max_y, min_z, max_z; each slab spans the full other two axes."""
    if isinstance(thickness, int):
        thickness = [thickness] * 6
    slabs = []
    for idx, (axis, side) in enumerate([(0, "-"), (0, "+"), (1, "-"), (1, "+"), (2, "-"), (2, "+")]):
        L = thickness[idx]
        if L <= 0:
            continue
        lo, hi = (0, L) if side == "-" else (shape[axis] - L, shape[axis])
        slabs.append(Slab(axis, side, L, lo, hi, cpml_profiles(L, side, dt, dx, dtype, **pml_params)))
    return slabs


def dipole_table(T, dt, c, center_frequency, spectral_width, inv_eps_local, amplitude=1.0,
                 static_amplitude_factor=1.0, dtype=np.float32, half_step=False, profile="gaussian",
                 period=None, phase_shift=0.0):
    """Per-step increment added to E[p, cell] by PointDipoleSource.update_E (dipole.py:195-234).

    time = time_step * dt evaluated in dtype (int32 * weak float); scale = c*amp*saf (python
    float) * amplitude(t); injection = scale * inv_eps_local; E += -1.0 * injection.
    """
    f = dtype
    n = np.arange(T, dtype=f)
    if half_step:
        n = (n + f(0.5)).astype(f)
    time = (n * f(dt)).astype(f)
    if profile == "gaussian":
        amp = gaussian_pulse(time, center_frequency, spectral_width, phase_shift, dtype=f)
    else:
        amp = single_frequency(time, period, phase_shift, dtype=f)
    scale = (f(c * amplitude * static_amplitude_factor) * amp).astype(f)
    inj = (scale * f(inv_eps_local)).astype(f)
    return (f(-1.0) * inj).astype(f)


# ---------------------------------------------------------------- stencils (core/physics/curl.py)
def _face(a, axis, idx):
    s = [slice(None)] * a.ndim
    s[axis] = idx
    return tuple(s)


def _prev(a, axis, wrap=False):
    """Value at index-1 along axis: a zero halo (pad_fields, core/misc.py:615-641), or the far
    side of the domain on a periodic axis (wrap padding, update.py:92-137), times the conjugate
    min-face Bloch phase on a Bloch axis (bloch.py:61-97)."""
    if wrap:  # True, or a Bloch pair
        out = np.roll(a, 1, axis)
        if isinstance(wrap, tuple):
            f0 = _face(out, axis, 0)
            out[f0] = out[f0] * np.conj(wrap[0])
        return out
    out = np.zeros_like(a)
    s_dst = [slice(None)] * 3
    s_src = [slice(None)] * 3
    s_dst[axis] = slice(1, None)
    s_src[axis] = slice(0, -1)
    out[tuple(s_dst)] = a[tuple(s_src)]
    return out


def _next(a, axis, wrap=False):
    """Value at index+1 along axis (the max-face Bloch phase on a Bloch axis)."""
    if wrap:  # True, or a Bloch pair
        out = np.roll(a, -1, axis)
        if isinstance(wrap, tuple):
            f1 = _face(out, axis, -1)
            out[f1] = out[f1] * wrap[1]
        return out
    out = np.zeros_like(a)
    s_dst = [slice(None)] * 3
    s_src = [slice(None)] * 3
    s_dst[axis] = slice(0, -1)
    s_src[axis] = slice(1, None)
    out[tuple(s_dst)] = a[tuple(s_src)]
    return out


NOWRAP = (False, False, False)
_prev_w = _prev


def curl_H_raw(H, wrap=NOWRAP):
    """curl.py:352-366: backward differences; returns (curl list, derivative dict)."""
    Hx, Hy, Hz = H[0], H[1], H[2]
    P = lambda a, ax: _prev(a, ax, wrap[ax])
    d = {
        "dyHz": Hz - P(Hz, 1), "dzHy": Hy - P(Hy, 2), "dzHx": Hx - P(Hx, 2),
        "dxHz": Hz - P(Hz, 0), "dxHy": Hy - P(Hy, 0), "dyHx": Hx - P(Hx, 1),
    }
    curl = [d["dyHz"] - d["dzHy"], d["dzHx"] - d["dxHz"], d["dxHy"] - d["dyHx"]]
    return curl, d


def curl_E_raw(E, wrap=NOWRAP):
    """curl.py:265-279: forward differences."""
    Ex, Ey, Ez = E[0], E[1], E[2]
    N = lambda a, ax: _next(a, ax, wrap[ax])
    d = {
        "dyEz": N(Ez, 1) - Ez, "dzEy": N(Ey, 2) - Ey, "dzEx": N(Ex, 2) - Ex,
        "dxEz": N(Ez, 0) - Ez, "dxEy": N(Ey, 0) - Ey, "dyEx": N(Ex, 1) - Ex,
    }
    curl = [d["dyEz"] - d["dzEy"], d["dzEx"] - d["dxEz"], d["dxEy"] - d["dyEx"]]
    return curl, d


# ---------------------------------------------------------------- full-tensor materials
def avg_E(F, comp, loc, wrap=NOWRAP):
    """avg_anisotropic_E_component (fdtd/misc.py:132-171, uniform grid): component `comp` of a
    field on E-type Yee positions averaged onto component `loc`'s position,
    (F[p] + F[p + e_loc] + F[p - e_comp] + F[p + e_loc - e_comp]) / 4, zero (or wrapped) beyond
    the domain."""
    s = F[comp]
    N = lambda a, ax: _next(a, ax, wrap[ax])
    P = lambda a, ax: _prev(a, ax, wrap[ax])
    return (((s + N(s, loc)) + P(s, comp)) + P(N(s, loc), comp)) / s.dtype.type(4)


def avg_H(F, comp, loc, wrap=NOWRAP):
    """avg_anisotropic_H_component (fdtd/misc.py:174-213, uniform grid):
    (F[p] + F[p - e_loc] + F[p + e_comp] + F[p - e_loc + e_comp]) / 4."""
    s = F[comp]
    N = lambda a, ax: _next(a, ax, wrap[ax])
    P = lambda a, ax: _prev(a, ax, wrap[ax])
    return (((s + P(s, loc)) + N(s, comp)) + N(P(s, loc), comp)) / s.dtype.type(4)


def avg_E_T(g, comp, loc, wrap=NOWRAP):
    """Transpose of avg_E: a cotangent g on `loc` positions back onto `comp` positions."""
    N = lambda a, ax: _next(a, ax, wrap[ax])
    P = lambda a, ax: _prev(a, ax, wrap[ax])
    return (((g + P(g, loc)) + N(g, comp)) + N(P(g, loc), comp)) / g.dtype.type(4)


def avg_H_T(g, comp, loc, wrap=NOWRAP):
    """Transpose of avg_H."""
    N = lambda a, ax: _next(a, ax, wrap[ax])
    P = lambda a, ax: _prev(a, ax, wrap[ax])
    return (((g + N(g, loc)) + P(g, comp)) + P(N(g, loc), comp)) / g.dtype.type(4)


def is_tensor(m):
    """True for a 9-component (full-tensor) material array; scalars and 1- or 3-component arrays
    are not."""
    return np.ndim(m) > 0 and np.shape(m)[0] == 9


def tensor_term(B, K, kind, wrap=NOWRAP):
    """Row i of B @ K with the off-diagonal K_j averaged onto component i's position
    (update.py:470-488, 796-820 with A = I): ((B_i0 T_0 + B_i1 T_1) + B_i2 T_2), T_i = K_i,
    T_j = avg(K_j -> i). B is (9, ...) row-major."""
    avg = avg_E if kind == "E" else avg_H
    rows = []
    for i in range(3):
        t = [B[3 * i + j] * (K[j] if j == i else avg(K, j, i, wrap)) for j in range(3)]
        rows.append((t[0] + t[1]) + t[2])
    return np.stack(rows)


def expand3x3(m, dtype):
    """expand_to_3x3 (core/misc.py:557-620) as a (9, ...) row-major array: a scalar or one
    component on the diagonal, three components as the diagonal, nine as they are."""
    m = np.asarray(m, dtype)
    if m.ndim == 0:
        m = m.reshape(1, 1, 1, 1)
    if m.shape[0] == 9:
        return m
    out = np.zeros((9, *m.shape[1:]), dtype)
    for i in range(3):
        out[4 * i] = m[0 if m.shape[0] == 1 else i]
    return out


def aniso_matrices(inv9, sig9, c, eta_factor, reverse=False):
    """compute_anisotropic_update_matrices(_reverse) (fdtd/misc.py:69-129): per cell,
    M1 = I + (c eta / 2) inv sigma, M2 = I - (c eta / 2) inv sigma, and (A, B) = (M1^-1 M2,
    c M1^-1 inv) forward or (M2^-1 M1, c M2^-1 inv) reverse; inputs and outputs (9, ...)
    row-major, evaluated in the inputs' precision."""
    f = inv9.dtype.type
    sp = inv9.shape[1:]
    inv = np.moveaxis(inv9.reshape(3, 3, *sp), (0, 1), (-2, -1))
    sig = np.moveaxis(sig9.reshape(3, 3, *sp), (0, 1), (-2, -1))
    eye = np.eye(3, dtype=f)
    fac = f(c * eta_factor / 2) * (inv @ sig)
    M1, M2 = eye + fac, eye - fac
    L, R = (M2, M1) if reverse else (M1, M2)
    A = np.linalg.solve(L, R)
    B = f(c) * np.linalg.solve(L, inv)
    back = lambda X: np.moveaxis(X.astype(f), (-2, -1), (0, 1)).reshape(9, *sp)
    return back(A), back(B)


def aniso_update(A, B, F, K, kind, sign, wrap=NOWRAP):
    """Row i of A avg(F) +- B avg(K) (update.py:470-488, 611-680, 796-820, 960-1000):
    ((A_i0 X_0 + A_i1 X_1) + A_i2 X_2) +- ((B_i0 Y_0 + B_i1 Y_1) + B_i2 Y_2), X_i = F_i and
    X_j = avg(F_j -> i), Y likewise from K."""
    avg = avg_E if kind == "E" else avg_H
    rows = []
    for i in range(3):
        x = [A[3 * i + j] * (F[j] if j == i else avg(F, j, i, wrap)) for j in range(3)]
        y = [B[3 * i + j] * (K[j] if j == i else avg(K, j, i, wrap)) for j in range(3)]
        a = (x[0] + x[1]) + x[2]
        b = (y[0] + y[1]) + y[2]
        rows.append(a + b if sign > 0 else a - b)
    return np.stack(rows)


def _avg_E1(s, comp, loc, wrap):
    """avg_E of one scalar field s sitting on component `comp`'s positions."""
    N = lambda a, ax: _next(a, ax, wrap[ax])
    P = lambda a, ax: _prev(a, ax, wrap[ax])
    return (((s + N(s, loc)) + P(s, comp)) + P(N(s, loc), comp)) / s.dtype.type(4)


def tensor_ade(st, scene, K):
    """The ADE block of the full-tensor E update (update.py:361-423): P_hat = c1 P^n + c2 P^{n-1}
    + coupling, the polarisation current delta = sum_p (P^n - P_hat) folded into the curl as
    K + delta / c. With 9-component couplings (oriented poles) row i is c_ii E_i + sum_{j != i}
    0.5 (c_ij avg(m_ij E_j) + [c_ij != 0] avg(c_ij E_j)), m = [c != 0], the means taken over the
    zero (or wrapped) halo of E. Returns the new K and advances st's polarisation."""
    f = scene.dtype
    c1, c2, c3 = scene.ade
    E = st.E
    wrap = scene.periodic
    if c3.shape[1] == 9:
        rows = []
        for i in range(3):
            row = c3[:, 4 * i] * E[i]
            for j in range(3):
                if j == i:
                    continue
                cij = c3[:, 3 * i + j]
                mask = (cij != 0).astype(f)
                a1 = np.stack([_avg_E1(mask[p] * E[j], j, i, wrap) for p in range(cij.shape[0])])
                a2 = np.stack([_avg_E1(cij[p] * E[j], j, i, wrap) for p in range(cij.shape[0])])
                row = row + f(0.5) * (cij * a1 + (cij != 0).astype(f) * a2)
            rows.append(row)
        coupling = np.stack(rows, axis=1)
    else:
        coupling = c3 * E
    P_hat = c1 * st.P_curr + c2 * st.P_prev + coupling
    delta = np.sum(st.P_curr - P_hat, axis=0)
    st.P_prev = st.P_curr
    st.P_curr = P_hat.astype(scene.field_dtype)
    return K + delta / f(scene.c)


def aniso_inv_gradient(mu, E, K, inv9, sig9, c, eta, wrap):
    """dL/d inv (9 components) of E' = A avg(E) + B avg(K) for the cotangent mu on E': with
    G_A = mu_r X_rq(E), G_B = mu_r Y_rq(K) (the averaged operands), M1 = I + f inv sigma,
    f = c eta / 2: dA = -M1^-1 dM1 (A + I), dB = -M1^-1 dM1 B + c M1^-1 dinv, dM1 = f dinv sigma,
    so dL/dinv = c M1^-T G_B + f G_M1 sigma^T, G_M1 = -M1^-T (G_A (A + I)^T + G_B B^T), per cell
    in float64."""
    sp = inv9.shape[1:]
    X = np.stack([mu[r] * (E[q] if q == r else avg_E(E, q, r, wrap)) for r in range(3) for q in range(3)])
    Y = np.stack([mu[r] * (K[q] if q == r else avg_E(K, q, r, wrap)) for r in range(3) for q in range(3)])
    mat = lambda a: np.moveaxis(np.asarray(a, np.float64).reshape(3, 3, *sp), (0, 1), (-2, -1))
    inv, sig, GA, GB = mat(inv9), mat(sig9), mat(X), mat(Y)
    fct = c * eta / 2
    eye = np.eye(3)
    M1 = eye + fct * (inv @ sig)
    M1i = np.linalg.inv(M1)
    A = M1i @ (eye - fct * (inv @ sig))
    B = c * (M1i @ inv)
    tr = lambda a: np.swapaxes(a, -1, -2)
    GM1 = -tr(M1i) @ (GA @ tr(A + eye) + GB @ tr(B))
    d = c * (tr(M1i) @ GB) + fct * (GM1 @ tr(sig))
    return np.moveaxis(d, (-2, -1), (0, 1)).reshape(9, *sp).astype(inv9.dtype)


def tensor_sigma(scene, which):
    """The conductivity of a half-step when it takes the full-tensor update: a 9-component
    sigma, or any sigma beside a full-tensor inverse material (update.py:296-297, 708-709)."""
    sig = scene.sigma_E if which == "E" else scene.sigma_H
    inv = scene.inv_eps if which == "E" else scene.inv_mu
    if sig is None:
        return None
    return sig if (is_tensor(sig) or is_tensor(inv)) else None


def tensor_term_T(B, g, kind, wrap=NOWRAP):
    """Transpose of tensor_term in K: component j = sum_i avg^T(B_ij g_i) (B_jj g_j for i = j)."""
    avgT = avg_E_T if kind == "E" else avg_H_T
    out = np.zeros_like(g)
    for j in range(3):
        for i in range(3):
            v = B[3 * i + j] * g[i]
            out[j] = out[j] + (v if i == j else avgT(v, j, i, wrap))
    return out


_PML_DERIV = {  # curl.py:289-294 and 375-380: (d_field_1, d_field_2) per PML axis
    "H": {0: ("dxHz", "dxHy"), 1: ("dyHx", "dyHz"), 2: ("dzHy", "dzHx")},
    "E": {0: ("dxEz", "dxEy"), 1: ("dyEx", "dyEz"), 2: ("dzEy", "dzEx")},
}


def apply_cpml(curl, d, psi, slabs, which, simulate_boundaries, shape):
    """curl.py:284-308 / 371-394 with step_cpml (PML.py:138-190).

    which = 'E' (curl of H, E coefficients) or 'H' (curl of E, H coefficients). psi is a list of
    (psi1, psi2) per slab in slab order; returns the new psi list. Corrections are scattered in
    slab order: curl[(a+1)%3] += -corr1, curl[(a+2)%3] += corr2.
    """
    new_psi = []
    key = "H" if which == "E" else "E"
    for s, (p1, p2) in zip(slabs, psi):
        bE, aE, ikE, bH, aH, ikH = s.coeffs
        b, a, ik = (bE, aE, ikE) if which == "E" else (bH, aH, ikH)
        b, a, ik = s.bcast(b), s.bcast(a), s.bcast(ik)
        sl = s.slc(shape)
        n1, n2 = _PML_DERIV[key][s.axis]
        d1, d2 = d[n1][sl], d[n2][sl]
        if simulate_boundaries:
            p1n = b * p1 + a * d1
            p2n = b * p2 + a * d2
        else:
            p1n, p2n = p1, p2
        if np.all(ik == 1):  # kappa_start == kappa_end == 1 (PML.py:184-188)
            c1, c2 = p1n, p2n
        else:
            c1 = (ik - 1.0).astype(p1.dtype) * d1 + p1n
            c2 = (ik - 1.0).astype(p1.dtype) * d2 + p2n
        i, j = (s.axis + 1) % 3, (s.axis + 2) % 3
        curl[i][sl] = curl[i][sl] + (-c1)
        curl[j][sl] = curl[j][sl] + c2
        new_psi.append((p1n, p2n))
    return new_psi


# ---------------------------------------------------------------- state
@dataclass
class State:
    E: np.ndarray
    H: np.ndarray
    psiE: list
    psiH: list
    P_curr: np.ndarray | None = None
    P_prev: np.ndarray | None = None


def init_state(scene: Scene) -> State:
    f = scene.field_dtype
    shp = (3, *scene.shape)
    psi = lambda: [(np.zeros(_slab_shape(s, scene.shape), f), np.zeros(_slab_shape(s, scene.shape), f)) for s in scene.slabs]
    P = None
    if scene.ade is not None:
        P = np.zeros((scene.ade[0].shape[0], *shp), f)
    return State(np.zeros(shp, f), np.zeros(shp, f), psi(), psi(), P, None if P is None else P.copy())


def _slab_shape(s, shape):
    shp = list(shape)
    shp[s.axis] = s.L
    return tuple(shp)


# ---------------------------------------------------------------- sources
def _apply_sources(field_arr, scene, n, which, inverse=False):
    """Source hooks in update_E/update_H (update.py:494-519, 824-849).

    dipole: E[p, cell] += table[n] (table already holds -scale*inv_eps, dipole.py:225-228).
    plane (TFSF, tfsf.py:193-409): per face; E[a] += s*((Hinc_b*ampH_b[n])*c)*inv_eps[a],
    E[b] += -s*((Hinc_a*ampH_a[n])*c)*inv_eps[b]; H[b] += s*((Einc_a*ampE_a[n])*c)*inv_mu,
    H[a] += -s*((Einc_b*ampE_b[n])*c)*inv_mu.
    """
    f = scene.dtype
    for src in scene.sources:
        if src["field"] != which:
            continue
        if src["kind"] == "dipole":
            i, j, k = src["cell"]
            v = src["table"][n]
            if inverse:
                field_arr[src["component"], i, j, k] -= v
            else:
                field_arr[src["component"], i, j, k] += v
        elif src["kind"] == "plane":
            sl = src["slc"]
            sgn = src["sign"] * (-1 if inverse else 1)
            for comp, src_comp, tsign in src["terms"]:
                prof = src["profile"][src_comp]
                amp = src["amp"][src_comp][n]
                term = ((prof * f(amp)).astype(f) * f(scene.c)).astype(f) * src["mult"][comp]
                field_arr[(comp, *sl)] = field_arr[(comp, *sl)] + f(sgn * tsign) * term.astype(f)


# ---------------------------------------------------------------- the forward step
def _tangential(axis):
    return [c for c in range(3) if c != axis]


def apply_walls(F, scene: Scene, kind: str):
    """apply_boundary_post_E/H_update (update.py:48-89): a PEC face zeroes the tangential E
    components on its layer (pec.py:70-77), a PMC face the tangential H components (pmc.py:62-75)."""
    for w_kind, axis, index in scene.walls:
        if w_kind != kind:
            continue
        sl = [slice(None)] * 3
        sl[axis] = slice(index, index + 1)
        for c in _tangential(axis):
            F[(c, *sl)] = 0


def loss_factors(scene: Scene, which: str):
    """The lossy update's x = c*sigma_E*eta0*inv_eps/2 (update.py:300-305) or
    x = c*sigma_H/eta0*inv_mu/2 (update.py:743-746), in that float order; the update multiplies
    the old field by 1 - x and divides the new one by 1 + x (Schneider 3.12). None without
    conductivity."""
    f = scene.dtype
    c = f(scene.c)
    if which == "E":
        if scene.sigma_E is None:
            return None
        return ((c * scene.sigma_E) * f(ETA0)) * scene.inv_eps / f(2)
    if scene.sigma_H is None:
        return None
    im = scene.inv_mu if not isinstance(scene.inv_mu, float) else f(scene.inv_mu)
    return ((c * scene.sigma_H) / f(ETA0)) * im / f(2)


def update_E(st: State, scene: Scene, n: int, simulate_boundaries=True):
    """update.py:256-354 (isotropic/diagonal branch) with conductivity, ADE and sources."""
    f = scene.dtype
    F = scene.field_dtype
    c = f(scene.c)
    curl, d = curl_H_raw(st.H, scene.wrap)
    st.psiE = apply_cpml(curl, d, st.psiE, scene.slabs, "E", simulate_boundaries, scene.shape)
    K = np.stack(curl)
    ie = scene.inv_eps
    x = loss_factors(scene, "E")
    sig = tensor_sigma(scene, "E")
    oriented = scene.ade is not None and scene.ade[2].shape[1] == 9
    if sig is not None:  # update.py:356-492 with conductivity: A avg(E) + B avg(K)
        A, B = aniso_matrices(expand3x3(ie, f), expand3x3(sig, f), scene.c, ETA0)
        if scene.ade is not None:
            K = tensor_ade(st, scene, K)
        E_new = aniso_update(A, B, st.E, K, "E", +1, scene.periodic)
        _apply_sources(E_new, scene, n, "E")
        E_new = E_new.astype(F)
        apply_walls(E_new, scene, "pec")
        st.E = E_new
        return K
    if is_tensor(ie) or oriented:  # update.py:356-492 without conductivity: E + B avg(K), B = c inv_eps
        if x is not None:
            raise NotImplementedError("oracle: full-tensor permittivity with diagonal-path conductivity")
        if scene.ade is not None:
            K = tensor_ade(st, scene, K)
        E_new = st.E + tensor_term(c * expand3x3(ie, f), K, "E", scene.periodic)
        _apply_sources(E_new, scene, n, "E")
        E_new = E_new.astype(F)
        apply_walls(E_new, scene, "pec")
        st.E = E_new
        return K
    E_new = (st.E if x is None else (f(1) - x) * st.E) + (c * K) * ie
    if scene.ade is not None:  # update.py:316-347
        c1, c2, c3 = scene.ade
        P_hat = c1 * st.P_curr + c2 * st.P_prev + c3 * st.E
        delta = np.sum(st.P_curr - P_hat, axis=0)
        E_new = E_new + ie * delta
        st.P_prev = st.P_curr
        st.P_curr = P_hat.astype(F)
    if x is not None:  # update.py:345-354
        E_new = E_new / (f(1) + x)
    _apply_sources(E_new, scene, n, "E")
    E_new = E_new.astype(F)
    apply_walls(E_new, scene, "pec")
    st.E = E_new
    return K


def update_H(st: State, scene: Scene, n: int, simulate_boundaries=True):
    """update.py:689-853 (isotropic/diagonal branch)."""
    f = scene.dtype
    F = scene.field_dtype
    c = f(scene.c)
    curl, d = curl_E_raw(st.E, scene.wrap)
    st.psiH = apply_cpml(curl, d, st.psiH, scene.slabs, "H", simulate_boundaries, scene.shape)
    K = np.stack(curl)
    im = scene.inv_mu if not isinstance(scene.inv_mu, float) else f(scene.inv_mu)
    x = loss_factors(scene, "H")
    sig = tensor_sigma(scene, "H")
    if sig is not None:  # update.py:752-822 with conductivity: A avg(H) - B avg(K)
        A, B = aniso_matrices(expand3x3(im, f), expand3x3(sig, f), scene.c, 1.0 / ETA0)
        H_new = aniso_update(A, B, st.H, K, "H", -1, scene.periodic)
        _apply_sources(H_new, scene, n, "H")
        H_new = H_new.astype(F)
        apply_walls(H_new, scene, "pmc")
        st.H = H_new
        return K
    if is_tensor(im):  # update.py:752-822 without conductivity: H - B avg(K), B = c inv_mu
        if x is not None:
            raise NotImplementedError("oracle: full-tensor permeability with conductivity")
        H_new = st.H - tensor_term(c * im, K, "H", scene.periodic)
        _apply_sources(H_new, scene, n, "H")
        H_new = H_new.astype(F)
        apply_walls(H_new, scene, "pmc")
        st.H = H_new
        return K
    H_new = (st.H if x is None else (f(1) - x) * st.H) - (c * K) * im
    if x is not None:  # update.py:748-750
        H_new = H_new / (f(1) + x)
    _apply_sources(H_new, scene, n, "H")
    H_new = H_new.astype(F)
    apply_walls(H_new, scene, "pmc")
    st.H = H_new
    return K


# ---------------------------------------------------------------- detectors
def interpolate_fields(E, Hc, wrap=NOWRAP):
    """curl.py:86-224 on the whole padded domain (zero halo, wrapped on periodic axes);
    Hc = (H_prev + H)/2."""
    f = E.dtype.type
    h = f(0.5)
    Ex, Ey, Ez = E
    Hx, Hy, Hz = Hc
    _prev = lambda a, ax: _prev_w(a, ax, wrap[ax])
    up = lambda a: _next(a, 2, wrap[2])
    ex_lo = h * (Ex + _prev(Ex, 0))
    ex_hi = h * (up(Ex) + _prev(up(Ex), 0))
    ex = (ex_lo + ex_hi) / f(2.0)
    ey_lo = h * (Ey + _prev(Ey, 1))
    ey_hi = h * (up(Ey) + _prev(up(Ey), 1))
    ey = (ey_lo + ey_hi) / f(2.0)
    hx = h * (Hx + _prev(Hx, 1))
    hy = h * (Hy + _prev(Hy, 0))
    lo_x = h * (Hz + _prev(Hz, 0))
    lo_xy = h * (lo_x + h * (_prev(Hz, 1) + _prev(_prev(Hz, 1), 0)))
    Hzu = up(Hz)
    hi_x = h * (Hzu + _prev(Hzu, 0))
    hi_xy = h * (hi_x + h * (_prev(Hzu, 1) + _prev(_prev(Hzu, 1), 0)))
    hz = (lo_xy + hi_xy) / f(2.0)
    return np.stack([ex, ey, Ez]), np.stack([hx, hy, hz])


COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def mirror_source(field_type, c, a):
    """pad_fields_with_symmetry_mirror (update.py:181-190) with core/physics/symmetry.py:18-108 for
    an electric (PEC) plane normal to axis a: the interior row the halo below the plane mirrors and
    the sign. Components sampled on the plane (tangential E, normal H) are odd and pair with row 1;
    the others are even and pair with row 0."""
    on_plane = (c != a) if field_type == "E" else (c == a)
    return (1, -1) if on_plane else (0, 1)


def _block(F, sl, wrap=NOWRAP, mirror=NOWRAP, field_type="E"):
    """F restricted to the box sl grown by one cell on every side: zero beyond the domain, or the
    far side of the domain on a periodic axis, or (below the min face of a mirror axis) the
    parity-signed mirror of the interior, axes taken in order."""
    shp = F.shape[1:]
    idx = []
    for a, s in enumerate(sl):
        r = np.arange(s.start - 1, s.stop + 1)
        idx.append(np.mod(r, shp[a]) if wrap[a] else r)
    out = np.zeros((F.shape[0], *[len(r) for r in idx]), F.dtype)
    valid = [((r >= 0) & (r < shp[a])) for a, r in enumerate(idx)]
    ii = np.ix_(*[r[v] for r, v in zip(idx, valid)])
    dst = np.ix_(*[np.nonzero(v)[0] for v in valid])
    for c in range(F.shape[0]):
        out[c][dst] = F[c][ii]
    for a, r in enumerate(idx):  # Bloch axes: the wrapped rows carry the face's phase
        if not isinstance(wrap[a], tuple):
            continue
        raw = np.arange(sl[a].start - 1, sl[a].stop + 1)
        ph = np.ones(len(raw), out.dtype)
        ph[raw < 0] = np.conj(wrap[a][0])
        ph[raw >= shp[a]] = wrap[a][1]
        shape = [1, 1, 1, 1]
        shape[a + 1] = len(raw)
        out = out * ph.reshape(shape)
    for a, s in enumerate(sl):
        if not mirror[a] or s.start != 0:
            continue
        for c in range(F.shape[0]):
            row, sign = mirror_source(field_type, c, a)
            tgt = [slice(None)] * 3
            src = [slice(None)] * 3
            tgt[a] = slice(0, 1)
            src[a] = slice(row + 1, row + 2)
            out[c][tuple(tgt)] = F.dtype.type(sign) * out[c][tuple(src)]
    return out


def mirror_pad(F, mirror, field_type):
    """F with one extra row before the min face of every mirror axis holding its parity-signed
    mirror (axes in order, so a corner takes both signs)."""
    for a in range(3):
        if not mirror[a]:
            continue
        rows = []
        for c in range(3):
            row, sign = mirror_source(field_type, c, a)
            rows.append(F.dtype.type(sign) * np.take(F[c], [row], axis=a))
        F = np.concatenate([np.stack(rows), F], axis=a + 1)
    return F


def mirror_pad_T(G, mirror, field_type):
    """Transpose of mirror_pad: each pad row folds back onto its mirror row with its sign (axes in
    reverse order), and the pad rows are dropped."""
    for a in reversed(range(3)):
        if not mirror[a]:
            continue
        G = G.copy()
        for c in range(3):
            row, sign = mirror_source(field_type, c, a)
            pad = np.take(G[c], [0], axis=a)
            idx = [slice(None)] * 3
            idx[a] = slice(row + 1, row + 2)
            G[c][tuple(idx)] += G.dtype.type(sign) * pad
        G = np.take(G, np.arange(1, G.shape[a + 1]), axis=a + 1)
    return G


def _abs2(F):
    """|F|^2 componentwise (jnp.square(jnp.abs(F)), metrics.py:56-66): F * F for real fields,
    re^2 + im^2 for complex ones."""
    if np.iscomplexobj(F):
        return F.real * F.real + F.imag * F.imag
    return F * F


def _flux(E, H, b1, b2):
    """Re(E x conj H) along the axis normal to (b1, b2): E_b1 H_b2 - E_b2 H_b1 for real fields."""
    if np.iscomplexobj(E):
        return (E[b1].real * H[b2].real + E[b1].imag * H[b2].imag) - (E[b2].real * H[b1].real + E[b2].imag * H[b1].imag)
    return E[b1] * H[b2] - E[b2] * H[b1]


def detector_update(det, state_arr, n, E, H_prev, H, scene):
    """update_detector_states (update.py:1040-1137) for one detector and its update():
    energy.py:90-143, field.py:31-63, poynting_flux.py:171-195, phasor.py:190-235."""
    f = scene.dtype
    idx = det["time_to_idx"][n]
    if idx < 0:
        return state_arr
    sl = det["slc"]
    if det.get("exact_interpolation", True):
        # interior detectors interpolate a haloed block (update.py:1096-1099); zero halo outside
        Eb = _block(E, sl, scene.wrap, scene.mirror, "E")
        Hb = (_block(H_prev, sl, scene.wrap, scene.mirror, "H")
              + _block(H, sl, scene.wrap, scene.mirror, "H")) / f(2)
        Ei, Hi = interpolate_fields(Eb, Hb)
        inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
        Er, Hr = Ei[inner], Hi[inner]
    else:
        Er, Hr = E[(slice(None), *sl)], H[(slice(None), *sl)]
    kind = det["kind"]
    if kind in ("energy", "energy_mean") and (is_tensor(scene.inv_eps) or is_tensor(scene.inv_mu)):
        raise NotImplementedError("oracle: energy detectors in full-tensor media")
    if kind == "energy_mean":  # energy.py:107-114: the per-cell energy averaged along one axis
        ie = scene.inv_eps[(slice(None), *sl)]
        eE = np.sum((f(0.5) * (f(1) / ie)) * _abs2(Er), axis=0)
        im = scene.inv_mu
        if isinstance(im, float):
            eH = np.sum(f(0.5 * (1 / im)) * _abs2(Hr), axis=0)
        else:
            eH = np.sum((f(0.5) * (f(1) / im[(slice(None), *sl)])) * _abs2(Hr), axis=0)
        state_arr[idx] = (eE + eH).astype(np.float64).mean(axis=det["axis"]).astype(f)
    elif kind == "energy":
        ie = scene.inv_eps[(slice(None), *sl)]
        eE = np.sum((f(0.5) * (f(1) / ie)) * _abs2(Er), axis=0)
        im = scene.inv_mu
        if isinstance(im, float):
            eH = np.sum(f(0.5 * (1 / im)) * _abs2(Hr), axis=0)
        else:
            eH = np.sum((f(0.5) * (f(1) / im[(slice(None), *sl)])) * _abs2(Hr), axis=0)
        en = eE + eH
        if det["reduce_volume"]:
            state_arr[idx] = np.sum((en * f(det["weight"])).astype(np.float64)).astype(f)
        else:
            state_arr[idx] = en
    elif kind == "field":  # a float32 state keeps the real part of complex fields
        both = np.real(np.concatenate([Er, Hr]))
        sel = np.stack([both[COMPONENTS.index(cn)] for cn in det["components"]])
        if det.get("reduce_volume", False):  # volume-weighted mean, uniform cells (detector.py:95-109)
            state_arr[idx] = sel.reshape(sel.shape[0], -1).astype(np.float64).mean(axis=1).astype(f)
        else:
            state_arr[idx] = sel
    elif kind == "flux":
        keep_all = det.get("keep_all_components", False)
        pfs = []
        for a in ((0, 1, 2) if keep_all else (det["axis"],)):
            b1, b2 = (a + 1) % 3, (a + 2) % 3
            pf = _flux(Er, Hr, b1, b2)  # compute_poynting_flux(E, H).real (metrics.py:99-117)
            if det["sign"] < 0:
                pf = -pf
            pfs.append(pf)
        if det["reduce_volume"]:
            r = [np.sum((pf * f(det["weight"])).astype(np.float64)).astype(f) for pf in pfs]
            state_arr[idx] = r
        else:
            state_arr[idx] = np.stack(pfs) if keep_all else pfs[0]
    elif kind == "closed_flux":  # net_poynting_flux_through_box (metrics.py:120-157), poynting_flux.py:265-286
        w = f(det["weight"])
        net = f(0)
        for a in det["axes"]:
            b1, b2 = (a + 1) % 3, (a + 2) % 3
            S = _flux(Er, Hr, b1, b2)
            mx = np.sum((np.take(S, -1, axis=a) * w).astype(np.float64)).astype(f)
            mn = np.sum((np.take(S, 0, axis=a) * w).astype(np.float64)).astype(f)
            net = f(f(net + mx) - mn)
        state_arr[idx] = -net if det.get("orientation", "outward") == "inward" else net
    elif kind == "phasor":
        both = np.concatenate([Er, Hr])
        sel = np.stack([both[COMPONENTS.index(cn)] for cn in det["components"]])
        cos_t, sin_t = det["cos"][:, n], det["sin"][:, n]
        scale = f(det["scale"])
        w = f(det["window"][n])
        re = ((sel[None] * cos_t.reshape(-1, *([1] * sel.ndim))) * scale) * w
        im_ = ((sel[None] * sin_t.reshape(-1, *([1] * sel.ndim))) * scale) * w
        if det.get("reduce_volume", False):  # box mean of the per-cell phasors (phasor.py:221-223)
            acc = np.complex128 if np.iscomplexobj(re) else np.float64
            F = scene.field_dtype
            re = re.reshape(*re.shape[:2], -1).astype(acc).mean(axis=-1).astype(F)
            im_ = im_.reshape(*im_.shape[:2], -1).astype(acc).mean(axis=-1).astype(F)
        sgn = -1 if det.get("inverse", False) else 1
        state_arr[0] = state_arr[0] + sgn * (re + 1j * im_).astype(np.complex64 if f == np.float32 else np.complex128)
    return state_arr


def init_detector_state(det, scene):
    f = scene.dtype
    shp = tuple(s.stop - s.start for s in det["slc"])
    nrec = int(np.sum(det["time_to_idx"] >= 0))
    if det["kind"] == "energy":
        return np.zeros((nrec, 1) if det["reduce_volume"] else (nrec, *shp), f)
    if det["kind"] == "energy_mean":
        return np.zeros((nrec, *[s_ for a, s_ in enumerate(shp) if a != det["axis"]]), f)
    if det["kind"] == "field":
        nc = len(det["components"])
        return np.zeros((nrec, nc) if det.get("reduce_volume", False) else (nrec, nc, *shp), f)
    if det["kind"] == "flux":
        m = 3 if det.get("keep_all_components", False) else 1
        if det["reduce_volume"]:
            return np.zeros((nrec, m), f)
        return np.zeros((nrec, m, *shp) if m == 3 else (nrec, *shp), f)
    if det["kind"] == "closed_flux":
        return np.zeros((nrec, 1), f)
    if det["kind"] == "phasor":
        cdt = np.complex64 if f == np.float32 else np.complex128
        return np.zeros((1, det["cos"].shape[0], len(det["components"]),
                         *(() if det.get("reduce_volume", False) else shp)), cdt)
    raise ValueError(det["kind"])


def phasor_poynting_flux(phasors, axis, sign, weight, keep_all=False, continuous=True):
    """PhasorPoyntingFluxDetector.compute_poynting_flux (poynting_flux.py:361-387): phasors
    (num_freqs, 6, *plane); S = Re(E x conj(H)) (58-71), signed, area-weighted sums over the plane,
    halved for continuous scaling."""
    E, H = phasors[:, :3], np.conj(phasors[:, 3:])
    S = np.stack([(E[:, (a + 1) % 3] * H[:, (a + 2) % 3] - E[:, (a + 2) % 3] * H[:, (a + 1) % 3]).real
                  for a in range(3)], axis=1) * sign
    if keep_all:
        flux = np.sum(S * weight, axis=(2, 3, 4))
    else:
        flux = np.sum(S[:, axis] * weight, axis=(1, 2, 3))
    return 0.5 * flux if continuous else flux


def phasor_poynting_flux_cotangent(phasors, axis, sign, weight, g, keep_all=False, continuous=True):
    """Cotangent (dL/dRe + i dL/dIm) on the phasors (num_freqs, 6, *plane) of L = sum g * flux with
    flux = phasor_poynting_flux(...): g has the flux's shape. Re(a conj(b)) = Re a Re b + Im a Im b,
    so dRe(a conj(b))/da = b and /db = a in this convention."""
    E, H = phasors[:, :3], phasors[:, 3:]
    out = np.zeros_like(phasors)
    scale = (0.5 if continuous else 1.0) * sign * weight
    for t, a in enumerate((0, 1, 2) if keep_all else (axis,)):
        ga = (g[:, t] if keep_all else g).reshape(-1, *([1] * (phasors.ndim - 2))) * scale
        b1, b2 = (a + 1) % 3, (a + 2) % 3
        out[:, b1] += ga * H[:, b2]
        out[:, b2] -= ga * H[:, b1]
        out[:, 3 + b2] += ga * E[:, b1]
        out[:, 3 + b1] -= ga * E[:, b2]
    return out


def closed_phasor_net_flux(faces, weight, inward=False, continuous=True):
    """ClosedSurfacePhasorPoyntingFluxDetector.compute_net_flux (poynting_flux.py:508-533): faces
    maps (axis, side) to the (1, num_freqs, 6, *plane) state of that face."""
    net = 0.0
    for (a, side), st in faces.items():
        net = net + (1 if side == "max" else -1) * phasor_poynting_flux(st[0], a, 1, weight, continuous=False)
    net = -net if inward else net
    return 0.5 * net if continuous else net


def gaussian_mode(u, v, axis, radius, direction, pol_axis, n, wavelength, divergence_deg=0.0):
    """Reference Gaussian cross-section of GaussianModeOverlapDetector (detectors/mode.py:26-141,
    normal incidence): E = A pol, H = n (k_hat x pol) A, A = exp(-r^2/radius^2) exp(i k r^2
    tan(div)/(2 radius)), k = 2 pi n / wavelength, r^2 = u^2 + v^2 from the transverse offsets."""
    r2 = np.asarray(u, np.float64) ** 2 + np.asarray(v, np.float64) ** 2
    A = np.exp(-r2 / radius**2) * np.exp(1j * np.pi * n / wavelength * math.tan(math.radians(divergence_deg))
                                         / radius * r2)
    e = np.zeros(3)
    e[pol_axis] = 1.0
    k = np.zeros(3)
    k[axis] = 1.0 if direction == "+" else -1.0
    h = np.cross(k, e)
    return A[None] * e.reshape(3, *([1] * r2.ndim)), A[None] * (n * h).reshape(3, *([1] * r2.ndim))


def unit_flux(mE, mH, axis):
    """Scale (mE, mH) so that |sum 1/2 Re(conj(E) x H)_axis| = 1 (metrics.py:160-221)."""
    a1, a2 = (axis + 1) % 3, (axis + 2) % 3
    p = abs(float(np.sum(0.5 * np.real(np.conj(mE[a1]) * mH[a2] - np.conj(mE[a2]) * mH[a1]))))
    return mE / math.sqrt(p), mH / math.sqrt(p)


def mode_overlap(ph, mE, mH, axis, pulse=False):
    """compute_overlap_to_mode (detectors/mode.py:379-421) on a uniform grid: ph is the (6, *plane)
    phasor of one frequency; sum((mE x conj H) + (conj E x mH))_axis, over 4 unless pulse."""
    E, H = ph[:3].astype(np.complex128), ph[3:].astype(np.complex128)
    a1, a2 = (axis + 1) % 3, (axis + 2) % 3
    t = (mE[a1] * np.conj(H[a2]) - mE[a2] * np.conj(H[a1])) + (np.conj(E[a1]) * mH[a2] - np.conj(E[a2]) * mH[a1])
    s = np.sum(t)
    return s if pulse else s / 4.0


def phasor_tables(omegas, T, dt, window=None, scaling_mode="continuous", dtype=np.float32):
    """phasor.py:139-175, 190-235: phase = omega(f32) * (n*dt)(f32); exp(i*phase)."""
    f = dtype
    w32 = np.asarray(omegas, dtype=f)
    time = (np.arange(T, dtype=f) * f(dt)).astype(f)
    phase = (w32[:, None] * time[None, :]).astype(f)
    cos_t, sin_t = np.cos(phase).astype(f), np.sin(phase).astype(f)
    win = np.ones(T, f) if window is None else np.asarray(window, f)
    scale = 2 / float(np.sum(win)) if scaling_mode == "continuous" else 1
    return cos_t, sin_t, win, scale


# ---------------------------------------------------------------- full runs
def forward_step(st, scene, n, det_states=None, record=None, simulate_boundaries=True):
    """forward.py:83-156: H_prev <- H; E half; H half; [record]; [detectors]."""
    H_prev = st.H.copy()
    K_E = update_E(st, scene, n, simulate_boundaries)
    update_H(st, scene, n, simulate_boundaries)
    if record is not None:
        record.append(collect_interfaces(st, scene))
    if det_states is not None:
        for i, det in enumerate(scene.detectors):
            if not det.get("inverse", False):  # update.py:1073: the forward records forward detectors
                det_states[i] = detector_update(det, det_states[i], n, st.E, H_prev, st.H, scene)
    return K_E


def run(scene, T, st=None, det_states=None, record=None, t0=0):
    if st is None:
        st = init_state(scene)
    if det_states is None:
        det_states = [init_detector_state(d, scene) for d in scene.detectors]
    for n in range(t0, T):
        forward_step(st, scene, n, det_states, record)
    return st, det_states


# ---------------------------------------------------------------- reversible pieces
def collect_interfaces(st, scene):
    """fdtd/misc.py:10-35: innermost PML layer of E and H per slab (E first, then H)."""
    out = []
    for name in ("E", "H"):
        arr = st.E if name == "E" else st.H
        for s in scene.slabs:
            out.append(arr[(slice(None), *s.interface_slc(scene.shape))].copy())
    return out


def interface_keys(scene):
    """Record entry names "{face}_{E|H}" in collect_interfaces order (fdtd/misc.py:31-35)."""
    faces = [f"{'min' if s.side == '-' else 'max'}_{'xyz'[s.axis]}" for s in scene.slabs]
    return [f"{f}_{name}" for name in ("E", "H") for f in faces]


def add_interfaces(st, scene, values):
    """fdtd/misc.py:38-66."""
    i = 0
    for name in ("E", "H"):
        arr = st.E if name == "E" else st.H
        for s in scene.slabs:
            arr[(slice(None), *s.interface_slc(scene.shape))] = values[i]
            i += 1


def reverse_step(st, scene, n, record_n):
    """backward.py:62-135 with reset_fields=False as used by the reversible VJP (fdtd.py:219-226):
    restore interfaces, update_H_reverse (update.py:856-1007), update_E_reverse (526-686),
    CPML psi frozen (simulate_boundaries=False)."""
    f = scene.dtype
    c = f(scene.c)
    add_interfaces(st, scene, record_n)
    H = st.H.copy()
    _apply_sources(H, scene, n, "H", inverse=True)
    curl, d = curl_E_raw(st.E, scene.periodic)
    apply_cpml(curl, d, st.psiH, scene.slabs, "H", False, scene.shape)
    im = scene.inv_mu if not isinstance(scene.inv_mu, float) else f(scene.inv_mu)
    x = loss_factors(scene, "H")
    sig = tensor_sigma(scene, "H")
    if sig is not None:  # update.py:930-1000: A_rev avg(H) + B_rev avg(K)
        A, B = aniso_matrices(expand3x3(im, f), expand3x3(sig, f), scene.c, 1.0 / ETA0, reverse=True)
        st.H = aniso_update(A, B, H, np.stack(curl), "H", +1, scene.periodic).astype(f)
    elif is_tensor(im):  # update.py:960-1000 with A = I
        st.H = (H + tensor_term(c * im, np.stack(curl), "H", scene.periodic)).astype(f)
    elif x is None:
        st.H = (H + (c * np.stack(curl)) * im).astype(f)
    else:  # update.py:921-926
        st.H = ((H * (f(1) + x) + (c * np.stack(curl)) * im) / (f(1) - x)).astype(f)
    apply_walls(st.H, scene, "pmc")  # update.py:1004
    E = st.E.copy()
    _apply_sources(E, scene, n, "E", inverse=True)
    curl, d = curl_H_raw(st.H, scene.periodic)
    apply_cpml(curl, d, st.psiE, scene.slabs, "E", False, scene.shape)
    x = loss_factors(scene, "E")
    sig = tensor_sigma(scene, "E")
    if sig is not None:  # update.py:611-680: A_rev avg(E) - B_rev avg(K)
        A, B = aniso_matrices(expand3x3(scene.inv_eps, f), expand3x3(sig, f), scene.c, ETA0, reverse=True)
        st.E = aniso_update(A, B, E, np.stack(curl), "E", -1, scene.periodic).astype(f)
    elif is_tensor(scene.inv_eps):  # update.py:611-680 with A = I
        st.E = (E - tensor_term(c * scene.inv_eps, np.stack(curl), "E", scene.periodic)).astype(f)
    elif x is None:
        st.E = (E - (c * np.stack(curl)) * scene.inv_eps).astype(f)
    else:  # update.py:600-606
        st.E = ((E * (f(1) + x) - (c * np.stack(curl)) * scene.inv_eps) / (f(1) - x)).astype(f)
    apply_walls(st.E, scene, "pec")  # update.py:683


def full_backward(st, scene, t1, t0, record, det_states=None, reset_fields=True):
    """backward.py:18-135: steps n = t1-1 .. t0 of the time-reversed run. Each restores the
    recorded interfaces, keeps H (after the restore) as H_prev, undoes H then E (reverse_step),
    zeroes E and H over every CPML slab when reset_fields (perfectly_matched_layer.py:227-229;
    periodic faces copy a layer onto itself, bloch.py:99-105), then records the inverse
    detectors at step n (update.py:1073)."""
    for n in range(t1 - 1, t0 - 1, -1):
        add_interfaces(st, scene, record[n])
        H_prev = st.H.copy()
        reverse_step(st, scene, n, record[n])
        if reset_fields:
            for s in scene.slabs:
                st.E[(slice(None), *s.slc(scene.shape))] = 0
                st.H[(slice(None), *s.slc(scene.shape))] = 0
        if det_states is not None:
            for i, det in enumerate(scene.detectors):
                if det.get("inverse", False):
                    det_states[i] = detector_update(det, det_states[i], n, st.E, H_prev, st.H, scene)
    return st, det_states


# ---------------------------------------------------------------- discrete adjoint
def _slab_scatter_T(vec3, s, which_sign_first=True):
    """Transpose of the CPML scatter: returns (g1, g2) on the slab from a 3-component vector."""
    sl = s.slc(None)
    i, j = (s.axis + 1) % 3, (s.axis + 2) % 3
    return -vec3[i][sl], vec3[j][sl]


def _deriv_T(name, v_full, wrap=NOWRAP):
    """Transpose of a single backward/forward difference along an axis (zero padded or wrapped).

    For backward 'd{ax}H*' (v[i]-v[i-1]) the transpose is u[i] - u[i+1]; for forward 'd{ax}E*'
    (v[i+1]-v[i]) it is u[i-1] - u[i]."""
    ax = "xyz".index(name[1])
    if name[2] == "H":
        return v_full - _next(v_full, ax, wrap[ax])
    return _prev(v_full, ax, wrap[ax]) - v_full


def adjoint_step(lam, scene, H_n, psiE_used, det_inject=None, E_n=None):
    """Transpose of one forward step at fixed materials (pure linear algebra on the Yee step).

    lam: dict with lamE, lamH, lampsiE (list), lampsiH (list) = dL/d(state after step n).
    H_n, E_n: reconstructed H and E at the start of step n (E_n is read with conductivity only);
    psiE_used: the psi_E fed to the step.
    With conductivity the E update is E' = ((1-x) E + c inv_eps K) / (1+x), x = a inv_eps, whose
    transposes scale by (1-x)/(1+x) (the old field) and 1/(1+x) (the curl), and whose inv_eps
    derivative is (c K - 2 a E) / (1+x)^2; sigma itself is a constant (initialization.py:287).
    det_inject: optional (vE, vHc, vH1) detector cotangents (detector_cotangent): vE and vH1 add
    to lamE and lamH before propagation; vHc/2 adds to lamH before and vHc/2 to the result (the
    H_prev half of the time-centred H).
    Returns (new lam, grad_inv_eps contribution with the inv_eps component shape).
    """
    f = scene.dtype
    c = f(scene.c)
    lamE, lamH = lam["lamE"].copy(), lam["lamH"].copy()
    if det_inject is not None:
        vE, vH, vH1 = det_inject
        lamE = lamE + vE
        lamH = lamH + f(0.5) * vH + vH1
    im = scene.inv_mu if not isinstance(scene.inv_mu, float) else f(scene.inv_mu)
    ie = scene.inv_eps
    shape = scene.shape
    # ---- transpose of the H update: H' = Z_H(H - c*mu^-1*(curlE(E') + S_H psiH')), with Z_H the
    # PMC zeroing (its own transpose), psiH' = bH psiH + aH dE(E')
    apply_walls(lamH, scene, "pmc")
    xH = loss_factors(scene, "H")
    xE = loss_factors(scene, "E")
    if is_tensor(im):  # the transposed averages carry each row's cotangent back to K_j
        wH = tensor_term_T(c * im, lamH, "H", scene.periodic)
    else:
        wH = (c * im) * lamH  # cotangent on K' with the minus sign folded in below
    if xH is not None:
        wH = wH / (f(1) + xH)
    new_lampsiH = []
    muE = lamE.copy()
    # -curlE^T(wH) = -curlH(wH)
    cr, _ = curl_H_raw(wH, scene.periodic)
    muE = muE - np.stack(cr)
    for s, (lp1, lp2) in zip(scene.slabs, lam["lampsiH"]):
        bE, aE, ikE, bH, aH, ikH = s.coeffs
        g1, g2 = _slab_scatter_T(-wH, s)  # S_H^T of (-wH)
        gam1, gam2 = lp1 + g1, lp2 + g2
        new_lampsiH.append((s.bcast(bH) * gam1, s.bcast(bH) * gam2))
        n1, n2 = _PML_DERIV["E"][s.axis]
        kap = not np.all(ikH == 1)  # corr = (1/kappa - 1) d + psi': d also feeds corr directly
        for nm, gam, g in ((n1, gam1, g1), (n2, gam2, g2)):
            full = np.zeros(shape, f)
            v = s.bcast(aH) * gam
            if kap:
                v = v + s.bcast((ikH - 1.0).astype(f)) * g
            full[s.slc(shape)] = v
            comp = "xyz".index(nm[3])
            muE[comp] = muE[comp] + _deriv_T(nm, full, scene.periodic)
    # ---- transpose of the E update: E' = Z_E(E + c*inv_eps*(curlH(H) + S_E psiE')), with Z_E the
    # PEC zeroing, psiE' = bE psiE + aE dH(H)
    apply_walls(muE, scene, "pec")
    sigE = tensor_sigma(scene, "E")
    if tensor_sigma(scene, "H") is not None:
        raise NotImplementedError("oracle: the adjoint through a magnetic conductivity tensor")
    if sigE is not None:  # E' = A avg(E) + B avg(K): the curl's cotangent through B, E's through A
        A_E, B_E = aniso_matrices(expand3x3(ie, f), expand3x3(sigE, f), scene.c, ETA0)
        wE = tensor_term_T(B_E, muE, "E", scene.periodic)
        xE = None
    else:
        wE = tensor_term_T(c * ie, muE, "E", scene.periodic) if is_tensor(ie) else (c * ie) * muE
    if xE is not None:
        wE = wE / (f(1) + xE)
    newH = lamH.copy() if xH is None else ((f(1) - xH) / (f(1) + xH)) * lamH
    cr, _ = curl_E_raw(wE, scene.periodic)  # curlH^T = curlE
    newH = newH + np.stack(cr)
    new_lampsiE = []
    for s, (lp1, lp2) in zip(scene.slabs, lam["lampsiE"]):
        bE, aE, ikE, bH, aH, ikH = s.coeffs
        g1, g2 = _slab_scatter_T(wE, s)
        gam1, gam2 = lp1 + g1, lp2 + g2
        new_lampsiE.append((s.bcast(bE) * gam1, s.bcast(bE) * gam2))
        n1, n2 = _PML_DERIV["H"][s.axis]
        kap = not np.all(ikE == 1)
        for nm, gam, g in ((n1, gam1, g1), (n2, gam2, g2)):
            full = np.zeros(shape, f)
            v = s.bcast(aE) * gam
            if kap:
                v = v + s.bcast((ikE - 1.0).astype(f)) * g
            full[s.slc(shape)] = v
            comp = "xyz".index(nm[3])
            newH[comp] = newH[comp] + _deriv_T(nm, full, scene.periodic)
    if det_inject is not None:
        newH = newH + f(0.5) * det_inject[1]
    # gradient wrt inv_eps: c * muE * K^{n+1/2}, K from H_n with psiE' = bE psiE_used + aE dH
    curl, d = curl_H_raw(H_n, scene.periodic)
    apply_cpml(curl, d, [(p1.copy(), p2.copy()) for p1, p2 in psiE_used], scene.slabs, "E", True, shape)
    K = np.stack(curl)
    if sigE is not None:  # through A = M1^-1 M2 and B = c M1^-1 inv_eps (fdtd/misc.py:69-98)
        g = aniso_inv_gradient(muE, E_n, K, expand3x3(ie, f), expand3x3(sigE, f), scene.c, ETA0, scene.periodic)
        if ie.shape[0] != 9:  # the diagonal of the expanded tensor
            g = g[[0, 4, 8]] if ie.shape[0] == 3 else (g[0] + g[4] + g[8])[None]
        lamE_n = tensor_term_T(A_E, muE, "E", scene.periodic)
    elif is_tensor(ie):  # d(B_ij avg(K_j))/d inv_eps_ij = c avg(K_j -> i)
        g = np.stack([c * muE[i] * (K[j] if i == j else avg_E(K, j, i, scene.periodic))
                      for i in range(3) for j in range(3)])
        lamE_n = muE
    elif xE is None:
        g = c * muE * K
        lamE_n = muE
    else:
        dE = f(1) + xE
        g = muE * ((c * K - (c * scene.sigma_E) * f(ETA0) * E_n) / (dE * dE))
        lamE_n = ((f(1) - xE) / dE) * muE
    if ie.shape[0] == 1:
        g = np.sum(g, axis=0, keepdims=True)
    return {"lamE": lamE_n, "lamH": newH, "lampsiE": new_lampsiE, "lampsiH": new_lampsiH}, g


def detector_cotangent(det, cot, n, E, H_prev, H, scene, grad_eps=False):
    """Transpose of detector_update at step n with respect to the fields it reads, for the cotangent
    `cot` of the whole detector state (its layout; a complex entry z takes dL/dRe z + i dL/dIm z).
    E and H are the fields after step n and H_prev the H before it. Returns (vE, vHc, vH1): the
    cotangents on E, on the time-centred Hc = (H_prev + H)/2 that exactly interpolating detectors
    read, and on H itself, which point-sampled detectors read. An energy detector also reads
    inv_eps (energy.py:90-143 takes it as an argument, so the reference's step VJP differentiates
    it); with grad_eps=True a fourth item holds that direct dL/d inv_eps (None otherwise)."""
    f = scene.dtype
    zero = np.zeros_like(E)
    idx = det["time_to_idx"][n]
    if idx < 0:
        return (zero, zero.copy(), zero.copy()) + ((None,) if grad_eps else ())
    sl = det["slc"]
    box = (slice(None), *sl)
    exact = det.get("exact_interpolation", True)
    mir = scene.mirror
    off = [1 if m else 0 for m in mir]  # the mirror pad shifts the box
    pbox = (slice(None), *(slice(s_.start + o, s_.stop + o) for s_, o in zip(sl, off)))
    if exact:
        Ei, Hi = interpolate_fields(mirror_pad(E, mir, "E"), mirror_pad((H_prev + H) / f(2), mir, "H"),
                                    scene.periodic)
        Er, Hr = Ei[pbox], Hi[pbox]
    else:
        Er, Hr = E[box], H[box]
    ncell = int(np.prod(Er.shape[1:]))
    red = det.get("reduce_volume", False)
    g6 = np.zeros((6, *Er.shape[1:]), f)  # cotangent on the six sampled components
    g_ie = None
    kind = det["kind"]
    if kind in ("energy", "energy_mean"):  # u = 1/2 (|E|^2 / inv_eps + |H|^2 / inv_mu): du/dE = E / inv_eps
        if kind == "energy_mean":  # the plane's cotangent spread over the averaged axis
            ax = det["axis"]
            g = np.repeat(np.expand_dims(cot[idx] / f(Er.shape[1 + ax]), ax), Er.shape[1 + ax], axis=ax)
        else:
            g = f(cot[idx, 0] * det["weight"]) if red else cot[idx]
        ie_b = scene.inv_eps[box]
        g6[:3] = g * ((f(1) / ie_b) * Er)
        r = f(1) / ie_b  # du/d inv_eps = -1/2 |E|^2 / inv_eps^2
        d_ie = g * ((f(-0.5) * (r * r)) * (Er * Er))
        g_ie = np.zeros_like(scene.inv_eps)
        g_ie[box] = d_ie if scene.inv_eps.shape[0] == 3 else np.sum(d_ie, axis=0, keepdims=True)
        im = scene.inv_mu
        g6[3:] = g * (Hr * f(1 / im) if isinstance(im, float) else (f(1) / im[box]) * Hr)
    elif kind == "field":
        for c, cn in enumerate(det["components"]):
            g6[COMPONENTS.index(cn)] += f(cot[idx, c]) / f(ncell) if red else cot[idx, c]
    elif kind == "flux":
        keep_all = det.get("keep_all_components", False)
        for t, a in enumerate((0, 1, 2) if keep_all else (det["axis"],)):
            if red:
                w = f(cot[idx, t] * det["weight"] * det["sign"])
            else:
                w = (cot[idx, t] if keep_all else cot[idx]) * f(det["sign"])
            b1, b2 = (a + 1) % 3, (a + 2) % 3
            g6[b1] += w * Hr[b2]
            g6[b2] -= w * Hr[b1]
            g6[3 + b2] += w * Er[b1]
            g6[3 + b1] -= w * Er[b2]
    elif kind == "closed_flux":
        g = f(cot[idx, 0] * det["weight"] * (-1 if det.get("orientation", "outward") == "inward" else 1))
        for a in det["axes"]:
            wc = np.zeros(Er.shape[1:], f)  # +g on the max face, -g on the min face
            hi = [slice(None)] * 3
            hi[a] = slice(-1, None)
            lo = [slice(None)] * 3
            lo[a] = slice(0, 1)
            wc[tuple(hi)] += g
            wc[tuple(lo)] -= g
            b1, b2 = (a + 1) % 3, (a + 2) % 3
            g6[b1] += wc * Hr[b2]
            g6[b2] -= wc * Hr[b1]
            g6[3 + b2] += wc * Er[b1]
            g6[3 + b1] -= wc * Er[b2]
    elif kind == "phasor":
        cos_t, sin_t = det["cos"][:, n], det["sin"][:, n]
        scale, w = f(det["scale"]), f(det["window"][n])
        sgn = f(-1 if det.get("inverse", False) else 1)
        for c, cn in enumerate(det["components"]):
            for fi in range(len(cos_t)):
                z = cot[0, fi, c]
                v = ((z.real.astype(f) * cos_t[fi] + z.imag.astype(f) * sin_t[fi]) * scale) * w
                if red:
                    v = v / f(ncell)
                g6[COMPONENTS.index(cn)] += sgn * v
    else:
        raise ValueError(kind)
    if exact:
        gE, gH = np.zeros_like(mirror_pad(E, mir, "E")), np.zeros_like(mirror_pad(H, mir, "H"))
        gE[pbox], gH[pbox] = g6[:3], g6[3:]
        vE, vHc = interpolate_fields_T(gE, gH, scene.periodic)
        out = (mirror_pad_T(vE, mir, "E"), mirror_pad_T(vHc, mir, "H"), zero)
    else:
        gE, gH = np.zeros_like(E), np.zeros_like(H)
        gE[box], gH[box] = g6[:3], g6[3:]
        out = (gE, zero, gH)
    return out + ((g_ie,) if grad_eps else ())


def interpolate_fields_T(gE, gH, wrap=NOWRAP):
    """Transpose of interpolate_fields (linear in E and Hc)."""
    f = gE.dtype.type
    h = f(0.5)
    Pt = lambda a, ax: _next(a, ax, wrap[ax])  # transpose of _prev
    Nt = lambda a, ax: _prev(a, ax, wrap[ax])  # transpose of _next (up)
    vE = np.zeros_like(gE)
    vH = np.zeros_like(gH)
    # ex = (h(Ex + P0 Ex) + h(U Ex + P0 U Ex)) / 2
    g = gE[0] / f(2.0)
    t = h * g
    vE[0] = t + Pt(t, 0) + Nt(t + Pt(t, 0), 2)
    g = gE[1] / f(2.0)
    t = h * g
    vE[1] = t + Pt(t, 1) + Nt(t + Pt(t, 1), 2)
    vE[2] = gE[2]
    t = h * gH[0]
    vH[0] = t + Pt(t, 1)
    t = h * gH[1]
    vH[1] = t + Pt(t, 0)
    g = gH[2] / f(2.0)
    # lo_xy = h*(h*(Hz + P0 Hz) + h*(P1 Hz + P0 P1 Hz)); hi same on U Hz
    t = h * h * g
    core = t + Pt(t, 0) + Pt(t, 1) + Pt(Pt(t, 1), 0)
    vH[2] = core + Nt(core, 2)
    return vE, vH


# ---------------------------------------------------------------- reversible gradient (fdtd.py:39-418)
def slice_boundaries(T, num_checkpoints):
    """Slice boundaries of the sliced reversible backward pass (fdtd.py:20-36)."""
    k = num_checkpoints + 1
    return [round(i * T / k) for i in range(k + 1)]


def _copy_state(st):
    return State(st.E.copy(), st.H.copy(), [(a.copy(), b.copy()) for a, b in st.psiE],
                 [(a.copy(), b.copy()) for a, b in st.psiH])


def reversible_gradient(scene, T, det_cotangents, num_checkpoints=0, recorder_modules=None, exact_states=False):
    """dL/d inv_eps by the reference's reversible scheme: forward with interface recording, then per
    reverse step restore interfaces, reconstruct with frozen psi, and apply the transposed step at
    the reconstructed state (fdtd.py:215-251). With num_checkpoints > 0 the run is cut into slices
    (fdtd.py:106-121): the forward stores the full field state (E, H, psi) at every interior
    boundary, and the reverse resets its reconstructed state to it on reaching that boundary
    (fdtd.py:295-313), so each slice reconstructs with the psi of its own end frozen.
    det_cotangents(det_states) -> list of cotangent arrays (or None) in the detector state
    layouts. recorder_modules: compression modules of the interface record as
    record_codec.RecordPipeline takes them (None: the exact record). exact_states=True keeps every
    forward state and applies the transposed steps at those states instead of the reconstructed
    ones (the adjoint of the forward trajectory itself, for checking the transposition where the
    reverse step is not the forward step's inverse). Returns (grad, det_states, final_state)."""
    from . import record_codec
    f = scene.dtype
    st = init_state(scene)
    ds = [init_detector_state(d, scene) for d in scene.detectors]
    bounds = slice_boundaries(T, num_checkpoints)
    interior = set(bounds[1:-1])
    ckpts = {}
    record = []
    pipe = None if not recorder_modules else record_codec.RecordPipeline(recorder_modules, T)
    keys = interface_keys(scene)
    states = {}
    for n in range(T):
        if exact_states:
            states[n] = _copy_state(st)
        forward_step(st, scene, n, ds, record)
        if pipe is not None:
            pipe.compress(dict(zip(keys, record[-1])), n)
            record[-1] = None
        if n + 1 in interior:
            ckpts[n + 1] = _copy_state(st)
    cots = det_cotangents(ds)
    final = _copy_state(st)
    psiE_frozen = [(a.copy(), b.copy()) for a, b in st.psiE]
    lam = {"lamE": np.zeros_like(st.E), "lamH": np.zeros_like(st.H),
           "lampsiE": [(np.zeros_like(a), np.zeros_like(b)) for a, b in st.psiE],
           "lampsiH": [(np.zeros_like(a), np.zeros_like(b)) for a, b in st.psiH]}
    grad = np.zeros_like(scene.inv_eps)
    for n in reversed(range(T)):
        if n + 1 in ckpts:  # the reconstructed state has reached a boundary: reset it
            st = _copy_state(ckpts[n + 1])
            psiE_frozen = [(a.copy(), b.copy()) for a, b in st.psiE]
        rec_n = record[n] if pipe is None else [v.astype(f) for v in (lambda d: [d[k] for k in keys])(pipe.decompress(n))]
        add_interfaces(st, scene, rec_n)
        E1, H1 = st.E.copy(), st.H.copy()
        if exact_states:
            st = _copy_state(states[n])
            psiE_frozen = [(a.copy(), b.copy()) for a, b in st.psiE]
        else:
            reverse_step(st, scene, n, rec_n)
        inject = None
        for det, cot in zip(scene.detectors, cots):
            if cot is None or det["time_to_idx"][n] < 0:
                continue
            *v, g_ie = detector_cotangent(det, cot, n, E1, st.H, H1, scene, grad_eps=True)
            inject = tuple(v) if inject is None else tuple(a + b for a, b in zip(inject, v))
            if g_ie is not None:
                grad = grad + g_ie
        lam, g = adjoint_step(lam, scene, st.H, psiE_frozen, inject, E_n=st.E)
        grad = grad + g
    return grad.astype(f), ds, final


def detector_losses_fd(spec_scene_fn, rho, idx, h, T, losses):
    """Central finite differences w.r.t. rho[idx] of each loss in losses(det_states) -> list of
    floats (FP64 scenes)."""
    vals = []
    for s in (+1, -1):
        r = rho.copy()
        r[idx] += s * h
        _, ds = run(spec_scene_fn(r), T)
        vals.append(np.asarray(losses(ds), np.float64))
    return (vals[0] - vals[1]) / (2 * h)


def flux_loss_and_grad_fd(spec_scene_fn, rho, idx, h, T):
    """Central finite difference of L = sum_n flux_n w.r.t. rho[idx] (FP64 scenes)."""
    vals = []
    for s in (+1, -1):
        r = rho.copy()
        r[idx] += s * h
        sc = spec_scene_fn(r)
        _, ds = run(sc, T)
        vals.append(float(np.sum(ds[0])))
    return (vals[0] - vals[1]) / (2 * h)
