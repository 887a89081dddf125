"""Build oracle Scenes from a neutral scenario spec (TEST INFRASTRUCTURE, see fdtd_oracle.py).

A spec is a plain dict that states a simulation the way a user of fdtdx would: shape, spacing,
materials in boxes, a continuous design region, CPML thickness, sources and detectors. The
oracle derives every table (ε⁻¹, CPML coefficients, source amplitudes, ADE coefficients, detector
schedules) from that spec on its own, following the reference files cited below, so the
product's host-side lowering can be checked against it.
"""

from __future__ import annotations

import math

import numpy as np

from . import fdtd_oracle as O


def on_list(T, dt, spec_switch=None):
    """core/switch.py:54-74 (+ is_on_at_time_step 108-222) for the supported keys."""
    sw = spec_switch or {}
    if sw.get("fixed_on_time_steps") is not None:
        lst = [False] * T
        for t in sw["fixed_on_time_steps"]:
            lst[t] = True
        return lst
    if sw.get("is_always_off"):
        return [False] * T
    start = sw.get("start_time", 0.0) or 0.0
    end = sw.get("end_time", math.inf)
    interval = sw.get("interval", 1)
    out = []
    for t in range(T):
        tp = t * dt
        out.append(bool(start <= tp <= end) and t % interval == 0)
    return out


def time_to_idx(lst):
    idx = np.full(len(lst), -1, np.int32)
    c = 0
    for t, on in enumerate(lst):
        if on:
            idx[t] = c
            c += 1
    return idx


def build_inv_eps(spec, dtype):
    """_init_arrays material fill (initialization.py:844-870) + continuous device (404-417)."""
    shape = tuple(spec["shape"])
    f = dtype
    def comps(e):
        return (e,) if np.isscalar(e) else tuple(e)
    all_eps = [spec.get("background_eps", 1.0)] + [b["eps"] for b in spec.get("boxes", [])]
    ncomp = 1 if all(np.isscalar(e) for e in all_eps) else (9 if any(len(comps(e)) == 9 for e in all_eps) else 3)
    if spec.get("design") is not None and ncomp == 3:
        raise NotImplementedError("diagonal design")
    inv = np.zeros((ncomp, *shape), f)

    def fill(slc, eps):
        e = comps(eps)
        if ncomp == 9:  # the inverse of the 3x3 tensor, row-major (initialization.py:866-873)
            m = np.diag(e * 3 if len(e) == 1 else e) if len(e) != 9 else np.asarray(e, np.float64).reshape(3, 3)
            val = np.linalg.inv(np.asarray(m, np.float64)).reshape(9).astype(f)
        else:
            if ncomp == 3 and len(e) == 1:
                e = e * 3
            val = (f(1) / np.array(e, dtype=f)).astype(f)
        inv[(slice(None), *slc)] = val[:, None, None, None]

    fill(tuple(slice(None) for _ in range(3)), spec.get("background_eps", 1.0))
    for b in spec.get("boxes", []):
        fill(tuple(slice(*r) for r in b["slc"]), b["eps"])
    d = spec.get("design")
    if d is not None:
        rho = np.asarray(d["rho"], f)
        ea, eb = f(d["eps_a"]), f(d["eps_b"])
        perm = (ea + rho * (eb - ea)).astype(f)
        dsl = tuple(slice(*r) for r in d["slc"])
        if ncomp == 9:  # an isotropic design in a tensor array: the diagonal, zero off-diagonals
            inv[(slice(None), *dsl)] = 0
            for q in (0, 4, 8):
                inv[(q, *dsl)] = (f(1) / perm).astype(f)
        else:
            inv[(0, *dsl)] = (f(1) / perm).astype(f)
    return inv


def build_inv_mu(spec, dtype):
    """Inverse permeability painted box by box (initialization.py:872-890): the scalar 1.0 when no
    box sets "mu", else one component when every value is scalar, else three."""
    boxes = spec.get("boxes", [])
    if not any("mu" in b for b in boxes):
        return 1.0
    f = dtype
    mus = [b.get("mu", 1.0) for b in boxes]
    ncomp = 1 if all(np.isscalar(v) for v in mus) else (9 if any(np.size(v) == 9 for v in mus) else 3)
    if ncomp == 9:
        inv = np.zeros((9, *tuple(spec["shape"])), f)
        inv[[0, 4, 8]] = 1
    else:
        inv = np.ones((ncomp, *tuple(spec["shape"])), f)
    for b in boxes:
        v = b.get("mu", 1.0)
        if ncomp == 9:
            m = np.diag((v,) * 3 if np.isscalar(v) else tuple(v)) if np.size(v) != 9 else np.reshape(v, (3, 3))
            val = np.linalg.inv(np.asarray(m, np.float64)).reshape(9).astype(f)
        else:
            v = (v,) * ncomp if np.isscalar(v) else tuple(v)
            val = (f(1) / np.asarray(v, f)).astype(f)
        inv[(slice(None), *tuple(slice(*r) for r in b["slc"]))] = val[:, None, None, None]
    return inv


def build_sigma(spec, dtype, key, dt, courant):
    """Electric ("sigma_E") or magnetic ("sigma_H") conductivity painted box by box in order
    (initialization.py:757-783, 892-930): allocated only when some box conducts, one component when
    every value is scalar, else three, each scaled by c0*dt/courant (initialization.py:786)."""
    boxes = spec.get("boxes", [])
    if not any(np.any(np.asarray(b.get(key, 0.0)) != 0) for b in boxes):
        return None
    f = dtype
    ncomp = 1 if all(np.isscalar(b.get(key, 0.0)) for b in boxes) else 3
    if any(not np.isscalar(b.get(key, 0.0)) and len(b.get(key)) == 9 for b in boxes):
        ncomp = 9  # a full conductivity tensor (row-major 3 x 3); scalars and 3-tuples on its diagonal
    spacing = O.C0 * dt / courant
    sig = np.zeros((ncomp, *tuple(spec["shape"])), f)
    for b in boxes:
        v = b.get(key, 0.0)
        if ncomp == 9:
            d = (v,) * 3 if np.isscalar(v) else tuple(v)
            v = d if len(d) == 9 else (d[0], 0, 0, 0, d[1], 0, 0, 0, d[2])
        v = (v,) * ncomp if np.isscalar(v) else tuple(v)
        sig[(slice(None), *tuple(slice(*r) for r in b["slc"]))] = \
            (np.asarray(v, f) * f(spacing)).astype(f)[:, None, None, None]
    return sig


def _orientation(p):
    """The unit orientation of an oriented pole tuple (its trailing 3-tuple), or None."""
    u = p[4] if p[0] == "lorentz" and len(p) > 4 else (p[3] if p[0] == "drude" and len(p) > 3 else None)
    if u is None:
        return None
    u = np.asarray(u, np.float64)
    return u / np.linalg.norm(u)


def ade_coefficients(poles, dt, tensor=False):
    """dispersion.py:789-857 for Lorentz ('lorentz', w0, gamma, d_eps[, u]) and Drude ('drude', wp,
    gamma[, u]); with tensor=True c3 is the row-major 3x3 coupling per pole
    (compute_pole_coefficients_tensor, dispersion.py:888-952): (K dt^2 / D) u u^T for a pole
    oriented along u, (K dt^2 / D) on the diagonal otherwise."""
    n = len(poles)
    c1 = np.zeros(n); c2 = np.zeros(n); c3 = np.zeros((n, 9)) if tensor else np.zeros(n)
    for i, p in enumerate(poles):
        if p[0] == "lorentz":
            w0, g, de = p[1], p[2], p[3]
            K = de * w0**2
        else:
            w0, g, K = 0.0, p[2], p[1] ** 2
        if K != 0.0 and w0 * dt >= 2.0:
            raise ValueError("omega_0 * dt >= 2")
        D = 1.0 + 0.5 * g * dt
        c1[i] = (2.0 - w0**2 * dt**2) / D
        c2[i] = -(1.0 - 0.5 * g * dt) / D
        u = _orientation(p)
        if tensor:
            c3[i] = ((K * dt**2 / D) * np.outer(u, u)).reshape(-1) if u is not None else \
                (K * dt**2 / D) * np.eye(3).reshape(-1)
        else:
            if u is not None:
                raise ValueError("an oriented pole needs tensor coefficients")
            c3[i] = (K * dt**2) / D
    return c1, c2, c3


def build_ade(spec, dtype, dt, ncomp_eps):
    boxes = [b for b in spec.get("boxes", []) if b.get("poles")]
    if not boxes:
        return None
    P = max(len(b["poles"]) for b in spec["boxes"] if b.get("poles"))
    shape = tuple(spec["shape"])
    f = dtype
    # oriented poles widen the coupling c3 to row-major 3x3 tensors (initialization.py:686-711)
    tensor = any(_orientation(p) is not None for b in boxes for p in b["poles"])
    c1 = np.zeros((P, 1, *shape), f); c2 = np.zeros_like(c1)
    c3 = np.zeros((P, 9 if tensor else 1, *shape), f)
    for b in spec.get("boxes", []):  # later boxes overwrite (initialization.py:948-999)
        slc = (slice(None), slice(None), *tuple(slice(*r) for r in b["slc"]))
        k1, k2, k3 = ade_coefficients(b.get("poles", []), dt, tensor)
        v1 = np.zeros(P); v2 = np.zeros(P); v3 = np.zeros((P, c3.shape[1]))
        v1[: len(k1)], v2[: len(k2)] = k1, k2
        v3[: len(k3)] = k3.reshape(len(k3), -1)
        c1[slc] = v1.astype(f)[:, None, None, None, None]
        c2[slc] = v2.astype(f)[:, None, None, None, None]
        c3[slc] = v3.astype(f)[:, :, None, None, None]
    return c1, c2, c3


def pole_parameters(c1, c2, c3):
    """dispersion.py:1033-1058: gamma*dt, omega0^2*dt^2 and a*dt^2 recovered from the ADE
    coefficients (D = 1 + gamma*dt/2), zero where no pole acts."""
    mask = (c1 != 0) | (c3 != 0)
    om = 1.0 - c2
    gdt = np.where(mask, 2.0 * (1.0 + c2) / np.where(om == 0.0, 1.0, om), 0.0)
    D = 1.0 + 0.5 * gdt
    return mask, gdt, np.where(mask, 2.0 - c1 * D, 0.0), np.where(mask, c3 * D, 0.0)


def chi_of(c1, c2, c3, omega, dt, complex_dtype=np.complex64):
    """dispersion.py:988-1066: chi(omega) = sum_p a / (omega0^2 - omega^2 - i gamma omega), in
    dt-scaled form."""
    real = np.float32 if complex_dtype == np.complex64 else np.float64
    c1, c2, c3 = (np.asarray(x, real) for x in (c1, c2, c3))
    mask, gdt, w0, a = pole_parameters(c1, c2, c3)
    wdt = real(omega * dt)
    den = (w0.astype(real) - wdt * wdt).astype(complex_dtype) - complex_dtype(1j) * (gdt.astype(real) * wdt)
    chi = np.where(mask, a.astype(complex_dtype) / np.where(mask, den, complex_dtype(1)), complex_dtype(0))
    return chi.sum(axis=0)


def effective_inv_eps(ie, ade_slice, omega, dt):
    """dispersion.py:1254-1307: 1 / (eps_inf + Re chi(omega)) in float32."""
    f = np.float32
    ie = np.asarray(ie, f)
    if ade_slice is None:
        return ie
    chi = chi_of(*ade_slice, omega, dt)
    return (f(1) / (f(1) / ie + np.real(chi).astype(f))).astype(f)


def broadband_H_profile(prof, dt, T, ade_slice, ie_inf):
    """tfsf.py:21-140 + dispersion.py:1067-1251: the Gaussian profile sampled at every step and
    filtered by sqrt(eps(omega)/eps(omega_c)) (eps averaged over the source cells; DC gain 1;
    real Nyquist bin) through a zero-padded FFT of length the power of two >= 2T."""
    c1, c2, c3 = (np.asarray(x, np.float64) for x in ade_slice)
    if not np.any(c3):
        return None
    raw = O.gaussian_pulse(np.arange(T, dtype=np.float64) * dt, prof["center_frequency"], prof["spectral_width"],
                           dtype=np.float32).astype(np.float64)
    m = 1
    while m < 2 * T:
        m *= 2
    mask, gdt, w0, a = pole_parameters(c1, c2, c3)
    eps_inf = np.mean(1.0 / np.asarray(ie_inf, np.float64), axis=0)

    def eps_mean(omegas):
        out = []
        for w in np.asarray(omegas, np.float64):
            wdt = w * dt
            chi = np.where(mask, a / np.where(mask, w0 - wdt * wdt - 1j * gdt * wdt, 1.0), 0.0)
            out.append(np.mean(eps_inf + chi.sum(axis=0).mean(axis=0)))
        return np.asarray(out)

    ratio = eps_mean(2 * np.pi * np.fft.rfftfreq(m, d=dt)) / eps_mean([2 * np.pi * prof["center_frequency"]])[0]
    g = np.sqrt(ratio.astype(np.complex128))
    g[0] = 1.0
    g[-1] = np.real(g[-1])
    x = np.zeros(m)
    x[:T] = raw
    return np.fft.irfft(np.fft.rfft(x) * g, n=m)[:T].astype(np.float32)


def plane_source(spec_src, scene, T, inv_eps, dispersive=True, broadband=True):
    """UniformPlaneSource at normal incidence: linear_polarization.py:168-358, 471-482 with
    _source_impedance (tfsf.py:144-190) and calculate_time_offset_yee (core/grid.py:720-880)."""
    f = scene.dtype
    ax = spec_src["axis"]
    idx = spec_src["index"]
    direction = spec_src.get("direction", "+")
    shape = scene.shape
    slc = [slice(None)] * 3
    slc[ax] = slice(idx, idx + 1)
    slc = tuple(slc)
    ie_inf = inv_eps[(slice(None), *slc)]
    ade_slc = None
    if scene.ade is not None and dispersive:
        ade_slc = tuple(c[(slice(None), slice(None), *slc)] for c in scene.ade)
        if not (np.any(ade_slc[0]) or np.any(ade_slc[2])):
            ade_slc = None
    prof = spec_src["profile"]
    ie = effective_inv_eps(ie_inf, ade_slc, 2 * np.pi * prof["center_frequency"], scene.dt).astype(np.float64)
    e_pol = np.asarray(spec_src["e_pol"], np.float64)
    e_pol = e_pol / np.linalg.norm(e_pol)
    k = np.zeros(3)
    k[ax] = 1.0 if direction == "+" else -1.0
    h_pol = np.cross(k, e_pol)
    plane_shape = tuple(1 if a == ax else shape[a] for a in range(3))
    amp = np.full(plane_shape, spec_src.get("amplitude", 1.0))
    E = amp[None] * e_pol[:, None, None, None]
    H = amp[None] * h_pol[:, None, None, None]
    ie_b = np.broadcast_to(ie, (3, *plane_shape)) if ie.shape[0] == 3 else np.broadcast_to(ie, (3, *plane_shape))
    energy = 0.5 * (1 / ie_b) * E**2
    energy = energy.sum(0) + (0.5 * H**2).sum(0)
    # the full-domain plane's energy: the reduced sum times 2 per symmetry plane it straddles
    # (tfsf.py:559-569, linear_polarization.py:295-306)
    root = math.sqrt(float(energy.sum()) * spec_src.get("multiplicity", 1))
    E = E / root
    H = H / root
    H = H / np.sqrt(ie[0:1] / 1.0)
    # Yee time offsets (normal incidence, local coordinates start at 0 on the plane cell)
    n_idx = 1 / np.sqrt(ie[0] * 1.0)
    vel = O.C0 / n_idx
    off_E = np.zeros((3,), np.float64)
    off_H = np.zeros((3,), np.float64)
    half = 0.5 * scene.dx
    # E_a at offset 0.5 along axis a; H_a at 0.5 along the two other axes
    kk = k[ax]
    for comp in range(3):
        pos_E = half if comp == ax else 0.0
        pos_H = 0.0 if comp == ax else half
        off_E[comp] = -(pos_E * kk)
        off_H[comp] = -(pos_H * kk)
    vel = float(np.mean(vel))
    tau_E = off_E / (vel * scene.dt)
    tau_H = off_H / (vel * scene.dt)
    a_ax, b_ax = (ax + 1) % 3, (ax + 2) % 3
    filt = None if ade_slc is None or not broadband else broadband_H_profile(prof, scene.dt, T, ade_slc, ie_inf)
    sign = 1 if direction == "+" else -1
    n = np.arange(T, dtype=f)

    def amp_table(half_step, tau):
        base = (n + f(0.5)).astype(f) if half_step else n
        t = ((base + f(tau)).astype(f) * f(scene.dt)).astype(f)
        return O.gaussian_pulse(t, prof["center_frequency"], prof["spectral_width"], dtype=f) * f(
            spec_src.get("static_amplitude_factor", 1.0))

    def filtered_table(tau):
        idx = (n + f(tau)).astype(f).astype(np.float64)
        return (np.interp(idx, np.arange(T, dtype=np.float64), filt.astype(np.float64), left=0.0, right=0.0)
                .astype(f) * f(spec_src.get("static_amplitude_factor", 1.0)))

    if filt is None:
        ampH = np.stack([amp_table(False, tau_H[c]) for c in range(3)]).astype(f)
    else:  # the E-side correction injects the filtered incident H (tfsf.py:258-263)
        ampH = np.stack([filtered_table(tau_H[c]) for c in range(3)]).astype(f)
    ampE = np.stack([amp_table(True, tau_E[c]) for c in range(3)]).astype(f)
    multE = {c: (inv_eps[(c if inv_eps.shape[0] == 3 else 0, *slc)]) for c in range(3)}
    im = f(1.0)
    multH = {c: im for c in range(3)}
    srcE = dict(kind="plane", field="E", slc=slc, sign=sign, profile=H.astype(f), amp=ampH,
                terms=[(a_ax, b_ax, 1), (b_ax, a_ax, -1)], mult=multE)
    srcH = dict(kind="plane", field="H", slc=slc, sign=sign, profile=E.astype(f), amp=ampE,
                terms=[(b_ax, a_ax, 1), (a_ax, b_ax, -1)], mult=multH)
    return [srcE, srcH], dict(E=E, H=H, tau_E=tau_E, tau_H=tau_H)


def mode_source(spec_src, scene, T, inv_eps, dispersive=True):
    """ModePlaneSource (mode.py:98-298): the TFSF plane tables of plane_source with the incident
    cross-section given by a mode, spec_src["mode"] = (E, H, neff) over the plane cells (already
    normalised to unit flux; the real parts are injected, mode.py:207-213) and the Yee time
    offsets at the phase velocity c0 / Re(neff) (core/grid.py:720-880 with effective_index)."""
    f = scene.dtype
    ax = spec_src["axis"]
    idx = spec_src["index"]
    direction = spec_src.get("direction", "+")
    slc = [slice(None)] * 3
    slc[ax] = slice(idx, idx + 1)
    slc = tuple(slc)
    mE, mH, neff = spec_src["mode"]
    E = np.real(np.asarray(mE)).astype(np.float64)
    H = np.real(np.asarray(mH)).astype(np.float64)
    ie_inf = inv_eps[(slice(None), *slc)]
    ade_slc = None
    if scene.ade is not None and dispersive:
        ade_slc = tuple(c[(slice(None), slice(None), *slc)] for c in scene.ade)
        if not (np.any(ade_slc[0]) or np.any(ade_slc[2])):
            ade_slc = None
    prof = spec_src["profile"]
    vel = O.C0 / float(np.real(neff))
    kk = 1.0 if direction == "+" else -1.0
    half = 0.5 * scene.dx
    tau_E = np.array([-(half * kk if c == ax else 0.0) for c in range(3)]) / (vel * scene.dt)
    tau_H = np.array([-(0.0 if c == ax else half * kk) for c in range(3)]) / (vel * scene.dt)
    a_ax, b_ax = (ax + 1) % 3, (ax + 2) % 3
    filt = None if ade_slc is None else broadband_H_profile(prof, scene.dt, T, ade_slc, ie_inf)
    sign = 1 if direction == "+" else -1
    n = np.arange(T, dtype=f)
    saf = f(spec_src.get("static_amplitude_factor", 1.0))

    def amp_table(half_step, tau):
        base = (n + f(0.5)).astype(f) if half_step else n
        t = ((base + f(tau)).astype(f) * f(scene.dt)).astype(f)
        return O.gaussian_pulse(t, prof["center_frequency"], prof["spectral_width"], dtype=f) * saf

    if filt is None:
        ampH = np.stack([amp_table(False, tau_H[c]) for c in range(3)]).astype(f)
    else:
        ampH = np.stack([(np.interp((n + f(tau_H[c])).astype(f).astype(np.float64), np.arange(T, dtype=np.float64),
                                    filt.astype(np.float64), left=0.0, right=0.0).astype(f) * saf)
                         for c in range(3)]).astype(f)
    ampE = np.stack([amp_table(True, tau_E[c]) for c in range(3)]).astype(f)
    multE = {c: (inv_eps[(c if inv_eps.shape[0] == 3 else 0, *slc)]) for c in range(3)}
    multH = {c: f(1.0) for c in range(3)}
    srcE = dict(kind="plane", field="E", slc=slc, sign=sign, profile=H.astype(f), amp=ampH,
                terms=[(a_ax, b_ax, 1), (b_ax, a_ax, -1)], mult=multE)
    srcH = dict(kind="plane", field="H", slc=slc, sign=sign, profile=E.astype(f), amp=ampE,
                terms=[(b_ax, a_ax, 1), (a_ax, b_ax, -1)], mult=multH)
    return [srcE, srcH], dict(E=E, H=H, tau_E=tau_E, tau_H=tau_H)


def _expand_detectors(dets):
    """Spec detectors as the oracle runs them: a "phasor_flux" detector is a phasor detector of all
    six components (its flux is formed after the run); a "closed_phasor" detector is one such
    phasor detector per boundary plane of each active axis (min, then max), named
    "{name}:phasor_axis{a}_{side}", whose steps take no apodization weight
    (poynting_flux.py:478-505)."""
    out = []
    six = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for d in dets:
        if d["kind"] in ("phasor_flux", "mode_gaussian", "mode"):
            out.append(dict(d, kind="phasor", components=six))
        elif d["kind"] == "closed_phasor":
            shp = [h - l for l, h in d["slc"]]
            axes = tuple(d["axes"]) if d.get("axes") is not None else tuple(a for a in range(3) if shp[a] > 1)
            for a in axes:
                for side in ("min", "max"):
                    slc = [tuple(r) for r in d["slc"]]
                    lo, hi = slc[a]
                    slc[a] = (hi - 1, hi) if side == "max" else (lo, lo + 1)
                    out.append(dict(d, kind="phasor", components=six, slc=tuple(slc), face=(a, side),
                                    name=f"{d['name']}:phasor_axis{a}_{side}", unweighted_steps=True))
        elif d["kind"] == "energy" and d.get("as_slices"):  # energy.py:76-132: three planes
            mean = d.get("aggregate") == "mean" or any(d.get(f"{c}_slice_index") is None for c in "xyz")
            for key, ax in (("XY Plane", 2), ("XZ Plane", 1), ("YZ Plane", 0)):
                if mean:
                    out.append(dict(d, kind="energy_mean", axis=ax, name=f"{d['name']}:{key}"))
                else:
                    slc = [tuple(r) for r in d["slc"]]
                    i0 = slc[ax][0] + d["xyz"[ax] + "_slice_index"]
                    slc[ax] = (i0, i0 + 1)
                    out.append(dict(d, slc=tuple(slc), reduce_volume=False, name=f"{d['name']}:{key}"))
        else:
            out.append(d)
    return out


def apodization_window(apo, T, dt):
    """TemporalWindow.get_window (core/window.py:16-55, 75-120) at t = n dt, float32.
    apo: dict(kind="gaussian", center_time, sigma_time) or dict(kind="tukey", start_time,
    end_time, alpha)."""
    f = np.float32
    t = (np.arange(T, dtype=f) * f(dt)).astype(f)
    if apo["kind"] == "gaussian":
        u = (t - f(apo["center_time"])).astype(f)
        return np.exp(-(u * u) / f(2.0 * apo["sigma_time"] ** 2)).astype(f)
    a = apo.get("alpha", 0.5)
    x = ((t - f(apo["start_time"])) / f(apo["end_time"] - apo["start_time"])).astype(f)
    inside = (x >= 0) & (x <= 1)
    if a <= 0:
        return inside.astype(f)
    h = a / 2.0
    up = f(0.5) * (f(1) + np.cos(f(np.pi) * (x / f(h) - f(1))))
    down = f(0.5) * (f(1) + np.cos(f(np.pi) * ((x - f(1)) / f(h) + f(1))))
    w = np.where(x < h, up, np.where(x > 1 - h, down, f(1)))
    return np.where(inside, w, f(0)).astype(f)


def reduce_spec(spec):
    """Mirror-symmetry reduction (fdtd/symmetry.py:84-264) restated on a scenario spec: along every
    axis a with symmetry[a] != 0 (-1 electric, +1 magnetic) the upper half of the (even) extent n is
    kept, everything is shifted by m = n/2 and clipped to it, objects wholly below are dropped, the
    min-face CPML is dropped, an electric plane gets a PEC face at index 0 and a magnetic one
    nothing (its mirror is the zero halo half a cell below). A plane source gains the multiplicity
    2 per transverse axis it straddles; detectors record the axes they straddle; a dipole in the
    first kept cell, or any periodic face on a symmetric axis, is rejected."""
    sym = tuple(spec.get("symmetry", (0, 0, 0)))
    if not any(sym):
        return spec
    shape = list(spec["shape"])
    mid = [0, 0, 0]
    for a in range(3):
        if sym[a]:
            if shape[a] % 2:
                raise ValueError(f"symmetric axis {a} needs an even number of cells, got {shape[a]}")
            mid[a] = shape[a] // 2
    full = list(spec["shape"])

    def clip(slc):
        out, straddle = [], []
        for a, (s0, s1) in enumerate(slc):
            if sym[a]:
                n0, n1 = max(s0, mid[a]) - mid[a], min(s1, full[a]) - mid[a]
                if n1 <= n0:
                    return None, None
                out.append((n0, n1))
                straddle.append(s0 < mid[a])
            else:
                out.append((s0, s1))
                straddle.append(False)
        return tuple(out), tuple(straddle)

    new = dict(spec)
    new.pop("symmetry")
    new["shape"] = tuple(full[a] - mid[a] for a in range(3))
    boxes = []
    for b in spec.get("boxes", []):
        c, _ = clip(b["slc"])
        if c is not None:
            boxes.append(dict(b, slc=c))
    new["boxes"] = boxes
    d = spec.get("design")
    if d is not None:
        c, _ = clip(d["slc"])
        if c is None:
            new["design"] = None
        else:
            rho = np.asarray(d["rho"])
            cut = tuple(slice(max(s0, mid[a]) - s0 if sym[a] else 0, None) for a, (s0, _) in enumerate(d["slc"]))
            new["design"] = dict(d, slc=c, rho=rho[cut])
    srcs = []
    for s_ in spec.get("sources", []):
        if s_["kind"] == "dipole":
            cell = list(s_["cell"])
            keep = True
            for a in range(3):
                if sym[a]:
                    cell[a] -= mid[a]
                    keep = keep and cell[a] >= 0
                    if cell[a] == 0:
                        raise ValueError("a point dipole may not sit on a symmetry plane (dipole.py:110-128)")
            if keep:
                srcs.append(dict(s_, cell=tuple(cell)))
        else:
            ax = s_["axis"]
            s2 = dict(s_)
            if sym[ax]:
                s2["index"] = s_["index"] - mid[ax]
                if s2["index"] < 0:
                    continue
            s2["multiplicity"] = 2 ** sum(1 for a in range(3) if a != ax and sym[a])
            srcs.append(s2)
    new["sources"] = srcs
    dets = []
    for dd in spec.get("detectors", []):
        c, st = clip(dd["slc"])
        if c is not None:
            dets.append(dict(dd, slc=c, straddle=st))
    new["detectors"] = dets
    pml = dict(spec.get("pml", {"thickness": 0}))
    types = dict(pml.get("types", {}))
    for a in range(3):
        if not sym[a]:
            continue
        if types.get(f"min_{'xyz'[a]}") == "periodic" or types.get(f"max_{'xyz'[a]}") == "periodic":
            raise ValueError("periodic boundaries and mirror symmetry exclude each other on one axis")
        types[f"min_{'xyz'[a]}"] = "pec" if sym[a] == -1 else "none"
    pml["types"] = types
    new["pml"] = pml
    new["mirror"] = tuple(s_ == -1 for s_ in sym)
    return new


def build_scene(spec, dtype=np.float32):
    spec = reduce_spec(spec)
    shape = tuple(spec["shape"])
    dx = spec["dx"]
    T = spec["T"]
    sc = O.Scene(shape=shape, dx=dx, dtype=dtype, courant_factor=spec.get("courant_factor", 0.99))
    dt = sc.dt
    sc.inv_eps = build_inv_eps(spec, dtype)
    sc.inv_mu = build_inv_mu(spec, dtype)
    sc.sigma_E = build_sigma(spec, dtype, "sigma_E", dt, sc.c)
    sc.sigma_H = build_sigma(spec, dtype, "sigma_H", dt, sc.c)
    pml = spec.get("pml", {"thickness": 0})
    faces = ["min_x", "max_x", "min_y", "max_y", "min_z", "max_z"]
    types = pml.get("types", {})
    if pml.get("thickness", 0) or types:
        th = [pml.get("thickness", 0) if types.get(fc, "pml") == "pml" else 0 for fc in faces]
        sc.slabs = O.make_slabs(shape, th, dt, dx, dtype, **pml.get("params", {}))
    for fc in faces:
        kind = types.get(fc, "pml")
        if kind in ("pec", "pmc"):
            axis = "xyz".index(fc[-1])
            sc.walls.append((kind, axis, 0 if fc.startswith("min") else shape[axis] - 1))
        elif kind in ("periodic", "bloch"):
            a = "xyz".index(fc[-1])
            per = list(sc.periodic)
            per[a] = True
            sc.periodic = tuple(per)
            if kind == "bloch":  # get_bloch_phase (bloch.py:140-153): k_a * L_a, L_a = n_a * dx
                k = float(spec.get("bloch_vector", (0.0, 0.0, 0.0))[a])
                b = list(sc.bloch)
                th = k * shape[a] * dx
                old = b[a] if b[a] is not None else (0.0, 0.0)
                b[a] = (th, old[1]) if fc.startswith("min") else (old[0], th)
                sc.bloch = tuple(b)
        elif kind not in ("pml", "none"):
            raise NotImplementedError(f"oracle: boundary type {kind!r}")
    # complex fields: requested, or implied by a nonzero Bloch phase (initialization.py:580-594)
    sc.complex_fields = bool(spec.get("complex_fields", False)) or any(
        b is not None and (b[0] != 0.0 or b[1] != 0.0) for b in sc.bloch)
    sc.mirror = tuple(spec.get("mirror", (False, False, False)))
    sc.ade = build_ade(spec, dtype, dt, sc.inv_eps.shape[0])
    for s in spec.get("sources", []):
        prof = s["profile"]
        if s["kind"] == "dipole":
            cell = tuple(s["cell"])
            p = s["polarization"]
            ie_cell = sc.inv_eps[(slice(None), *tuple(slice(c, c + 1) for c in cell))]
            ade_cell = None
            if sc.ade is not None:
                ade_cell = tuple(c[(slice(None), slice(None), *tuple(slice(x, x + 1) for x in cell))] for c in sc.ade)
                if not (np.any(ade_cell[0]) or np.any(ade_cell[2])):
                    ade_cell = None
            if sc.inv_eps.shape[0] == 9:  # dipole.py:15-37, 219-234: every row of column p
                for axis in range(3):
                    tab = O.dipole_table(T, dt, sc.c, prof["center_frequency"], prof["spectral_width"],
                                         sc.inv_eps[(3 * axis + p, *cell)], amplitude=s.get("amplitude", 1.0),
                                         dtype=dtype)
                    sc.sources.append(dict(kind="dipole", field="E", cell=cell, component=axis, table=tab))
                continue
            single = prof.get("kind") == "single"
            fc = prof["frequency"] if single else prof["center_frequency"]
            ie_local = effective_inv_eps(ie_cell, ade_cell, 2 * np.pi * fc, dt)[
                (p if sc.inv_eps.shape[0] == 3 else 0, 0, 0, 0)]  # dipole.py:163-191
            if single:  # SingleFrequencyProfile with the wave character's period and phase shift
                tab = O.dipole_table(T, dt, sc.c, None, None, ie_local, amplitude=s.get("amplitude", 1.0),
                                     dtype=dtype, profile="single", period=1.0 / fc,
                                     phase_shift=prof.get("phase_shift", 0.0))
            else:
                tab = O.dipole_table(T, dt, sc.c, prof["center_frequency"], prof["spectral_width"], ie_local,
                                     amplitude=s.get("amplitude", 1.0), dtype=dtype)
            sc.sources.append(dict(kind="dipole", field="E", cell=cell, component=p, table=tab))
        elif s["kind"] == "plane":
            srcs, _ = plane_source(s, sc, T, sc.inv_eps)
            sc.sources.extend(srcs)
        elif s["kind"] == "mode":
            srcs, _ = mode_source(s, sc, T, sc.inv_eps)
            sc.sources.extend(srcs)
    for d in _expand_detectors(spec.get("detectors", [])):
        det = dict(d)
        det["slc"] = tuple(slice(*r) for r in d["slc"])
        det["time_to_idx"] = time_to_idx(on_list(T, dt, d.get("switch")))
        if d["kind"] in ("energy", "energy_mean"):
            det["weight"] = dx**3
        if d["kind"] == "flux":
            det["weight"] = dx**2
            det["sign"] = 1 if d.get("direction", "+") == "+" else -1
        if d["kind"] == "closed_flux":
            det["weight"] = dx**2
            shp = [h - l for l, h in d["slc"]]
            det["axes"] = tuple(d["axes"]) if d.get("axes") is not None else tuple(a for a in range(3) if shp[a] > 1)
        if d["kind"] == "flux" and d.get("fixed_propagation_axis") is not None:
            det["axis"] = d["fixed_propagation_axis"]
        if d["kind"] == "phasor":
            om = [2 * math.pi * fr for fr in d["frequencies"]]
            lst = on_list(T, dt, d.get("switch"))
            stride = d.get("dft_subsample", 1)
            if stride > 1:  # every stride-th active step (phasor.py:105-116)
                act = [t for t, o in enumerate(lst) if o]
                lst = [False] * T
                for t in act[::stride]:
                    lst[t] = True
            win = np.asarray(lst, np.float32)
            if d.get("apodization") is not None:
                win = (apodization_window(d["apodization"], T, dt) * win).astype(np.float32)
            cos_t, sin_t, w, scale = O.phasor_tables(om, T, dt, window=win, dtype=dtype,
                                                     scaling_mode=d.get("scaling_mode", "continuous"))
            if d.get("unweighted_steps"):
                w = np.asarray(lst, np.float32)
            if d.get("scaling_mode") == "pulse":
                scale = stride
            det.update(cos=cos_t, sin=sin_t, window=w, scale=scale)
            det["time_to_idx"] = np.where(np.asarray(lst), 0, -1).astype(np.int32)
        sc.detectors.append(det)
    return sc
