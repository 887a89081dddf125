"""Pins the CPU oracle (oracle/fdtd_oracle.py) before it is trusted as the parity reference.

Each check is chosen so that a plausible mistake in the restatement (dropped term, wrong sign or
index, transposed operand) fails: the reference's own known-answer tests (tests/unit/core/physics/
test_curl.py:162-183, 339-360), closed forms, physical invariants (CPML absorption, time reversal,
Lorentz susceptibility of the ADE recurrence), a dot-product test of the discrete adjoint and FP64
central finite differences of a full solve.
"""

import math

import numpy as np
import pytest

from fdtdx_amd import scenarios as W
from oracle import fdtd_oracle as O
from oracle import scenes as S
from scenarios import (PERIODIC_XY, TENSOR_EPS, TENSOR_MU, detector_zoo_losses, detector_zoo_spec,
                       energy_slices_spec, full_spec, periodic_spec, surface_zoo_losses, surface_zoo_spec,
                       symmetry_spec)


# ---------------------------------------------------------------- curl stencils
def test_curl_E_linear_field_kat():
    """test_curl.py:162-183: E = (y, -x, 0) gives curl_z = -2 away from the upper faces."""
    n = 6
    X, Y, _ = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    E = np.stack([Y, -X, np.zeros_like(X)]).astype(np.float64)
    curl, _ = O.curl_E_raw(E)
    assert np.allclose(curl[2][:-1, :-1], -2.0)
    inner = (slice(0, -1),) * 3
    assert np.allclose(curl[0][inner], 0.0) and np.allclose(curl[1][inner], 0.0)


def test_curl_H_linear_field_kat():
    """test_curl.py:339-360: H = (z, 0, -x) gives curl_y = 2 away from the lower faces."""
    n = 6
    X, _, Z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    H = np.stack([Z, np.zeros_like(X), -X]).astype(np.float64)
    curl, _ = O.curl_H_raw(H)
    assert np.allclose(curl[1][1:, :, 1:], 2.0)


def test_curl_integer_fields_exact_and_dual():
    """Integer fields: FP32 curls are exact; and <curlH a, b> = <a, curlE b> (Yee duality with the
    zero halo of pad_fields)."""
    rng = np.random.default_rng(0)
    a = rng.integers(-50, 50, size=(3, 7, 6, 5)).astype(np.float64)
    b = rng.integers(-50, 50, size=(3, 7, 6, 5)).astype(np.float64)
    c64, _ = O.curl_H_raw(a)
    c32, _ = O.curl_H_raw(a.astype(np.float32))
    assert np.array_equal(np.stack(c64), np.stack(c32).astype(np.float64))
    lhs = np.sum(np.stack(O.curl_H_raw(a)[0]) * b)
    rhs = np.sum(a * np.stack(O.curl_E_raw(b)[0]))
    assert lhs == rhs


# ---------------------------------------------------------------- CPML coefficients
def test_cpml_profile_closed_forms():
    dx = 50e-9
    dt = O.time_step_duration(dx)
    L = 10
    for side in ("-", "+"):
        bE, aE, ikE, bH, aH, ikH = O.cpml_profiles(L, side, dt, dx, np.float64)
        alpha_s = 0.01 * 2 * math.pi * O.C0 / 1.55e-6 * O.EPS0
        # interface cell has sigma = 0: a = 0 and b = exp(-dt/eps0 * alpha_start)
        iface = L - 1 if side == "-" else 0
        assert aE[iface] == 0.0 and aH[iface] == 0.0
        assert math.isclose(bE[iface], math.exp(-dt / O.EPS0 * alpha_s), rel_tol=1e-12)
        # outermost cell: sigma = sigma_max (d/L)^3 with the reference's sigma_end
        sig_end = -4 * math.log(1e-6) / (2 * O.ETA0 * L * dx)
        d = L - 1 if side == "-" else L - 1.5
        outer = 0 if side == "-" else L - 1
        sig = sig_end * (d / L) ** 3
        alp = alpha_s * (1 - d / L)
        b_ref = math.exp(-dt / O.EPS0 * (sig + alp))
        assert math.isclose(bE[outer], b_ref, rel_tol=1e-12)
        assert math.isclose(aE[outer], (b_ref - 1) * sig / (sig + alp), rel_tol=1e-12)
        assert np.all(ikE == 1.0) and np.all(ikH == 1.0)
        # grading: sigma grows towards the outer face
        order = np.argsort(np.abs(aE))
        assert (order[-1] == outer)


def test_cpml_noop_without_loss():
    b1, a1, *_ = O.cpml_profiles(6, "-", 1e-16, 50e-9, np.float64, sigma_end=0.0, alpha_start=0.0)
    assert np.all(b1 == 1.0) and np.all(a1 == 0.0)


def test_cpml_absorbs_outgoing_pulse():
    """Physics: a dipole pulse radiates into a CPML-lined box; after it has left, < 0.1 % of the
    peak electromagnetic energy remains (reflections from the CPML would keep it in the box)."""
    spec = W.scenario_c1(T=700, n=30, pml=8, cube=4)
    spec["boxes"] = []
    sc = S.build_scene(spec, np.float64)
    _, ds = O.run(sc, spec["T"])
    en = ds[0][:, 0]
    assert en[-1] < 1e-3 * en.max()


# ---------------------------------------------------------------- sources
def test_gaussian_pulse_closed_form():
    f0 = O.C0 / 1.55e-6
    t = np.linspace(0, 60e-15, 257)
    a = O.gaussian_pulse(t, f0, 0.2 * f0, dtype=np.float64)
    st = 1 / (2 * math.pi * 0.2 * f0)
    ref = np.exp(-((t - 6 * st) ** 2) / (2 * st**2)) * np.cos(2 * math.pi * f0 * t)
    assert np.allclose(a, ref, rtol=0, atol=1e-12)
    a32 = O.gaussian_pulse(t.astype(np.float32), f0, 0.2 * f0)
    assert np.max(np.abs(a32 - ref)) < 1e-4


def test_first_step_closed_form():
    """One step from zero fields: E holds only the dipole increment and H = -c * curl_E(E)."""
    spec = W.scenario_c1(T=3, n=12, pml=0, cube=2)
    sc = S.build_scene(spec, np.float64)
    st = O.init_state(sc)
    O.forward_step(st, sc, 0)
    src = sc.sources[0]
    i, j, k = src["cell"]
    assert st.E[2, i, j, k] == src["table"][0]
    assert np.count_nonzero(st.E) == (1 if src["table"][0] != 0 else 0)
    curl, _ = O.curl_E_raw(st.E)
    assert np.array_equal(st.H, -(sc.c * np.stack(curl)))


# ---------------------------------------------------------------- ADE
def test_ade_recurrence_reproduces_lorentz_susceptibility():
    """The discrete recurrence P^{n+1} = c1 P^n + c2 P^{n-1} + c3 E^n driven by E^n = z^n has
    chi_d = c3 / (z - c1 - c2/z); for omega*dt << 1 it must approach the Lorentz/Drude chi."""
    dt = O.time_step_duration(50e-9)
    w0, de = 2 * math.pi * O.C0 / 1.2e-6, 1.5
    g = 0.05 * w0
    wp, gd = 2 * math.pi * O.C0 / 0.3e-6, 1e14
    (c1, c2, c3) = S.ade_coefficients([("lorentz", w0, g, de), ("drude", wp, gd)], dt)
    for w in (0.5 * w0, 0.9 * w0, 1.3 * w0):
        z = np.exp(1j * w * dt)
        chi_d = c3 / (z - c1 - c2 / z)
        chi_l = de * w0**2 / (w0**2 - w**2 + 1j * g * w)
        chi_dr = -(wp**2) / (w**2 - 1j * gd * w)
        assert abs(chi_d[0] - chi_l) < 0.02 * abs(chi_l)
        assert abs(chi_d[1] - chi_dr) < 0.02 * abs(chi_dr)


# ---------------------------------------------------------------- detectors
def test_interpolation_matches_brute_force():
    rng = np.random.default_rng(2)
    E = rng.standard_normal((3, 5, 4, 6))
    H = rng.standard_normal((3, 5, 4, 6))
    Ei, Hi = O.interpolate_fields(E, H)

    def g(F, c, i, j, k):
        if 0 <= i < 5 and 0 <= j < 4 and 0 <= k < 6:
            return F[c, i, j, k]
        return 0.0

    for (i, j, k) in [(0, 0, 0), (2, 1, 3), (4, 3, 5), (3, 0, 5)]:
        ex = 0.25 * (g(E, 0, i, j, k) + g(E, 0, i - 1, j, k) + g(E, 0, i, j, k + 1) + g(E, 0, i - 1, j, k + 1))
        hz = sum(g(H, 2, i - a, j - b, k + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)) / 8
        assert math.isclose(Ei[0, i, j, k], ex, rel_tol=1e-12, abs_tol=1e-15)
        assert math.isclose(Hi[2, i, j, k], hz, rel_tol=1e-12, abs_tol=1e-15)
        assert Ei[2, i, j, k] == E[2, i, j, k]


def test_phasor_of_pure_tone():
    """phasor.py: sum_n A cos(w n dt + phi) exp(i w n dt) * 2/N -> A exp(-i phi) over whole periods."""
    N, per = 400, 40
    dt = 1.0
    w = 2 * math.pi / per
    cos_t, sin_t, win, scale = O.phasor_tables([w], N, dt, dtype=np.float64)
    A, phi = 1.7, 0.4
    sig = A * np.cos(w * np.arange(N) * dt + phi)
    acc = np.sum(sig * (cos_t[0] + 1j * sin_t[0])) * scale
    assert abs(acc - A * np.exp(-1j * phi)) < 1e-9


def test_apodization_windows_closed_forms():
    """window.py:16-55: the Gaussian window is 1 at its centre and exp(-1/2) one sigma away; the
    Tukey window is the indicator of [start, end] for alpha = 0, the Hann window
    (1 - cos 2 pi x) / 2 for alpha = 1, and 1 on its flat top. Both restatements (the oracle's
    and the host's) are held to the closed forms."""
    from fdtdx_amd.core import GaussianWindow, TukeyWindow

    T, dt = 400, 1e-16
    t = np.arange(T) * dt
    x = (t - 50 * dt) / (300 * dt)
    hann = np.where((x >= 0) & (x <= 1), 0.5 * (1 - np.cos(2 * np.pi * x)), 0.0)
    cases = [(dict(kind="gaussian", center_time=200 * dt, sigma_time=30 * dt), np.exp(-((t - 200 * dt) ** 2) /
                                                                                     (2 * (30 * dt) ** 2))),
             (dict(kind="tukey", start_time=50 * dt, end_time=350 * dt, alpha=0.0),
              ((x >= 0) & (x <= 1)).astype(float)),
             (dict(kind="tukey", start_time=50 * dt, end_time=350 * dt, alpha=1.0), hann)]
    for apo, want in cases:
        kw = {k: v for k, v in apo.items() if k != "kind"}
        host = (GaussianWindow if apo["kind"] == "gaussian" else TukeyWindow)(**kw)
        for got in (S.apodization_window(apo, T, dt), host.get_window(np.arange(T, dtype=np.float32)
                                                                       * np.float32(dt))):
            assert got.dtype == np.float32
            assert np.max(np.abs(got - want)) < 2e-6, apo
    w = S.apodization_window(dict(kind="tukey", start_time=50 * dt, end_time=350 * dt, alpha=0.5), T, dt)
    assert np.all(w[(x >= 0.25) & (x <= 0.75)] == 1.0) and w[49] == 0.0 and w[351] == 0.0
    g = S.apodization_window(dict(kind="gaussian", center_time=200 * dt, sigma_time=30 * dt), T, dt)
    assert g[200] == 1.0 and abs(g[230] - math.exp(-0.5)) < 1e-6 and abs(g[170] - g[230]) < 1e-6


def test_apodized_phasor_of_pure_tone():
    """phasor.py:139-175: with a Hann window and the coherent-gain scale 2 / sum(w), the phasor of
    A cos(w t + phi) is A exp(-i phi) up to the window's small leakage of the 2w term."""
    N, per, dt = 400, 40, 1.0
    w = 2 * math.pi / per
    win = S.apodization_window(dict(kind="tukey", start_time=0.0, end_time=N - 1.0, alpha=1.0), N, dt)
    cos_t, sin_t, win, scale = O.phasor_tables([w], N, dt, window=win, dtype=np.float64)
    assert math.isclose(scale, 2 / np.sum(win))
    A, phi = 1.7, 0.4
    sig = A * np.cos(w * np.arange(N) * dt + phi)
    acc = np.sum(sig * (cos_t[0] + 1j * sin_t[0]) * win) * scale
    assert abs(acc - A * np.exp(-1j * phi)) < 1e-3


# ---------------------------------------------------------------- detector options and transposes
def _detector_scene(mu):
    spec = W.scenario_c1(T=4, n=10, pml=3, cube=2)
    spec["sources"] = []
    box = dict(slc=((2, 7), (3, 8), (1, 6)), eps=(2.0, 3.0, 4.0))
    if mu:
        box["mu"] = (1.5, 2.5, 3.0)
    spec["boxes"] = spec["boxes"] + [box]
    return S.build_scene(spec, np.float64)


def _detector(sc, kind, slc=((1, 6), (0, 4), (2, 10)), T=4, **kw):
    """A detector dict as oracle/scenes.py builds it, recording steps 0 and 2 (slots 0 and 1)."""
    det = dict(kind=kind, slc=tuple(slice(*r) for r in slc), time_to_idx=np.array([0, -1, 1, -1], np.int32),
               reduce_volume=False, exact_interpolation=True)
    det.update(kw)
    if kind in ("energy", "energy_mean"):
        det["weight"] = 0.37
    if kind == "flux":
        det.setdefault("axis", 1)
        det.setdefault("sign", -1)
        det["weight"] = 0.61
    if kind == "closed_flux":
        det.setdefault("axes", (0, 1, 2))
        det["weight"] = 0.61
    if kind == "phasor":
        rng = np.random.default_rng(3)
        win = rng.uniform(0.2, 1.0, T)
        cos_t, sin_t, w, scale = O.phasor_tables([0.7, 1.9], T, 0.3, window=win, dtype=np.float64)
        det.update(cos=cos_t, sin=sin_t, window=w, scale=scale, time_to_idx=np.array([0, -1, 0, -1], np.int32))
    return det


DETECTOR_CASES = [
    dict(kind="energy"), dict(kind="energy", reduce_volume=True), dict(kind="energy", exact_interpolation=False),
    dict(kind="energy", reduce_volume=True, mu=True),
    dict(kind="field", components=("Ey", "Hz")), dict(kind="field", components=("Ex", "Ez", "Hx"), reduce_volume=True),
    dict(kind="field", components=("Ey", "Hy"), exact_interpolation=False, reduce_volume=True),
    dict(kind="flux"), dict(kind="flux", reduce_volume=True), dict(kind="flux", keep_all_components=True),
    dict(kind="flux", keep_all_components=True, reduce_volume=True, exact_interpolation=False),
    dict(kind="closed_flux"), dict(kind="closed_flux", orientation="inward", exact_interpolation=False, axes=(2, 0)),
    dict(kind="energy_mean", axis=1), dict(kind="energy_mean", axis=0, exact_interpolation=False, mu=True),
    dict(kind="phasor", components=("Ex", "Hy", "Hz")), dict(kind="phasor", components=("Ez", "Hx"), reduce_volume=True),
    dict(kind="phasor", components=("Ey",), inverse=True, exact_interpolation=False),
]


def _case_id(c):
    return "-".join(f"{k}={v}" if not isinstance(v, tuple) else f"{k}={'+'.join(map(str, v))}" for k, v in c.items())


@pytest.mark.parametrize("case", DETECTOR_CASES, ids=[_case_id(c) for c in DETECTOR_CASES])
def test_detector_cotangent_is_the_transposed_update(case):
    """detector_cotangent against FP64 central differences of L = <cot, detector_update(E, H_prev,
    H)> along a random direction: exact for the linear detectors (field, phasor) and, the update
    being quadratic, also for energy and flux. A complex state entry z enters L as
    Re(conj(cot) z), so its cotangent is dL/dRe z + i dL/dIm z."""
    case = dict(case)
    sc = _detector_scene(case.pop("mu", False))
    det = _detector(sc, **case)
    rng = np.random.default_rng(11)
    shp = (3, *sc.shape)
    E, Hp, H, dE, dHp, dH = (rng.standard_normal(shp) for _ in range(6))
    st0 = O.init_detector_state(det, sc)
    cot = rng.standard_normal(st0.shape)
    if np.iscomplexobj(st0):
        cot = cot + 1j * rng.standard_normal(st0.shape)
    n = 2

    def L(e, hp, h):
        st = O.detector_update(det, st0.copy(), n, e, hp, h, sc)
        return float(np.sum(np.real(np.conj(cot) * st)))

    h = 1e-3
    fd = (L(E + h * dE, Hp + h * dHp, H + h * dH) - L(E - h * dE, Hp - h * dHp, H - h * dH)) / (2 * h)
    vE, vHc, vH1, g_ie = O.detector_cotangent(det, cot, n, E, Hp, H, sc, grad_eps=True)
    adj = np.sum(vE * dE) + np.sum(vHc * (dHp + dH) / 2) + np.sum(vH1 * dH)
    assert abs(fd) > 1e-6
    assert math.isclose(adj, fd, rel_tol=1e-8), (adj, fd)
    # the direct inv_eps term (energy only) against differences along a random inv_eps direction
    assert (g_ie is not None) == (det["kind"] in ("energy", "energy_mean"))
    if g_ie is not None:
        die = rng.standard_normal(sc.inv_eps.shape)
        ie0 = sc.inv_eps.copy()
        vals = []
        for sgn in (1, -1):
            sc.inv_eps = ie0 + sgn * 1e-4 * die
            vals.append(L(E, Hp, H))
        sc.inv_eps = ie0
        fd_ie = (vals[0] - vals[1]) / 2e-4
        assert math.isclose(np.sum(g_ie * die), fd_ie, rel_tol=1e-6), (np.sum(g_ie * die), fd_ie)
    # the step a detector does not record contributes nothing
    z = O.detector_cotangent(det, cot, 1, E, Hp, H, sc)
    assert all(not np.any(v) for v in z)


def test_reduced_detectors_are_box_means_and_sums_of_cell_values():
    """reduce_volume against the per-cell states of the same detector: field and phasor keep the
    box mean of each component (detector.py:95-109 with uniform cells), flux with
    keep_all_components sums each component with the face-area weight, and its propagation-axis
    component equals the plain flux detector."""
    sc = _detector_scene(False)
    rng = np.random.default_rng(4)
    E, Hp, H = (rng.standard_normal((3, *sc.shape)) for _ in range(3))
    for kind, kw in (("field", dict(components=("Ex", "Hy", "Hz"))), ("phasor", dict(components=("Ey", "Hx")))):
        cell = _detector(sc, kind, **kw)
        red = _detector(sc, kind, reduce_volume=True, **kw)
        a = O.detector_update(cell, O.init_detector_state(cell, sc), 2, E, Hp, H, sc)
        b = O.detector_update(red, O.init_detector_state(red, sc), 2, E, Hp, H, sc)
        lead = a.shape[:3] if kind == "phasor" else a.shape[:2]
        assert b.shape == lead
        assert np.allclose(b, a.reshape(*lead, -1).mean(axis=-1), rtol=1e-12, atol=1e-14)
    one = _detector(sc, "flux", axis=1)
    allc = _detector(sc, "flux", axis=1, keep_all_components=True)
    allr = _detector(sc, "flux", axis=1, keep_all_components=True, reduce_volume=True)
    a = O.detector_update(one, O.init_detector_state(one, sc), 2, E, Hp, H, sc)
    b = O.detector_update(allc, O.init_detector_state(allc, sc), 2, E, Hp, H, sc)
    c = O.detector_update(allr, O.init_detector_state(allr, sc), 2, E, Hp, H, sc)
    assert np.array_equal(b[1, 1], a[1])
    assert np.allclose(c[1], 0.61 * b[1].reshape(3, -1).sum(axis=-1), rtol=1e-12)
    # every component against its own cross product: S_x = Ey Hz - Ez Hy, sign -1
    Ei, Hi = O.interpolate_fields(E, (Hp + H) / 2)
    box = (slice(None), *allc["slc"])
    Ei, Hi = Ei[box], Hi[box]
    assert np.allclose(b[1, 0], -(Ei[1] * Hi[2] - Ei[2] * Hi[1]), rtol=1e-12, atol=1e-14)
    assert np.allclose(b[1, 2], -(Ei[0] * Hi[1] - Ei[1] * Hi[0]), rtol=1e-12, atol=1e-14)


def _flux_vector(Ei, Hi):
    return np.stack([Ei[(a + 1) % 3] * Hi[(a + 2) % 3] - Ei[(a + 2) % 3] * Hi[(a + 1) % 3] for a in range(3)])


def test_energy_slices_are_means_and_planes_of_the_cell_energy():
    """energy.py:107-132: as_slices keeps, per recorded step, either the means of the per-cell
    energy along each collapsed axis (aggregate "mean" or a missing slice position) or its planes
    at the slice cells; against the per-cell energy detector over the same box."""
    sc = _detector_scene(False)
    rng = np.random.default_rng(8)
    E, Hp, H = (rng.standard_normal((3, *sc.shape)) for _ in range(3))
    box = ((1, 6), (0, 4), (2, 10))
    cell = _detector(sc, "energy", slc=box)
    full = O.detector_update(cell, O.init_detector_state(cell, sc), 2, E, Hp, H, sc)[1]
    for ax in range(3):
        m = _detector(sc, "energy_mean", slc=box, axis=ax)
        got = O.detector_update(m, O.init_detector_state(m, sc), 2, E, Hp, H, sc)[1]
        assert np.allclose(got, full.mean(axis=ax), rtol=1e-12)
    spec = energy_slices_spec()
    dets = {d["name"]: d for d in S.build_scene(spec).detectors}
    assert [dets[f"planes:{k}"]["slc"][ax] for k, ax in (("XY Plane", 2), ("XZ Plane", 1), ("YZ Plane", 0))] == \
        [slice(5 + 5, 5 + 6), slice(6 + 6, 6 + 7), slice(5 + 4, 5 + 5)]


def test_closed_surface_flux_is_the_divergence_sum():
    """The net outward flux through the faces of a box equals the sum over the box of the discrete
    divergence of the sampled Poynting vector (the telescoping sums of S_a[i+1] - S_a[i] along each
    axis); "inward" negates it; axes of size one are skipped by default."""
    sc = _detector_scene(False)
    rng = np.random.default_rng(6)
    E, Hp, H = (rng.standard_normal((3, *sc.shape)) for _ in range(3))
    slc = ((1, 6), (0, 4), (2, 10))
    det = _detector(sc, "closed_flux", slc=slc)
    got = O.detector_update(det, O.init_detector_state(det, sc), 2, E, Hp, H, sc)[1, 0]
    Ei, Hi = O.interpolate_fields(E, (Hp + H) / 2)
    Sv = _flux_vector(Ei, Hi)[(slice(None), *det["slc"])]
    div = sum(np.sum(np.diff(Sv[a], axis=a)) for a in range(3)) * 0.61
    assert math.isclose(got, div, rel_tol=1e-12)
    inward = _detector(sc, "closed_flux", slc=slc, orientation="inward")
    assert O.detector_update(inward, O.init_detector_state(inward, sc), 2, E, Hp, H, sc)[1, 0] == -got
    spec = W.scenario_c1(T=4, n=10, pml=3, cube=2)
    spec["detectors"] = [dict(name="cf", kind="closed_flux", slc=((2, 7), (4, 5), (1, 9)))]
    assert S.build_scene(spec).detectors[0]["axes"] == (0, 2)


def _pure_tones(rng, shape, w, t):
    A, ph = rng.uniform(0.5, 2.0, shape), rng.uniform(-np.pi, np.pi, shape)
    sig = A[..., None] * np.cos(w * t + ph[..., None])  # (..., time)
    return sig, A * np.exp(-1j * ph)  # the continuous-scaled DFT of a whole-period tone


def test_phasor_flux_is_the_time_averaged_flux_of_pure_tones():
    """poynting_flux.py:361-387 (oracle and host): for fields that are pure tones on a plane, the
    phasor flux 1/2 Re(E x H*) with continuous scaling equals the time average over a period of
    the instantaneous flux E x H, per axis and signed; pulse scaling drops the 1/2. The net phasor
    flux of a box's faces equals the telescoping divergence sum of the phasor flux over the box."""
    import types

    import torch

    import fdtdx_amd as fx

    rng = np.random.default_rng(9)
    per, w = 16, 2 * np.pi / 16
    t = np.arange(per)
    shape = (3, 1, 4, 3)
    Es, Ep = _pure_tones(rng, shape, w, t)
    Hs, Hp = _pure_tones(rng, shape, w, t)
    inst = _flux_vector(Es, Hs).mean(axis=-1)  # (3, 1, 4, 3)
    ph = np.concatenate([Ep, Hp])[None]  # (1 freq, 6, 1, 4, 3)
    wgt = 0.37
    cfg = types.SimpleNamespace(uniform_spacing=lambda: math.sqrt(wgt))
    for axis, sign in ((0, 1), (1, -1)):
        want = sign * wgt * np.sum(inst[axis])
        assert math.isclose(O.phasor_poynting_flux(ph, axis, sign, wgt)[0], want, rel_tol=1e-12)
        assert math.isclose(O.phasor_poynting_flux(ph, axis, sign, wgt, continuous=False)[0], 2 * want,
                            rel_tol=1e-12)
        det = fx.PhasorPoyntingFluxDetector(wave_characters=[fx.WaveCharacter(frequency=1.0)],
                                            direction="+" if sign > 0 else "-", fixed_propagation_axis=axis)
        det = det.place_on_grid(((0, 1), (0, 4), (0, 3)), cfg)
        host = det.compute_poynting_flux({"phasor": torch.tensor(ph[None])})
        assert math.isclose(float(host[0]), want, rel_tol=1e-6)
    allc = O.phasor_poynting_flux(ph, 0, 1, wgt, keep_all=True)
    assert np.allclose(allc[0], wgt * inst.reshape(3, -1).sum(axis=1), rtol=1e-12)
    # closed box: faces of a volume of tones against the divergence of its time-averaged flux
    vshape = (3, 4, 5, 3)
    Es, Ep = _pure_tones(rng, vshape, w, t)
    Hs, Hp = _pure_tones(rng, vshape, w, t)
    Sbar = _flux_vector(Es, Hs).mean(axis=-1)
    vol = np.concatenate([Ep, Hp])[None, None]  # (1, 1 freq, 6, 4, 5, 3)
    faces = {}
    for a in (0, 1, 2):
        for side in ("min", "max"):
            sl = [slice(None)] * 6
            sl[3 + a] = slice(-1, None) if side == "max" else slice(0, 1)
            faces[(a, side)] = vol[tuple(sl)]
    div = sum(np.sum(np.diff(Sbar[a], axis=a)) for a in range(3)) * wgt
    assert math.isclose(O.closed_phasor_net_flux(faces, wgt)[0], div, rel_tol=1e-10)
    assert math.isclose(O.closed_phasor_net_flux(faces, wgt, inward=True)[0], -div, rel_tol=1e-10)
    det = fx.ClosedSurfacePhasorPoyntingFluxDetector(wave_characters=[fx.WaveCharacter(frequency=1.0)])
    det = det.place_on_grid(((0, 4), (0, 5), (0, 3)), cfg)
    st = {f"phasor_axis{a}_{side}": torch.tensor(v) for (a, side), v in faces.items()}
    assert math.isclose(float(det.compute_net_flux(st)[0]), div, rel_tol=1e-6)


@pytest.mark.parametrize("keep_all", [False, True])
def test_phasor_flux_cotangent_matches_finite_differences(keep_all):
    """phasor_poynting_flux_cotangent against FP64 central differences of sum g * flux along random
    complex directions of the phasors (dL/dRe + i dL/dIm convention)."""
    rng = np.random.default_rng(12)
    shp = (2, 6, 1, 3, 4)
    ph = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    d = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    g = rng.standard_normal((2, 3) if keep_all else (2,))

    def L(p):
        return float(np.sum(g * O.phasor_poynting_flux(p, 0, -1, 0.3, keep_all=keep_all)))

    h = 1e-4
    fd = (L(ph + h * d) - L(ph - h * d)) / (2 * h)
    cot = O.phasor_poynting_flux_cotangent(ph, 0, -1, 0.3, g, keep_all=keep_all)
    adj = float(np.sum(cot.real * d.real + cot.imag * d.imag))
    assert math.isclose(adj, fd, rel_tol=1e-8)


def test_gradient_through_surface_detectors_matches_finite_differences():
    """The reversible gradient of the net flux through a closed box, of the phasor flux through a
    plane and of the net phasor flux through the faces of a box (each loss in turn) against FP64
    central differences (h = 1e-3) w.r.t. design voxels."""
    spec = surface_zoo_spec()
    sc = S.build_scene(spec, np.float64)
    names = [d["name"] for d in sc.detectors]
    T = spec["T"]
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    deps = -(d["eps_b"] - d["eps_a"]) / eps**2

    def scene_fn(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        return S.build_scene(s2, np.float64)

    keys = [x["name"] for x in spec["detectors"]]
    grads = {}
    for key in keys:
        def cot_fn(ds, key=key):
            return [surface_zoo_losses(ds, names, spec, spec["dx"])[1][key].get(nm) for nm in names]
        g, _, _ = O.reversible_gradient(sc, T, cot_fn)
        grads[key] = g[(0, *slc)] * deps
    for idx in [(1, 2, 1), (4, 0, 3)]:
        fd = O.detector_losses_fd(scene_fn, rho, idx, 1e-3, T,
                                  lambda ds: [surface_zoo_losses(ds, names, spec, spec["dx"])[0][k] for k in keys])
        for q, key in enumerate(keys):
            assert abs(fd[q]) > 0
            assert abs(grads[key][idx] - fd[q]) <= 1e-3 * abs(fd[q]), (key, idx, grads[key][idx], fd[q])


# ---------------------------------------------------------------- full-tensor materials
def _yee_linear(shape, comp, kind, coef):
    """A linear function sampled at component `comp`'s Yee positions: E_c at p + e_c/2, H_c at
    p + (1,1,1)/2 - e_c/2 (curl.py:227-397 stencils)."""
    idx = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    off = np.full(3, 0.0 if kind == "E" else 0.5)
    off[comp] = 0.5 if kind == "E" else 0.0
    return sum(coef[a] * (idx[a] + off[a]) for a in range(3)) + coef[3]


def test_tensor_averages_are_exact_for_linear_fields():
    """fdtd/misc.py:132-213: the four-point means carry a component to another component's Yee
    position, so for a linear field they equal the field evaluated there (cells whose stencil
    stays inside the domain); their transposes satisfy <avg F, g> = <F, avg^T g>."""
    shape = (6, 5, 7)
    coef = (0.7, -1.3, 2.1, 0.4)
    rng = np.random.default_rng(3)
    for kind, avg, avgT in (("E", O.avg_E, O.avg_E_T), ("H", O.avg_H, O.avg_H_T)):
        F = np.stack([_yee_linear(shape, c, kind, coef) for c in range(3)])
        for comp in range(3):
            for loc in range(3):
                if comp == loc:
                    continue
                got = avg(F, comp, loc)
                want = _yee_linear(shape, loc, kind, coef)
                inner = [slice(1, -1)] * 3
                assert np.allclose(got[tuple(inner)], want[tuple(inner)], rtol=0, atol=1e-12), (kind, comp, loc)
                a = rng.standard_normal((3, *shape))
                g = rng.standard_normal(shape)
                assert math.isclose(np.sum(avg(a, comp, loc) * g), np.sum(a[comp] * avgT(g, comp, loc)),
                                    rel_tol=1e-12)
                per = (True, False, True)
                assert math.isclose(np.sum(avg(a, comp, loc, per) * g), np.sum(a[comp] * avgT(g, comp, loc, per)),
                                    rel_tol=1e-12)


def test_diagonal_tensor_equals_the_diagonal_update():
    """A 9-component tensor with zero off-diagonal entries runs the full-tensor branch
    (update.py:356-492, 752-822) and must reproduce the diagonal branch (FP64), dipole included
    (its off-diagonal injections are zero)."""
    spec = W.scenario_c1(T=50, n=16, pml=4, cube=4)
    spec["detectors"] = []
    spec["boxes"] = [dict(b, eps=(2.0, 3.0, 4.0), mu=(1.5, 1.2, 2.0)) for b in spec["boxes"]]
    diag = S.build_scene(spec, np.float64)
    spec9 = dict(spec, boxes=[dict(b, eps=(2.0, 0, 0, 0, 3.0, 0, 0, 0, 4.0), mu=(1.5, 0, 0, 0, 1.2, 0, 0, 0, 2.0))
                              for b in spec["boxes"]])
    full = S.build_scene(spec9, np.float64)
    assert full.inv_eps.shape[0] == 9 and full.inv_mu.shape[0] == 9
    a, _ = O.run(diag, spec["T"])
    b, _ = O.run(full, spec["T"])
    scale = np.abs(a.E).max()
    assert np.abs(a.E - b.E).max() < 1e-12 * scale
    assert np.abs(a.H - b.H).max() < 1e-12 * np.abs(a.H).max()


@pytest.mark.parametrize("which", ["eps", "mu"])
def test_nonsymmetric_tensor_is_the_diagonal_medium_in_its_eigenbasis(which):
    """For fields uniform in x and y (periodic x, y; CPML along z) the xy block M = V D V^-1 of a
    non-symmetric tensor decouples: with E = V E~ and H = J^-1 V J H~ (J the quarter turn that the
    z-curl applies), the tensor run equals the run of the diagonal medium D (FP64). Using M^T,
    swapping rows and columns of the flattened tensor, or a wrong off-diagonal sign breaks it."""
    V = np.array([[1.0, 0.4], [0.3, 1.0]])
    D = np.array([0.5, 0.3])  # inverse permittivities / permeabilities along the eigenvectors
    M = V @ np.diag(D) @ np.linalg.inv(V)
    J = np.array([[0.0, -1.0], [1.0, 0.0]])
    inv9 = np.zeros(9)
    inv9[[0, 1, 3, 4]] = M.reshape(4)
    inv9[8] = 0.25
    base = dict(shape=(2, 2, 40), dx=50e-9, T=40, background_eps=1.0,
                pml={"thickness": 6, "types": dict(PERIODIC_XY)}, sources=[], detectors=[])
    t9 = tuple(np.linalg.inv(inv9.reshape(3, 3)).reshape(9))
    tdiag = (1 / D[0], 1 / D[1], 4.0)
    box = ((0, 2), (0, 2), (10, 30))
    key = "eps" if which == "eps" else "mu"
    sc_t = S.build_scene(dict(base, boxes=[dict(dict(slc=box, eps=1.0), **{key: t9})]), np.float64)
    sc_d = S.build_scene(dict(base, boxes=[dict(dict(slc=box, eps=1.0), **{key: tdiag})]), np.float64)
    rng = np.random.default_rng(7)
    z = np.arange(40)
    prof = lambda: np.exp(-((z - 20 + rng.uniform(-3, 3)) / 4.0) ** 2) * rng.uniform(0.5, 1.5)
    Et = np.stack([prof(), prof()])
    Ht = np.stack([prof(), prof()])
    # eps: E = V E~, H = J^-1 V J H~; mu: H = V H~, E = J^-1 V J E~
    TE, TH = (V, np.linalg.inv(J) @ V @ J) if which == "eps" else (np.linalg.inv(J) @ V @ J, V)
    sd, st_ = O.init_state(sc_d), O.init_state(sc_t)
    for state, tE, tH in ((sd, np.eye(2), np.eye(2)), (st_, TE, TH)):
        state.E[:2] = np.einsum("ab,bz->az", tE, Et)[:, None, None, :]
        state.H[:2] = np.einsum("ab,bz->az", tH, Ht)[:, None, None, :]
    for n in range(base["T"]):
        O.forward_step(sd, sc_d, n)
        O.forward_step(st_, sc_t, n)
    wantE = np.einsum("ab,bxyz->axyz", TE, sd.E[:2])
    wantH = np.einsum("ab,bxyz->axyz", TH, sd.H[:2])
    assert np.abs(sd.E).max() > 1e-3
    assert np.abs(st_.E[:2] - wantE).max() < 1e-12 * np.abs(wantE).max()
    assert np.abs(st_.H[:2] - wantH).max() < 1e-12 * np.abs(wantH).max()


def test_tensor_gradient_matches_finite_differences():
    """The reversible gradient in a full-tensor scene (a non-symmetric tensor box beside the design,
    whose diagonal entries the isotropic design writes): design voxels through rho, and an
    off-diagonal entry of inv_eps directly, against FP64 central differences."""
    spec = W.scenario_c3(T=90, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    spec["boxes"] = [dict(slc=((5, 9), (6, 16), (6, 14)), eps=TENSOR_EPS, mu=TENSOR_MU)]
    sc = S.build_scene(spec, np.float64)
    assert sc.inv_eps.shape[0] == 9
    grad, ds, _ = O.reversible_gradient(sc, spec["T"], lambda ds: [np.ones_like(ds[0])])
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    g = (grad[0] + grad[4] + grad[8])[slc] * (-(d["eps_b"] - d["eps_a"]) / eps**2)

    def scene_fn(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        return S.build_scene(s2, np.float64)

    for idx in [(1, 2, 1), (4, 0, 3)]:
        fd = O.flux_loss_and_grad_fd(scene_fn, rho, idx, 1e-3, spec["T"])
        assert abs(g[idx] - fd) <= 1e-3 * abs(fd), (idx, g[idx], fd)
    for comp, cell in ((1, (7, 10, 9)), (5, (6, 12, 10)), (7, (8, 9, 8))):
        vals = []
        for sgn in (1, -1):
            s2 = S.build_scene(spec, np.float64)
            s2.inv_eps[(comp, *cell)] += sgn * 1e-4
            _, dd = O.run(s2, spec["T"])
            vals.append(np.sum(dd[0]))
        fd = (vals[0] - vals[1]) / 2e-4
        assert abs(grad[(comp, *cell)] - fd) <= 1e-3 * abs(fd), (comp, cell, grad[(comp, *cell)], fd)


# ---------------------------------------------------------------- mirror symmetry
def test_mirror_pad_is_the_transposed_fold():
    """<mirror_pad F, G> = <F, mirror_pad^T G> with two mirror axes, for E and for H."""
    rng = np.random.default_rng(21)
    F = rng.standard_normal((3, 4, 5, 6))
    for ft in ("E", "H"):
        P = O.mirror_pad(F, (False, True, True), ft)
        G = rng.standard_normal(P.shape)
        assert math.isclose(np.sum(P * G), np.sum(F * O.mirror_pad_T(G, (False, True, True), ft)), rel_tol=1e-12)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


SYMMETRY_CASES = [((0, 1, -1), (0, 0, 1)), ((0, -1, 1), (0, 1, 0))]


@pytest.mark.parametrize("symmetry,e_pol", SYMMETRY_CASES)
def test_symmetry_reduced_run_matches_the_full_run(symmetry, e_pol):
    """fdtd/symmetry.py:84-264: for a scene mirror-symmetric about the y and z mid-planes (with a
    wave of the matching parity), the reduced run (kept upper quarter, PEC face on the electric
    plane, zero halo below the magnetic one, CPML only on the far faces, plane-source energy
    normalised over the full plane) reproduces the kept quarter of the full run, and four times its
    flux. The agreement is that of the discretisation, not exact: per-cell rasterisation and the
    staggered CPML are mirror symmetric only to first order about the planes (measured 0.1-0.5 %
    for fields and detectors, 1 % for the flux). The detector stencils read the mirror halo of an
    electric y plane (update.py:139-190): with a zero halo there the snapshot and phasors are off
    by 13 %, with the wrong parity by 27 %; a PMC in place of the PEC face or a missing
    multiplicity are off by 60-300 %."""
    spec = symmetry_spec(symmetry=symmetry, e_pol=e_pol)
    red = S.build_scene(spec, np.float64)
    ful = S.build_scene(full_spec(spec), np.float64)
    assert red.shape == (36, 12, 10) and red.mirror == (False, symmetry[1] == -1, symmetry[2] == -1)
    assert [w[:2] for w in red.walls] == [("pec", 1 if symmetry[1] == -1 else 2)]
    sr, dr = O.run(red, spec["T"])
    sf, df = O.run(ful, spec["T"])
    kept = (slice(None), slice(None), slice(12, None), slice(10, None))
    assert _rel(sr.E, sf.E[kept]) < 5e-3 and _rel(sr.H, sf.H[kept]) < 5e-3
    assert _rel(4 * dr[0], df[0]) < 3e-2
    assert _rel(dr[1], df[1][..., 12:, 10:]) < 1e-2
    assert _rel(dr[2], df[2][..., 12:, 10:]) < 1e-2


def test_symmetry_reduced_gradient_matches_the_full_gradient():
    """A design voxel of the reduced run stands for its four mirror images and the reduced flux is
    a quarter of the full one, so the reduced gradient of the reduced flux matches the full-domain
    gradient of the full flux at each kept voxel, to the discretisation's first-order mirror
    asymmetry; the detector term reads the electric plane's mirror halo through the transposed
    fold (mirror_pad_T)."""
    spec = symmetry_spec(symmetry=(0, -1, 1), e_pol=(0, 1, 0))
    red = S.build_scene(spec, np.float64)
    ful = S.build_scene(full_spec(spec), np.float64)
    gr, _, _ = O.reversible_gradient(red, spec["T"], lambda ds: [np.ones_like(ds[0]), None, None])
    gf, _, _ = O.reversible_gradient(ful, spec["T"], lambda ds: [np.ones_like(ds[0]), None, None])
    d = spec["design"]
    (x0, x1), (y0, y1), (z0, z1) = d["slc"]
    g_red = gr[0][x0:x1, 0:y1 - 12, 0:z1 - 10]
    g_full = gf[0][x0:x1, 12:y1, 10:z1]
    assert np.abs(g_full).max() > 0
    assert _rel(g_red, g_full) < 3e-2, _rel(g_red, g_full)


# ---------------------------------------------------------------- reversibility and adjoint
# a weakly conducting cube (x = 0.01 per step for E), magnetic loss matched to it
LOSSY_CUBE = dict(sigma_E=1850.0, sigma_H=1850.0 * O.ETA0**2)


@pytest.mark.parametrize("lossy", [False, True, "tensor"])
def test_time_reversal_round_trip(lossy):
    """test_time_reversal.py:251-472 style: run forward with interface recording, reverse all steps;
    the interior returns to the zero initial state (FP64: to roundoff), also through the lossy
    reverse update (update.py:600-606, 921-926)."""
    spec = W.scenario_c1(T=60, n=20, pml=5, cube=4)
    spec["detectors"] = []
    if lossy == "tensor":  # full-tensor eps and mu around the dipole (update.py:611-680, 960-1000)
        spec["boxes"] = [dict(b, slc=((6, 14), (6, 14), (6, 14)), eps=TENSOR_EPS, mu=TENSOR_MU) for b in spec["boxes"]]
    elif lossy:
        spec["boxes"] = [dict(b, slc=((6, 14), (6, 14), (6, 14)), **LOSSY_CUBE) for b in spec["boxes"]]
    sc = S.build_scene(spec, np.float64)
    st = O.init_state(sc)
    rec = []
    for n in range(spec["T"]):
        O.forward_step(st, sc, n, record=rec)
    peak = np.abs(st.E).max()
    for n in reversed(range(spec["T"])):
        O.reverse_step(st, sc, n, rec[n])
    inner = (slice(None), slice(5, 15), slice(5, 15), slice(5, 15))
    assert np.abs(st.E[inner]).max() < 1e-9 * peak
    assert np.abs(st.H[inner]).max() < 1e-9 * peak


@pytest.mark.parametrize("walls", [{}, {"max_x": "pec", "min_y": "pmc", "max_z": "pec", "min_z": "pmc"},
                                   PERIODIC_XY, "lossy", "kappa_mu", "tensor", "tensor_periodic"])
def test_adjoint_step_dot_product(walls):
    """<J d, l> = <d, J^T l> for one sourceless step, with J the linear map (E, H, psi) -> next state,
    with CPML on every face, with PEC / PMC faces replacing some of them, with periodic axes and
    with conducting cells (diagonal sigma_E, isotropic sigma_H) that reach into the CPML."""
    spec = W.scenario_c1(T=2, n=14, pml=4, cube=2)
    spec["sources"] = []
    if walls == "lossy":
        spec["boxes"] = spec["boxes"] + [dict(slc=((2, 9), (3, 12), (5, 7)), eps=(2.0, 3.0, 4.0),
                                              sigma_E=(1e5, 2e5, 5e4), sigma_H=4e9)]
        walls = {}
    if walls == "kappa_mu":  # CPML with kappa > 1 (PML.py:184-188) and a magnetic box reaching into it
        spec["pml"] = dict(spec["pml"], params=dict(kappa_start=1.0, kappa_end=4.0))
        spec["boxes"] = spec["boxes"] + [dict(slc=((1, 8), (4, 12), (3, 9)), eps=2.0, mu=(1.5, 2.5, 3.0))]
        walls = {}
    if walls in ("tensor", "tensor_periodic"):  # non-symmetric eps and mu tensors reaching into the CPML
        spec["boxes"] = spec["boxes"] + [dict(slc=((1, 8), (4, 12), (3, 9)), eps=TENSOR_EPS, mu=TENSOR_MU)]
        walls = {} if walls == "tensor" else dict(PERIODIC_XY)
    spec["pml"] = dict(spec["pml"], types=walls)
    sc = S.build_scene(spec, np.float64)
    rng = np.random.default_rng(5)

    def rand_state():
        st = O.init_state(sc)
        st.E = rng.standard_normal(st.E.shape)
        st.H = rng.standard_normal(st.H.shape)
        st.psiE = [(rng.standard_normal(a.shape), rng.standard_normal(b.shape)) for a, b in st.psiE]
        st.psiH = [(rng.standard_normal(a.shape), rng.standard_normal(b.shape)) for a, b in st.psiH]
        return st

    d = rand_state()
    l = rand_state()
    x = O.State(d.E.copy(), d.H.copy(), [(a.copy(), b.copy()) for a, b in d.psiE],
                [(a.copy(), b.copy()) for a, b in d.psiH])
    O.forward_step(x, sc, 0)

    def dot(s1, s2):
        v = np.sum(s1.E * s2.E) + np.sum(s1.H * s2.H)
        v += sum(np.sum(a * c) + np.sum(b * e) for (a, b), (c, e) in zip(s1.psiE, s2.psiE))
        v += sum(np.sum(a * c) + np.sum(b * e) for (a, b), (c, e) in zip(s1.psiH, s2.psiH))
        return v

    lam = {"lamE": l.E, "lamH": l.H, "lampsiE": l.psiE, "lampsiH": l.psiH}
    out, _ = O.adjoint_step(lam, sc, d.H, d.psiE, E_n=d.E)
    back = O.State(out["lamE"], out["lamH"], out["lampsiE"], out["lampsiH"])
    assert math.isclose(dot(x, l), dot(d, back), rel_tol=1e-10)


@pytest.mark.parametrize("num_checkpoints,walls", [(0, {}), (2, {}), (0, {"max_x": "pec", "min_z": "pmc"})])
def test_reversible_gradient_matches_finite_differences(num_checkpoints, walls):
    """FP64 central differences (h = 1e-3) of L = sum_n flux_n w.r.t. design voxels (as
    test_time_reversal.py:652-685, rel < 1e-3), for the classic single-slice reverse pass, for
    the sliced pass with two interior checkpoints (fdtd.py:106-121), and with PEC / PMC faces."""
    spec = W.scenario_c3(T=90, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    spec["pml"] = dict(spec["pml"], types=walls)
    sc = S.build_scene(spec, np.float64)
    grad, ds, _ = O.reversible_gradient(sc, spec["T"], lambda ds: [np.ones_like(ds[0])], num_checkpoints)
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    g = grad[(0, *slc)] * (-(d["eps_b"] - d["eps_a"]) / eps**2)

    def scene_fn(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        return S.build_scene(s2, np.float64)

    for idx in [(1, 2, 1), (4, 0, 3)]:
        fd = O.flux_loss_and_grad_fd(scene_fn, rho, idx, 1e-3, spec["T"])
        assert abs(g[idx] - fd) <= 1e-3 * abs(fd), (idx, g[idx], fd)


def test_gradient_through_every_detector_kind_matches_finite_differences():
    """The reversible gradient with the cotangent of each detector kind and option in turn against
    FP64 central differences (h = 1e-3) of that detector's loss w.r.t. design voxels."""
    spec = detector_zoo_spec()
    sc = S.build_scene(spec, np.float64)
    T = spec["T"]
    rng = np.random.default_rng(8)
    st = [O.init_detector_state(d, sc) for d in sc.detectors]
    weights = dict(fld=rng.standard_normal(st[3].shape), fl3=rng.standard_normal(st[5].shape))
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    deps = -(d["eps_b"] - d["eps_a"]) / eps**2

    def scene_fn(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        return S.build_scene(s2, np.float64)

    grads = []
    for q in range(len(sc.detectors)):
        def cot_fn(ds, q=q):
            _, cots = detector_zoo_losses(ds, weights)
            return [c if i == q else None for i, c in enumerate(cots)]
        g, _, _ = O.reversible_gradient(sc, T, cot_fn)
        grads.append(g[(0, *slc)] * deps)
    for idx in [(1, 2, 1), (4, 0, 3)]:
        fd = O.detector_losses_fd(scene_fn, rho, idx, 1e-3, T, lambda ds: detector_zoo_losses(ds, weights)[0])
        for q, g in enumerate(grads):
            assert abs(fd[q]) > 0
            assert abs(g[idx] - fd[q]) <= 1e-3 * abs(fd[q]), (sc.detectors[q]["name"], idx, g[idx], fd[q])


@pytest.mark.parametrize("num_checkpoints", [0, 2])
def test_lossy_gradient_matches_finite_differences(num_checkpoints):
    """The inv_eps derivative through the loss factor and divisor, (c K - 2 a E) / (1 + x)^2,
    with a conducting box overlapping the design region, against FP64 central differences."""
    spec = W.scenario_c3(T=90, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    (x0, x1), (y0, y1), (z0, z1) = spec["design"]["slc"]
    spec["boxes"] = [dict(slc=((x0 - 1, x1 - 2), (y0 + 1, y1 + 2), (z0 - 2, z1)), eps=1.0, sigma_E=9000.0,
                          sigma_H=3e9)]
    sc = S.build_scene(spec, np.float64)
    assert sc.sigma_E is not None and sc.sigma_H is not None
    grad, ds, _ = O.reversible_gradient(sc, spec["T"], lambda ds: [np.ones_like(ds[0])], num_checkpoints)
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    g = grad[(0, *slc)] * (-(d["eps_b"] - d["eps_a"]) / eps**2)

    def scene_fn(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        return S.build_scene(s2, np.float64)

    for idx in [(1, 2, 1), (4, 0, 3)]:  # inside and outside the conducting box
        fd = O.flux_loss_and_grad_fd(scene_fn, rho, idx, 1e-3, spec["T"])
        assert abs(g[idx] - fd) <= 1e-3 * abs(fd), (idx, g[idx], fd)


@pytest.mark.parametrize("num_checkpoints", [0, 2])
def test_kappa_and_magnetic_gradient_matches_finite_differences(num_checkpoints):
    """The adjoint through CPML with kappa > 1 (the derivative feeds the correction directly,
    (1/kappa - 1) d) and through a diagonal permeability box over part of the design."""
    spec = W.scenario_c3(T=90, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    (x0, x1), (y0, y1), (z0, z1) = spec["design"]["slc"]
    spec["pml"] = dict(spec["pml"], params=dict(kappa_start=1.0, kappa_end=3.0))
    spec["boxes"] = [dict(slc=((x0 - 1, x1 - 2), (y0 + 1, y1 + 2), (z0 - 2, z1)), eps=1.0, mu=(1.4, 2.2, 1.8))]
    sc = S.build_scene(spec, np.float64)
    assert not isinstance(sc.inv_mu, float) and any(not np.all(s.coeffs[2] == 1) for s in sc.slabs)
    grad, ds, _ = O.reversible_gradient(sc, spec["T"], lambda ds: [np.ones_like(ds[0])], num_checkpoints)
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    g = grad[(0, *slc)] * (-(d["eps_b"] - d["eps_a"]) / eps**2)

    def scene_fn(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        return S.build_scene(s2, np.float64)

    for idx in [(1, 2, 1), (4, 0, 3)]:
        fd = O.flux_loss_and_grad_fd(scene_fn, rho, idx, 1e-3, spec["T"])
        assert abs(g[idx] - fd) <= 1e-3 * abs(fd), (idx, g[idx], fd)


ALL_PERIODIC = {f: "periodic" for f in ("min_x", "max_x", "min_y", "max_y", "min_z", "max_z")}


def test_uniform_field_in_conductor_decays_by_closed_form():
    """With every axis periodic a uniform field has zero curl, so each step only applies the loss:
    E_n = ((1 - x) / (1 + x))^n E_0 with x = c sigma dx eta0 / (2 eps_r) and H likewise with
    x = c sigma_H dx / (2 eta0) (Schneider 3.12), sigma scaled by the grid spacing."""
    eps_r, sE, sH, T = 2.0, 5000.0, 3e8, 10
    spec = dict(shape=(4, 4, 4), dx=50e-9, T=T, background_eps=eps_r, pml={"thickness": 0, "types": ALL_PERIODIC},
                boxes=[dict(slc=((0, 4), (0, 4), (0, 4)), eps=eps_r, sigma_E=sE, sigma_H=sH)], sources=[],
                detectors=[])
    sc = S.build_scene(spec, np.float64)
    st = O.init_state(sc)
    e0, h0 = np.array([1.0, 2.0, 3.0]), np.array([-1.0, 0.5, 2.0])
    st.E[:] = e0[:, None, None, None]
    st.H[:] = h0[:, None, None, None]
    for n in range(T):
        O.forward_step(st, sc, n)
    cn = 0.99 / math.sqrt(3)
    xe = cn * sE * 50e-9 * O.ETA0 / (2 * eps_r)
    xh = cn * sH * 50e-9 / (2 * O.ETA0)
    np.testing.assert_allclose(st.E, e0[:, None, None, None] * ((1 - xe) / (1 + xe)) ** T * np.ones((1, 4, 4, 4)),
                               rtol=1e-12)
    np.testing.assert_allclose(st.H, h0[:, None, None, None] * ((1 - xh) / (1 + xh)) ** T * np.ones((1, 4, 4, 4)),
                               rtol=1e-12)


def _line_phasor(sigma_E, sigma_H, nx=400, box_lo=150, src=60, T=2600, lam_cells=40):
    """A 1-D line (periodic y, z of one cell, CPML on x): a dipole sheet at ``src`` launches plane
    waves; cells from ``box_lo`` on conduct. Returns the Ey phasor at the centre frequency."""
    dx = 50e-9
    lam = lam_cells * dx
    boxes = [] if not (sigma_E or sigma_H) else [dict(slc=((box_lo, nx), (0, 1), (0, 1)), eps=1.0,
                                                      sigma_E=sigma_E, sigma_H=sigma_H)]
    spec = dict(shape=(nx, 1, 1), dx=dx, T=T, background_eps=1.0, boxes=boxes,
                pml={"thickness": 20, "types": {f: "periodic" for f in ("min_y", "max_y", "min_z", "max_z")}},
                sources=[dict(kind="dipole", cell=(src, 0, 0), polarization=1,
                              profile=W.gaussian_profile(wavelength=lam, bandwidth=0.3))], detectors=[])
    sc = S.build_scene(spec, np.float64)
    st = O.init_state(sc)
    w = 2 * math.pi * O.C0 / lam
    ph = np.zeros(nx, complex)
    for n in range(T):
        O.forward_step(st, sc, n)
        ph += st.E[1, :, 0, 0] * np.exp(-1j * w * (n + 1) * sc.dt)
    return ph, w, dx


def test_conductor_attenuation_and_reflection_match_continuum_theory():
    """A plane wave entering a conductor with sigma / (omega eps0) = 0.5: its attenuation constant
    and the reflection at the interface match the continuum values (alpha within 1 %, |Gamma|
    within 3 %); with matched magnetic loss (sigma_H = eta0^2 sigma_E, so the medium's impedance
    is eta0) the attenuation is sigma eta0 and the reflection drops to discretisation level."""
    lam = 40 * 50e-9
    w = 2 * math.pi * O.C0 / lam
    sig = 0.5 * w * O.EPS0
    vac, _, dx = _line_phasor(0.0, 0.0)
    ph, _, _ = _line_phasor(sig, 0.0)
    phm, _, _ = _line_phasor(sig, sig * O.ETA0**2)
    x1, x2 = 200, 300
    alpha = w * math.sqrt(O.MU0 * O.EPS0 / 2) * math.sqrt(math.sqrt(1 + (sig / (w * O.EPS0)) ** 2) - 1)
    a_meas = math.log(abs(ph[x1]) / abs(ph[x2])) / ((x2 - x1) * dx)
    assert abs(a_meas / alpha - 1) < 0.01, (a_meas, alpha)
    eta = np.sqrt(1j * w * O.MU0 / (sig + 1j * w * O.EPS0))
    gamma = abs((eta - O.ETA0) / (eta + O.ETA0))
    g_meas = abs(ph[40] - vac[40]) / abs(vac[40])
    assert abs(g_meas / gamma - 1) < 0.03, (g_meas, gamma)
    am = math.log(abs(phm[x1]) / abs(phm[x2])) / ((x2 - x1) * dx)
    assert abs(am / (sig * O.ETA0) - 1) < 0.01, (am, sig * O.ETA0)
    assert abs(phm[40] - vac[40]) / abs(vac[40]) < gamma / 4


def test_slice_boundaries_match_reference_rounding():
    # fdtd.py:20-36: round(i * T / k); strictly increasing, first 0, last T
    assert O.slice_boundaries(10, 0) == [0, 10]
    assert O.slice_boundaries(10, 2) == [0, 3, 7, 10]
    assert O.slice_boundaries(3000, 3) == [0, 750, 1500, 2250, 3000]
    b = O.slice_boundaries(7, 6)
    assert b == list(range(8))


def test_sliced_reverse_resets_to_the_exact_state():
    """With a checkpoint at every step the reverse pass never reconstructs: each step's VJP runs
    at the exact forward state, so the gradient equals the adjoint of the stored trajectory,
    which the single-slice pass only approximates (its reconstruction drifts in the CPML)."""
    spec = W.scenario_c3(T=24, shape=(20, 18, 16), pml=4, design=(4, 4, 4))
    sc = S.build_scene(spec, np.float64)
    cot = lambda ds: [np.ones_like(ds[0])]
    g0, _, _ = O.reversible_gradient(sc, spec["T"], cot, 0)
    g_all, _, _ = O.reversible_gradient(sc, spec["T"], cot, spec["T"] - 1)
    d = tuple(slice(*r) for r in spec["design"]["slc"])
    a, b = g0[(0, *d)], g_all[(0, *d)]
    assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b)
    assert not np.array_equal(g0, g_all)  # the resets are taken


def test_fp32_oracle_tracks_fp64():
    spec = W.scenario_c1(T=150, n=24, pml=5, cube=6)
    s32, _ = O.run(S.build_scene(spec, np.float32), spec["T"])
    s64, _ = O.run(S.build_scene(spec, np.float64), spec["T"])
    rel = np.linalg.norm(s32.E - s64.E) / np.linalg.norm(s64.E)
    assert rel < 1e-5


@pytest.mark.parametrize("eps,wavelength,T", [(1.0, 0.5e-6, 300), (4.0, 1.0e-6, 400)])
def test_plane_source_is_one_way_with_medium_impedance_and_unit_energy(eps, wavelength, T):
    """Pins the normal-incidence restatement of UniformPlaneSource (linear_polarization.py:168-358,
    tfsf.py:144-409, core/grid.py:720-880) by physics, not by retyping it:
      - one-way injection: the field behind the plane stays below 2 % of the field ahead (a
        flipped term sign gives 336 %, dropped Yee time offsets 16 %, a missing half step 9 %);
      - impedance: the peak E / H ahead of the plane is 1 / sqrt(eps_r) within 3 %;
      - unit-energy normalisation: the peak E is the closed-form 1 / sqrt(cells * (eps_r + 1) / 2)
        times the pulse peak, within 5 %."""
    prof = W.gaussian_profile(wavelength=wavelength, bandwidth=0.5)
    spec = dict(shape=(120, 32, 32), dx=50e-9, T=T, background_eps=eps, pml={"thickness": 6},
                sources=[dict(kind="plane", axis=0, index=30, direction="+", e_pol=(0, 0, 1), profile=prof)],
                detectors=[])
    sc = S.build_scene(spec, np.float64)
    st = O.init_state(sc)
    behind, ahead, h_ahead = [], [], []
    for n in range(T):
        O.forward_step(st, sc, n, [])
        behind.append(st.E[2, 22, 16, 16])
        ahead.append(st.E[2, 60, 16, 16])
        h_ahead.append(st.H[1, 60, 16, 16])
    b, a, h = (np.abs(np.asarray(v)).max() for v in (behind, ahead, h_ahead))
    assert b / a < 0.02
    assert abs(a / h - 1 / np.sqrt(eps)) < 0.03 / np.sqrt(eps)
    tt = np.linspace(0, T * sc.dt, 20000)
    g = np.abs(O.gaussian_pulse(tt, prof["center_frequency"], prof["spectral_width"], dtype=np.float64)).max()
    e0 = 1.0 / np.sqrt(32 * 32 * 0.5 * (eps + 1.0))
    assert abs(a / (e0 * g) - 1) < 0.05


def test_pec_and_pmc_faces_hold_tangential_fields_at_zero():
    """pec.py:70-77 / pmc.py:62-75: after every step the tangential E on a PEC face and the
    tangential H on a PMC face are exactly zero, while the wave has reached the cells next to
    both faces."""
    spec = W.scenario_c1(T=80, n=20, pml=4, cube=4)
    spec["pml"] = dict(spec["pml"], types={"max_x": "pec", "min_y": "pmc"})
    spec["sources"][0]["profile"] = W.gaussian_profile(wavelength=0.4e-6, bandwidth=0.8)
    sc = S.build_scene(spec, np.float64)
    assert sc.walls == [("pec", 0, 19), ("pmc", 1, 0)]
    st = O.init_state(sc)
    for n in range(spec["T"]):
        O.forward_step(st, sc, n)
        assert not st.E[1:, 19].any() and not st.H[0, :, 0].any() and not st.H[2, :, 0].any()
    assert np.abs(st.E[1:, 18]).max() > 1e-3 * np.abs(st.E).max()
    assert np.abs(st.H[0, :, 1]).max() > 1e-3 * np.abs(st.H).max()


def test_periodic_domain_equals_its_doubled_copy():
    """Translation invariance pins the wrap: a domain periodic in x of length n with one dipole
    evolves exactly as the left half of a periodic domain of length 2n with the dipole repeated."""
    base = dict(dx=50e-9, T=60, background_eps=1.0, pml={"thickness": 4, "types": dict(PERIODIC_XY)},
                detectors=[])
    prof = W.gaussian_profile(wavelength=0.4e-6, bandwidth=0.8)
    a = dict(base, shape=(10, 12, 20), sources=[dict(kind="dipole", cell=(3, 6, 10), polarization=1, profile=prof)])
    b = dict(base, shape=(20, 12, 20), sources=[dict(kind="dipole", cell=(3, 6, 10), polarization=1, profile=prof),
                                                dict(kind="dipole", cell=(13, 6, 10), polarization=1, profile=prof)])
    sa, sb = S.build_scene(a, np.float64), S.build_scene(b, np.float64)
    assert sa.periodic == (True, True, False)
    st_a, _ = O.run(sa, a["T"])
    st_b, _ = O.run(sb, b["T"])
    assert np.abs(st_a.E).max() > 0
    np.testing.assert_allclose(st_b.E[:, :10], st_a.E, rtol=0, atol=1e-12 * np.abs(st_a.E).max())
    np.testing.assert_allclose(st_b.E[:, 10:], st_a.E, rtol=0, atol=1e-12 * np.abs(st_a.E).max())
    np.testing.assert_allclose(st_b.H[:, :10], st_a.H, rtol=0, atol=1e-12 * np.abs(st_a.H).max())


def test_periodic_gradient_matches_finite_differences():
    """Reversible gradient through wrapped curls and the wrapped detector stencil (the flux plane
    spans the periodic cross-section), against FP64 central differences."""
    spec = periodic_spec()
    sc = S.build_scene(spec, np.float64)
    grad, ds, _ = O.reversible_gradient(sc, spec["T"], lambda ds: [np.ones_like(ds[0])])
    d = spec["design"]
    slc = tuple(slice(*r) for r in d["slc"])
    rho = np.asarray(d["rho"], np.float64)
    eps = d["eps_a"] + rho * (d["eps_b"] - d["eps_a"])
    g = grad[(0, *slc)] * (-(d["eps_b"] - d["eps_a"]) / eps**2)

    def scene_fn(r):
        s2 = dict(spec)
        s2["design"] = dict(d, rho=r)
        return S.build_scene(s2, np.float64)

    for idx in [(1, 2, 1), (4, 5, 2)]:
        fd = O.flux_loss_and_grad_fd(scene_fn, rho, idx, 1e-3, spec["T"])
        assert abs(g[idx] - fd) <= 1e-3 * abs(fd), (idx, g[idx], fd)


def test_susceptibility_inversion_recovers_poles():
    """dispersion.py:988-1066 inverts the ADE coefficients of dispersion.py:789-857: the recovered
    gamma dt, omega0^2 dt^2 and a dt^2 equal the pole parameters, and chi(omega) equals the
    continuum Lorentz / Drude susceptibility in FP64."""
    dt = O.time_step_duration(50e-9)
    w = 2 * math.pi * O.C0 / 1.55e-6
    for pole, a_exp, w0 in [(("lorentz", 1.2 * w, 0.07 * w, 1.7), 1.7 * (1.2 * w) ** 2, 1.2 * w),
                            (("drude", 2.3 * w, 0.2 * w), (2.3 * w) ** 2, 0.0)]:
        c1, c2, c3 = (np.asarray(v, np.float64).reshape(1, 1, 1) for v in S.ade_coefficients([pole], dt))
        _, gdt, w0dt2, adt2 = S.pole_parameters(c1, c2, c3)
        assert math.isclose(gdt[0, 0, 0], pole[2] * dt, rel_tol=1e-9)
        assert math.isclose(adt2[0, 0, 0], a_exp * dt**2, rel_tol=1e-9)
        assert abs(w0dt2[0, 0, 0] - (w0 * dt) ** 2) <= 1e-9 * max((w0 * dt) ** 2, 1e-30) + 1e-15
        for om in (0.8 * w, w, 1.3 * w):
            chi = S.chi_of(c1, c2, c3, om, dt, np.complex128)[0, 0]
            ref = a_exp / (w0**2 - om**2 - 1j * pole[2] * om)
            assert abs(chi - ref) <= 1e-8 * abs(ref)


def _line_injection(pole, variant):
    """Backward / forward energy of a TFSF plane source in a Lorentz medium filling a 1-D line."""
    PER_YZ = {"min_y": "periodic", "max_y": "periodic", "min_z": "periodic", "max_z": "periodic"}
    dx = 50e-9
    lam = 40 * dx
    nx, T, x0 = 700, 900, 300
    spec = dict(shape=(nx, 1, 1), dx=dx, T=T, background_eps=2.0, pml={"thickness": 20, "types": PER_YZ},
                boxes=[dict(slc=((0, nx), (0, 1), (0, 1)), eps=2.0, poles=[pole])],
                sources=[dict(kind="plane", axis=0, index=x0, direction="+", e_pol=(0, 1, 0),
                              profile=W.gaussian_profile(wavelength=lam, bandwidth=0.5))], detectors=[])
    sc = S.build_scene(spec, np.float64)
    sc.sources, _ = S.plane_source(spec["sources"][0], sc, T, sc.inv_eps, **variant)
    st = O.init_state(sc)
    back = fwd = 0.0
    for n in range(T):
        O.forward_step(st, sc, n)
        back += float(np.sum(st.E[1, 100:x0 - 5, 0, 0] ** 2))
        fwd += float(np.sum(st.E[1, x0 + 5:nx - 100, 0, 0] ** 2))
    return back / fwd


def test_plane_source_in_lorentz_medium_injects_one_way():
    """linear_polarization.py:191-210, 330-351: in a dispersive background the source impedance
    uses the effective permittivity at the carrier and the E-side correction the broadband
    sqrt(eps(omega)/eps(omega_c)) filtered profile. Measured backward/forward energy of a 50 %
    bandwidth pulse: away from resonance 0.9 % (carrier-only matching 1.2 %, eps_inf 3.2 %); with
    the band reaching the resonance 2.4 % (carrier-only 8.1 %)."""
    w = 2 * math.pi * O.C0 / (40 * 50e-9)
    far = ("lorentz", 1.3 * w, 0.1 * w, 1.5)
    r_bb, r_c, r_inf = (_line_injection(far, v) for v in ({}, dict(broadband=False), dict(dispersive=False)))
    assert r_bb < 0.012 and r_bb < r_c < r_inf, (r_bb, r_c, r_inf)
    near = ("lorentz", 1.15 * w, 0.05 * w, 2.0)
    r_bb, r_c = (_line_injection(near, v) for v in ({}, dict(broadband=False)))
    assert r_bb < 0.03 and r_bb < r_c / 2, (r_bb, r_c)


# ---------------------------------------------------------------- time-reversed run (backward.py)
def rel(a, b):
    a, b = np.asarray(a, np.complex128), np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _backward_case(T=80):
    import scenarios as SC

    spec = SC.backward_spec(T)
    sc = S.build_scene(spec)
    rec = []
    st, ds = O.run(sc, T, record=rec)
    return spec, sc, rec, st, ds


@pytest.mark.parametrize("reset", [True, False])
def test_full_backward_reconstructs_the_interior_state(reset):
    """The fields after reversing steps T-1 .. T0 equal those of a forward run of T0 steps inside
    the CPML (the restored interface layers feed the interior exactly); with reset_fields the CPML
    regions are zero."""
    T, T0 = 80, 30
    spec, sc, rec, st, ds = _backward_case(T)
    mid, _ = O.run(sc, T0)
    O.full_backward(st, sc, T, T0, rec, ds, reset_fields=reset)
    p = max(s.hi - s.lo for s in sc.slabs)
    inner = (slice(None), slice(p, -p), slice(p, -p), slice(p, -p))
    assert rel(st.E[inner], mid.E[inner]) < 1e-4
    assert rel(st.H[inner], mid.H[inner]) < 1e-4
    slab = st.E[:, :p] if reset else None
    if reset:
        assert not np.any(slab) and not np.any(st.H[:, :, -p:])
    else:
        assert np.any(st.E[:, :p])


def test_inverse_detectors_record_the_state_before_each_step():
    """Reversed step n leaves the state before forward step n, which the forward field detector
    recorded at step n-1: the inverse field detector's row n equals the forward row n-1, row 0 the
    initial zero state. The inverse phasor subtracts (phasor.py:226-229) the per-step terms of those
    same fields at step n's phase; the inverse energy is the energy of those fields."""
    T = 80
    spec, sc, rec, st, ds = _backward_case(T)
    names = [d["name"] for d in sc.detectors]
    O.full_backward(st, sc, T, 0, rec, ds, reset_fields=True)
    fwd, ifld, iph, ien = (ds[names.index(k)] for k in ("fwd", "ifld", "iph", "ien"))
    scale = float(np.abs(fwd).max())
    assert np.abs(ifld[1:] - fwd[:-1]).max() < 1e-4 * scale
    assert np.abs(ifld[0]).max() < 1e-4 * scale
    d = sc.detectors[names.index("iph")]
    sel = ifld[:, [2, 4]].astype(np.float64)  # Ez, Hy
    want = np.zeros(iph.shape[1:], np.complex128)
    for n in range(T):
        ph = (d["cos"][:, n] + 1j * d["sin"][:, n]).astype(np.complex128) * float(d["scale"]) * float(d["window"][n])
        want -= ph.reshape(-1, 1, 1, 1, 1) * sel[n][None]
    assert rel(iph[0], want) < 1e-4
    de = sc.detectors[names.index("ien")]
    E, H = ifld[:, :3].astype(np.float64), ifld[:, 3:].astype(np.float64)
    box = tuple(slice(*r) for r in spec["detectors"][3]["slc"])
    ie = np.broadcast_to(sc.inv_eps, (3, *sc.shape))[(slice(None), *box)].astype(np.float64)
    u = 0.5 * ((E ** 2) / ie[None]).sum(axis=1) + 0.5 * (H ** 2).sum(axis=1) / float(np.mean(sc.inv_mu))
    want_e = (u.reshape(T, -1).sum(axis=1) * de["weight"])
    got = np.asarray(ien, np.float64).reshape(-1)
    assert want_e.max() > 0 and np.abs(got - want_e).max() < 1e-4 * want_e.max()


def test_single_frequency_known_values():
    """SingleFrequencyProfile.get_amplitude (profile.py:263-275, linear_rampup window.py:25-30):
    the real part of exp(-i(2 pi t/P + phi + pi)) ramped up linearly over 4 periods. Pinned by
    values the definition fixes without evaluating it: at t = kP the carrier is cos(phi + pi) =
    -cos(phi), at t = (k + 1/2)P it is +cos(phi); the ramp is m/4 at t = mP for m <= 4 and 1 after;
    and the steady part projects onto the frequency 1/P with unit amplitude and phase phi + pi (a
    dropped profile phase flips every sign, a ramp over the wrong number of periods fails the
    m/4 values, a period in place of a frequency fails the projection)."""
    import math

    P = 5.1e-15
    for phi in (0.0, 0.3, -1.1):
        for dt in (np.float64, np.float32):
            tol = 1e-12 if dt == np.float64 else 1e-5  # float32 phase of up to 75 rad
            k = np.arange(0, 12)
            at_k = O.single_frequency((k * P).astype(dt), P, phi, dtype=dt).astype(np.float64)
            ramp = np.minimum(k / 4.0, 1.0)
            assert np.allclose(at_k, -math.cos(phi) * ramp, rtol=0, atol=tol)
            half = ((k + 0.5) * P).astype(dt)
            at_h = O.single_frequency(half, P, phi, dtype=dt).astype(np.float64)
            assert np.allclose(at_h, math.cos(phi) * np.minimum((k + 0.5) / 4.0, 1.0), rtol=0, atol=tol)
        t = np.linspace(4 * P, 14 * P, 4001)[:-1]
        v = O.single_frequency(t, P, phi, dtype=np.float64)
        z = 2.0 * np.mean(v * np.exp(1j * 2 * math.pi * t / P))
        assert abs(abs(z) - 1.0) < 1e-9
        assert abs(np.angle(z * np.exp(1j * (phi + math.pi)))) < 1e-9  # z = exp(-i(phi + pi))


def test_mode_overlap_known_answers():
    """compute_overlap_to_mode (detectors/mode.py:379-421) and the unit-flux normalisation
    (metrics.py:160-221), pinned by identities the definitions fix: the overlap of a mode
    normalised to unit flux with itself is 1 (4 in pulse scaling) and -1 for a backward mode;
    the overlap is conjugate-linear in the phasors; a mode of the crossed polarisation overlaps
    zero; the Gaussian cross-section (detectors/mode.py:26-141) has amplitude 1 at its centre and
    1/e at one radius, E and H transverse and orthogonal, |H| = n |E|, and a divergence phase of
    k r^2 tan(theta) / (2 radius)."""
    import math

    rng = np.random.default_rng(7)
    shp = (3, 1, 9, 11)
    mE = rng.normal(size=shp) + 1j * rng.normal(size=shp)
    mH = rng.normal(size=shp) + 1j * rng.normal(size=shp)
    for axis in (0,):
        e1, h1 = O.unit_flux(mE, mH, axis)
        ph = np.concatenate([e1, h1])
        a = O.mode_overlap(ph, e1, h1, axis)
        a1, a2 = 1, 2
        flux = np.sum(0.5 * np.real(np.conj(e1[a1]) * h1[a2] - np.conj(e1[a2]) * h1[a1]))
        assert abs(abs(flux) - 1) < 1e-12
        assert abs(a - np.sign(flux)) < 1e-12
        assert abs(O.mode_overlap(ph, e1, h1, axis, pulse=True) - 4 * np.sign(flux)) < 1e-12
        c = 0.3 - 1.7j
        assert abs(O.mode_overlap(c * ph, e1, h1, axis) - np.conj(c) * a) < 1e-12
    u, v = np.meshgrid((np.arange(31) - 15.0) * 0.1, (np.arange(27) - 13.0) * 0.1, indexing="ij")
    u, v = u[None], v[None]  # plane normal to x: (1, ny, nz)
    E, H = O.gaussian_mode(u, v, 0, radius=0.8, direction="+", pol_axis=1, n=1.5, wavelength=1.0)
    assert abs(E[1, 0, 15, 13] - 1) < 1e-15 and np.all(E[0] == 0) and np.all(E[2] == 0)
    r = np.hypot(u, v)
    k = np.argmin(np.abs(r[0] - 0.8).ravel())
    assert abs(abs(E[1, 0].ravel()[k]) - math.exp(-(r[0].ravel()[k] / 0.8) ** 2)) < 1e-15
    assert np.allclose(np.abs(H[2]), 1.5 * np.abs(E[1])) and np.all(H[0] == 0) and np.all(H[1] == 0)
    Ez, Hz = O.gaussian_mode(u, v, 0, radius=0.8, direction="+", pol_axis=2, n=1.5, wavelength=1.0)
    e1, h1 = O.unit_flux(E, H, 0)
    ez, hz = O.unit_flux(Ez, Hz, 0)
    assert abs(O.mode_overlap(np.concatenate([ez, hz]), e1, h1, 0)) < 1e-14
    assert abs(O.mode_overlap(np.concatenate([e1, h1]), e1, h1, 0) - 1) < 1e-12
    Eb, Hb = O.unit_flux(*O.gaussian_mode(u, v, 0, 0.8, "-", 1, 1.5, 1.0), 0)
    assert abs(O.mode_overlap(np.concatenate([Eb, Hb]), Eb, Hb, 0) + 1) < 1e-12
    Ed, _ = O.gaussian_mode(u, v, 0, radius=0.8, direction="+", pol_axis=1, n=1.5, wavelength=1.0, divergence_deg=10)
    ph_d = np.angle(Ed[1, 0].ravel()[k])
    want = (2 * math.pi * 1.5 / 1.0) * r[0].ravel()[k] ** 2 * math.tan(math.radians(10)) / (2 * 0.8)
    assert abs(ph_d - want) < 1e-12
