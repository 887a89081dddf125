"""The mode solver (fdtdx_amd/modes.py), the mode source and the solver-based mode-overlap
detector on the host, pinned by what the mathematics fixes:

  - a homogeneous cross-section has the closed-form spectrum of the grid's own difference
    operators: beta^2 = k0^2 eps mu - 4 sin^2((2a+1) pi / (2 (2 n1 + 1))) - 4 sin^2((2b+1) pi /
    (2 (2 n2 + 1))), each value twice (both transverse polarisations), on the dense and the
    sparse (shift-invert) paths;
  - a mode of an anisotropic, magnetic waveguide, repeated along the propagation axis with the
    phase e^{-j beta s} (2 sin(beta/2) = neff k0 in cell units), satisfies the oracle's discrete
    curl equations curl E = -j k0 mu H and curl H = j k0 eps E at every interior cell, for each
    propagation axis (a transposed operator, a misplaced material component or a wrong sign of a
    longitudinal component fails it);
  - a dielectric slab converges to the analytic TE0 index at second order in the cell size;
  - unit Poynting flux, the reciprocity transformation of a backward mode, the polarisation filter.
"""

import math

import numpy as np
import pytest

from fdtdx_amd import modes as M

C0 = 299792458.0


def _lam(a, n):
    return 4.0 * math.sin((2 * a + 1) * math.pi / (2 * (2 * n + 1))) ** 2


@pytest.mark.parametrize("n1,n2", [(9, 7), (24, 22)])
def test_homogeneous_cross_section_has_the_closed_form_spectrum(n1, n2):
    eps = 4.0
    ie = np.full((1, 1, n1, n2), 1 / eps)
    dx, f = 50e-9, C0 / 1.55e-6
    k0 = 2 * math.pi * f * dx / C0
    ms = M.solve_modes(ie, 1.0, f, dx, 12)
    ref = sorted([k0 * k0 * eps - _lam(a, n1) - _lam(b, n2) for a in range(n1) for b in range(n2)] * 2, reverse=True)
    got = np.real([(m.neff * k0) ** 2 for m in ms])
    np.testing.assert_allclose(got, ref[:12], rtol=0, atol=1e-9 * k0 * k0 * eps)


def _curl_residuals(s, E, H, neff, ie, im, k0, L=6):
    from oracle import fdtd_oracle as O

    beta = 2 * math.asin((neff * k0).real / 2)
    E3 = np.repeat(E, L, axis=s + 1).astype(np.complex128)
    H3 = np.repeat(H, L, axis=s + 1).astype(np.complex128)
    w = np.arange(L).reshape([L if a == s else 1 for a in range(3)])
    for c in range(3):
        E3[c] = E3[c] * np.exp(-1j * beta * (w + (0.5 if c == s else 0.0)))
        H3[c] = H3[c] * np.exp(-1j * beta * (w + (0.0 if c == s else 0.5)))
    cE, _ = O.curl_E_raw(E3)
    cH, _ = O.curl_H_raw(H3)
    eps3 = 1 / np.repeat(ie, L, axis=s + 1)
    mu3 = 1 / np.repeat(im, L, axis=s + 1)
    inner = [slice(None)] * 3
    inner[s] = slice(1, L - 1)
    inner = tuple(inner)
    sc = np.abs(E3).max() * k0
    r1 = max(np.abs(cE[c][inner] + 1j * k0 * mu3[c][inner] * H3[c][inner]).max() for c in range(3)) / sc
    r2 = max(np.abs(cH[c][inner] - 1j * k0 * eps3[c][inner] * E3[c][inner]).max() for c in range(3)) / sc
    return r1, r2


@pytest.mark.parametrize("s", [0, 1, 2])
def test_mode_is_an_eigenmode_of_the_grid_curls(s):
    dx, f = 50e-9, C0 / 1.55e-6
    k0 = 2 * math.pi * f * dx / C0
    shp = [10, 12, 9]
    shp[s] = 1
    ie = np.full((3, *shp), 1 / 2.1)
    core = [slice(3, 7), slice(4, 9), slice(2, 6)]
    core[s] = slice(None)
    ie[(slice(None), *core)] = 1 / np.array([12.0, 11.0, 10.0])[:, None, None, None]
    im = np.full((3, *shp), 1.0)
    im[(slice(None), *core)] = 1 / np.array([1.2, 1.1, 1.3])[:, None, None, None]
    E, H, neff = M.compute_mode(f, ie, im, dx)
    assert E.shape == (3, *shp) and H.shape == (3, *shp)
    assert math.sqrt(2.1) < neff.real < math.sqrt(12.0 * 1.3)
    r1, r2 = _curl_residuals(s, E, H, neff, ie, im, k0)
    assert r1 < 1e-10 and r2 < 1e-10
    assert abs(M.integrated_power(E, H, s) - 1.0) < 1e-12


def test_slab_converges_to_the_analytic_te0_index():
    from scipy.optimize import brentq

    lam = 1.55e-6
    f = C0 / lam
    ncore, nclad, d = 3.4, 1.45, 0.3e-6
    k0 = 2 * math.pi / lam

    def f_te(n):  # symmetric slab, even TE mode: tan(kappa d / 2) = gamma / kappa
        kap = k0 * math.sqrt(ncore**2 - n**2)
        gam = k0 * math.sqrt(n**2 - nclad**2)
        return math.tan(kap * d / 2) - gam / kap

    lo = math.sqrt(max(nclad**2, ncore**2 - (math.pi / (k0 * d)) ** 2)) + 1e-9
    n_te = brentq(f_te, lo, ncore - 1e-9)
    errs = []
    for dx in (50e-9, 25e-9, 12.5e-9):
        nz, nc = int(round(3.0e-6 / dx)), int(round(d / dx))
        ie = np.full((1, 1, 2, nz), 1 / nclad**2)  # an axis of two cells: a 2-D cross-section
        z0 = nz // 2 - nc // 2
        ie[..., z0:z0 + nc] = 1 / ncore**2
        E, H, neff = M.compute_mode(f, ie, 1.0, dx, filter_pol="te")
        assert np.array_equal(E[:, :, 0], E[:, :, 1])
        assert np.abs(E[2]).max() < 1e-12 * np.abs(E[1]).max()  # TE: E along the slab, no Ez
        errs.append(neff.real / n_te - 1)
    assert errs[-1] < 5e-4
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0


def test_backward_mode_is_the_reciprocal_of_the_forward_one():
    dx, f = 50e-9, C0 / 1.55e-6
    ie = np.full((1, 1, 12, 10), 1 / 2.1)
    ie[0, 0, 4:8, 3:6] = 1 / 12.1
    Ef, Hf, nf = M.compute_mode(f, ie, 1.0, dx, direction="+")
    Eb, Hb, nb = M.compute_mode(f, ie, 1.0, dx, direction="-")
    assert nf == nb
    np.testing.assert_array_equal(Eb[1:], Ef[1:])
    np.testing.assert_array_equal(Eb[0], -Ef[0])
    np.testing.assert_array_equal(Hb[1:], -Hf[1:])
    np.testing.assert_array_equal(Hb[0], Hf[0])
    a1, a2 = 1, 2
    flux = lambda E, H: float(np.sum(0.5 * np.real(np.conj(E[a1]) * H[a2] - np.conj(E[a2]) * H[a1])))
    assert abs(flux(Ef, Hf) - 1) < 1e-12 and abs(flux(Eb, Hb) + 1) < 1e-12


def test_polarisation_filter_orders_te_and_tm():
    dx, f = 50e-9, C0 / 1.55e-6
    ie = np.full((1, 1, 20, 14), 1 / 2.1)
    ie[0, 0, 5:15, 5:9] = 1 / 12.1  # wide along y, thin along z: the y-polarised mode leads
    E0, _, n0 = M.compute_mode(f, ie, 1.0, dx)
    Et, _, nt = M.compute_mode(f, ie, 1.0, dx, filter_pol="te")
    Em, _, nm = M.compute_mode(f, ie, 1.0, dx, filter_pol="tm")
    assert n0 == nt and nm.real < nt.real
    share = lambda E, c: np.sum(np.abs(E[c]) ** 2) / np.sum(np.abs(E[1]) ** 2 + np.abs(E[2]) ** 2)
    assert share(Et, 1) > 0.5 and share(Em, 2) > 0.5


def test_mode_source_lowers_the_solved_mode():
    """ModePlaneSource.apply: the injected cross-section is the real part of the mode of the source
    plane, the Yee offsets are half a cell at c0 / neff (H components transverse to x, E along x),
    and caller-supplied mode fields replace the solve."""
    import scenarios

    import fdtdx_amd as fx

    spec = scenarios.waveguide_spec(T=20)
    objects, arrays, params, config = scenarios.build(spec)
    src = next(o for o in objects.sources)
    E, H, neff = scenarios.with_oracle_mode(spec)["sources"][0]["mode"]
    np.testing.assert_array_equal(src._E, np.real(E).astype(np.float32))
    np.testing.assert_array_equal(src._H, np.real(H).astype(np.float32))
    assert src._neff == neff
    dt = config.time_step_duration
    tau = 0.5 * 50e-9 / (C0 / neff.real * dt)
    assert np.allclose(src._tau_E[0], -tau) and not src._tau_E[1:].any()
    assert np.allclose(src._tau_H[1:], -tau) and not src._tau_H[0].any()
    spec2 = dict(spec, sources=[dict(spec["sources"][0], mode_fields=(2 * E, 2 * H, neff))])
    objects2, *_ = scenarios.build(spec2)
    src2 = next(o for o in objects2.sources)
    np.testing.assert_array_equal(src2._E, np.real(2 * E).astype(np.float32))
    assert isinstance(src2, fx.ModePlaneSource)
