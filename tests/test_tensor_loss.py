"""Full-tensor conductivity (the A/B update matrices, fdtd/misc.py:69-129, update.py:356-492,
752-822) in the oracle, pinned by what the physics fixes:

  - a uniform field in a periodic box has zero curl, so Ohm's law eps dE/dt = -sigma E leaves
    dE/dt = -c eta0 inv_eps sigma E (grid units), and the update is its trapezoidal (Crank-
    Nicolson) step: along each eigenvector of inv_eps sigma the field decays by
    (1 - f lam) / (1 + f lam) per step, f = c eta0 / 2 (1 / eta0 for H). The eigenvectors come from
    numpy's eigensolver, not from the oracle's linear solves; with non-commuting inv_eps and sigma a
    transposed product (sigma inv_eps) fails it;
  - the reverse matrices undo that decay exactly;
  - a diagonal permittivity and conductivity written as full tensors give the diagonal lossy
    update (Schneider 3.12), which the conductivity tests pin on their own.
"""

import numpy as np
import pytest

from oracle import fdtd_oracle as O
from oracle import scenes as S

INV_EPS = np.array([[0.5, 0.1, 0.0], [0.1, 0.4, 0.05], [0.0, 0.05, 0.3]])
SIGMA = np.array([[3.0, 1.0, -0.5], [1.0, 2.0, 0.8], [-0.5, 0.8, 2.5]]) * 1e4


def _uniform_scene(which, sig, inv):
    n = 4
    sc = O.Scene(shape=(n, n, n), dx=50e-9, dtype=np.float64)
    sc.periodic = (True, True, True)
    full = lambda m: np.broadcast_to(m.reshape(9, 1, 1, 1), (9, n, n, n)).copy()
    spacing = O.C0 * sc.dt / sc.c  # the conductivity scaling of the arrays (initialization.py:786)
    if which == "E":
        sc.inv_eps = full(inv)
        sc.sigma_E = full(sig * spacing)
    else:
        sc.inv_eps = np.ones((1, n, n, n))
        sc.inv_mu = full(inv)
        sc.sigma_H = full(sig * 1e5 * spacing)
    return sc


def _decay(sc, which, inv, sig, steps):
    spacing = O.C0 * sc.dt / sc.c
    M = inv @ (sig * (1 if which == "E" else 1e5) * spacing)
    f = sc.c * (O.ETA0 if which == "E" else 1 / O.ETA0) / 2
    lam, V = np.linalg.eig(M)
    g = ((1 - f * lam) / (1 + f * lam)) ** steps
    return np.real(V @ np.diag(g) @ np.linalg.inv(V)), M


@pytest.mark.parametrize("which", ["E", "H"])
def test_uniform_field_decays_by_the_trapezoidal_rule(which):
    sc = _uniform_scene(which, SIGMA, INV_EPS)
    st = O.init_state(sc)
    v0 = np.array([1.0, -0.4, 0.7])
    F = st.E if which == "E" else st.H
    F[:] = v0.reshape(3, 1, 1, 1)
    T = 25
    for t in range(T):
        O.forward_step(st, sc, t)
    G, M = _decay(sc, which, INV_EPS, SIGMA, T)
    want = G @ v0
    got = (st.E if which == "E" else st.H)[:, 1, 2, 3]
    assert np.abs(want - v0).max() > 0.05  # a real decay, coupled between components
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    other = st.H if which == "E" else st.E
    assert np.abs(other).max() < 1e-12  # a uniform field has no curl
    # the transposed product (sigma inv_eps) decays differently
    lam, V = np.linalg.eig(M.T)
    f = sc.c * (O.ETA0 if which == "E" else 1 / O.ETA0) / 2
    wrong = np.real(V @ np.diag(((1 - f * lam) / (1 + f * lam)) ** T) @ np.linalg.inv(V)) @ v0
    assert np.abs(wrong - want).max() > 1e-3


def test_reverse_matrices_undo_the_decay():
    sc = _uniform_scene("E", SIGMA, INV_EPS)
    st = O.init_state(sc)
    v0 = np.array([0.3, 1.0, -0.2])
    st.E[:] = v0.reshape(3, 1, 1, 1)
    for t in range(10):
        O.forward_step(st, sc, t)
    for t in range(9, -1, -1):
        O.reverse_step(st, sc, t, [])
    np.testing.assert_allclose(st.E[:, 0, 0, 0], v0, rtol=0, atol=1e-12)


def test_diagonal_tensors_equal_the_diagonal_lossy_update():
    import scenarios

    spec = scenarios.scenario_c1(T=60, n=20, pml=4, cube=6)
    spec["boxes"] = [dict(slc=((6, 14), (5, 15), (7, 13)), eps=(2.0, 3.0, 4.0), sigma_E=(3e4, 1e4, 2e4))]
    spec["detectors"] = []
    sc_d = S.build_scene(spec, dtype=np.float64)
    assert sc_d.sigma_E.shape[0] == 3 and sc_d.inv_eps.shape[0] == 3
    sc_t = S.build_scene(spec, dtype=np.float64)
    sc_t.inv_eps = O.expand3x3(sc_d.inv_eps, np.float64)
    sc_t.sigma_E = O.expand3x3(sc_d.sigma_E, np.float64)
    st_d, _ = O.run(sc_d, spec["T"])
    st_t, _ = O.run(sc_t, spec["T"])
    scale = np.abs(st_d.E).max()
    assert scale > 0
    np.testing.assert_allclose(st_t.E, st_d.E, rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(st_t.H, st_d.H, rtol=0, atol=1e-11 * np.abs(st_d.H).max())


def test_full_conductivity_tensor_lowering():
    import scenarios

    from fdtdx_amd import fdtd

    spec = scenarios.tensor_loss_spec(T=20)
    objects, arrays, params, config = scenarios.build(spec)
    low = fdtd.lower(objects, config, arrays)
    assert low["eps_comps"] == 9 and low["mu_comps"] == 9
    assert arrays.electric_conductivity.shape[0] == 9 and arrays.magnetic_conductivity.shape[0] == 9
    sc = S.build_scene(spec)
    np.testing.assert_allclose(arrays.electric_conductivity.cpu().numpy(), sc.sigma_E, rtol=1e-6)
    np.testing.assert_allclose(arrays.magnetic_conductivity.cpu().numpy(), sc.sigma_H, rtol=1e-6)


# ---------------------------------------------------------------- ADE poles in the full-tensor update
W0, GAMMA, DEPS = 2 * np.pi * 1.5e14, 2e13, 2.0


def _periodic_box_scene(n, tensor_c3, ie=1.0 / 1.5):
    """A uniform medium filling an n^3 periodic box: inv_eps ie (nine components) and one Lorentz pole
    whose coupling is tensor_c3 (row-major 3x3, already scaled by dt^2 / D)."""
    sc = O.Scene(shape=(n, n, n), dx=50e-9, dtype=np.float64)
    sc.periodic = (True, True, True)
    sc.inv_eps = O.expand3x3(np.full((1, n, n, n), ie), np.float64)
    k1, k2, _ = S.ade_coefficients([("lorentz", W0, GAMMA, DEPS)], sc.dt)
    c1 = np.full((1, 1, n, n, n), k1[0])
    c2 = np.full((1, 1, n, n, n), k2[0])
    c3 = np.broadcast_to(np.asarray(tensor_c3, np.float64).reshape(1, 9, 1, 1, 1), (1, 9, n, n, n)).copy()
    sc.ade = (c1, c2, c3)
    return sc


def _run_uniform(sc, e0, T=40):
    st = O.init_state(sc)
    st.E[:] = np.asarray(e0, np.float64).reshape(3, 1, 1, 1)
    out = []
    for t in range(T):
        O.forward_step(st, sc, t)
        out.append(st.E[:, 1, 2, 0].copy())
    return np.array(out), st


def test_oriented_pole_polarises_only_along_its_axis():
    """A uniform field in a uniform oriented medium (periodic box: no curl, every four-point mean of a
    uniform field is the field): the component along u evolves exactly as in the isotropic medium
    with the same pole, the components across u do not see the pole at all."""
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    k3 = S.ade_coefficients([("lorentz", W0, GAMMA, DEPS)], O.Scene(shape=(1, 1, 1), dx=50e-9).dt)[2][0]
    oriented = _periodic_box_scene(4, k3 * np.outer(u, u))
    iso = _periodic_box_scene(4, k3 * np.eye(3))
    along, _ = _run_uniform(oriented, u)
    ref, _ = _run_uniform(iso, u)
    assert np.abs(ref - ref[0]).max() > 1e-3  # the pole acts
    np.testing.assert_allclose(along, ref, rtol=0, atol=1e-13)
    w = np.array([2.0, -1.0, 0.0]) / np.sqrt(5.0)  # across u
    across, st = _run_uniform(oriented, w)
    np.testing.assert_allclose(across, np.broadcast_to(w, across.shape), rtol=0, atol=1e-13)
    assert np.abs(st.P_curr).max() < 1e-20


def _coupling(sc, E):
    """The coupling term of tensor_ade alone (zero state, c1 = c2 = 0): -c (K' - K)."""
    c1, c2, c3 = sc.ade
    sc2 = O.Scene(shape=sc.shape, dx=sc.dx, dtype=sc.dtype)
    sc2.periodic = sc.periodic
    sc2.ade = (np.zeros_like(c1), np.zeros_like(c2), c3)
    st = O.init_state(sc2)
    st.E = E.copy()
    st.P_curr = np.zeros((c1.shape[0], 3, *sc.shape))
    st.P_prev = np.zeros_like(st.P_curr)
    K0 = np.zeros((3, *sc.shape))
    return -sc.c * O.tensor_ade(st, sc2, K0)


@pytest.mark.parametrize("periodic", [False, True])
def test_oriented_coupling_is_symmetric_where_the_medium_varies(periodic):
    """With a symmetric coupling tensor u u^T the symmetrized off-diagonal weights make the coupling
    operator self-adjoint, <C E, F> = <E, C F>, also across the medium's edge (a plain
    c(r) avg(E) would not be); the zero or wrapped halo keeps it so at the domain faces."""
    rng = np.random.default_rng(3)
    n = (7, 6, 5)
    sc = O.Scene(shape=n, dx=50e-9, dtype=np.float64)
    sc.periodic = (periodic,) * 3
    u = np.array([0.3, -0.8, 0.52])
    u /= np.linalg.norm(u)
    c3 = np.zeros((1, 9, *n))
    c3[0, :, 1:5, 0:4, 2:5] = (np.outer(u, u).reshape(9, 1, 1, 1) * rng.uniform(0.5, 1.5, (1, 4, 4, 3)))
    sc.ade = (np.zeros((1, 1, *n)), np.zeros((1, 1, *n)), c3)
    E, F = rng.standard_normal((2, 3, *n))
    CE, CF = _coupling(sc, E), _coupling(sc, F)
    lhs, rhs = np.sum(CE * F), np.sum(E * CF)
    assert abs(lhs) > 1e-3
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # the one-sided form c(r) avg(E_j) breaks the symmetry
    def naive(E):
        rows = []
        for i in range(3):
            row = c3[0, 4 * i] * E[i]
            for j in range(3):
                if j != i:
                    row = row + c3[0, 3 * i + j] * O._avg_E1(E[j], j, i, sc.periodic)
            rows.append(row)
        return np.stack(rows)
    assert abs(np.sum(naive(E) * F) - np.sum(E * naive(F))) > 1e-6 * abs(lhs)


def test_axis_aligned_ade_in_a_diagonal_tensor_equals_the_diagonal_update():
    """Isotropic poles in a diagonal permittivity written as a full tensor: the polarisation current
    folded into the curl (delta / c, then B = c inv_eps) is the diagonal update's inv_eps * delta."""
    import scenarios

    spec = scenarios.scenario_c4(T=60, n=24, pml=4, slab=(10, 10, 4), cube=4)
    spec["detectors"] = []
    sc_d = S.build_scene(spec, dtype=np.float64)
    assert sc_d.ade is not None and sc_d.inv_eps.shape[0] == 3
    sc_t = S.build_scene(spec, dtype=np.float64)
    sc_t.inv_eps = O.expand3x3(sc_d.inv_eps, np.float64)
    st_d, _ = O.run(sc_d, spec["T"])
    st_t, _ = O.run(sc_t, spec["T"])
    scale = np.abs(st_d.E).max()
    assert scale > 0
    np.testing.assert_allclose(st_t.E, st_d.E, rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(st_t.P_curr, st_d.P_curr, rtol=0, atol=1e-10 * np.abs(st_d.P_curr).max())


def test_conductivity_tensor_gradient_matches_finite_differences():
    """The gradient through a full-tensor box with a full electric conductivity tensor: the adjoint
    carries mu_E through A^T (transposed four-point means), the curl's cotangent through B^T, and
    dL/d inv_eps through the matrices' dependence on it (FP64 central differences: design voxels
    through rho, and diagonal and off-diagonal inv_eps entries inside the conducting box). The
    transposed steps run at the stored forward states: the reverse matrices undo a step only where
    the field is uniform (the averages do not commute with A), so the reconstructed states of the
    reversible scheme drift inside the conducting box."""
    import scenarios

    spec = scenarios.scenario_c3(T=90, shape=(24, 22, 20), pml=4, design=(6, 6, 4))
    spec["boxes"] = [dict(slc=((5, 9), (6, 16), (6, 14)), eps=scenarios.SYM_TENSOR_EPS,
                          sigma_E=scenarios.SIGMA_E_TENSOR)]
    sc = S.build_scene(spec, np.float64)
    assert sc.inv_eps.shape[0] == 9 and sc.sigma_E.shape[0] == 9
    grad, ds, _ = O.reversible_gradient(sc, spec["T"], lambda ds: [np.ones_like(ds[0])], exact_states=True)
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
    for comp, cell in ((0, (7, 10, 9)), (1, (6, 12, 10)), (5, (8, 9, 8)), (7, (7, 14, 12))):
        vals = []
        for sgn in (1, -1):
            s2 = S.build_scene(spec, np.float64)
            s2.inv_eps[(comp, *cell)] += sgn * 1e-4
            _, dd = O.run(s2, spec["T"])
            vals.append(np.sum(dd[0]))
        fd = (vals[0] - vals[1]) / 2e-4
        assert abs(grad[(comp, *cell)] - fd) <= 1e-3 * abs(fd), (comp, cell, grad[(comp, *cell)], fd)
