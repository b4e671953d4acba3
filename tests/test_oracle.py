"""Pin the CPU oracle: against the fixtures tdgl3d itself wrote, against
closed forms and textbook solvers, and by planting mistakes that must fail."""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

import specs as S
from conftest import golden
from helpers import oracle_film, rand_state, rel_err
from oracle import tdgl_oracle as O

# ---------------------------------------------------------------------------
# index maps, holes, material (bit-exact, ordered)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("name", list(S.INDEX_CASES))
def test_indices_match_reference(name):
    g = golden(f"indices_{name}")
    f = oracle_film(S.INDEX_CASES[name])
    for k in g.files:
        ref = g[k]
        got = getattr(f.material, k) if k in ("sc_mask", "interior_sc_mask", "kappa") else f.idx[k]
        assert np.array_equal(got, ref), k


def test_film128_hole_masks_and_digests():
    g = golden("indices_film128_hole")
    f = oracle_film(S.FILM128)
    for k in g.files:
        if k.startswith("hole_"):
            assert np.array_equal(f.idx[k], g[k]), k
    for k, dig in zip(g["digest_keys"], g["digest_vals"]):
        got = hashlib.sha256(np.ascontiguousarray(np.asarray(f.idx[str(k)], np.int64)).tobytes())
        assert got.hexdigest() == str(dig), k
    # hand-derived counts: hole nodes i, j in 49..80 (32 x 32) over k = 0..16
    assert f.idx["hole_x_bc_mask"].size == 2 * 32 * 17
    assert f.idx["hole_y_bc_mask"].size == 2 * 32 * 17
    assert f.idx["hole_z_bc_mask"].size == 0
    assert f.idx["hole_x_bc_interior"].size == 2 * 32 * 15
    assert int(np.sum(f.material.sc_mask == 0.0)) == int(g["n_sc_zero"]) == 32 * 32 * 17


def _winding_inside(x, y, verts):
    """Independent point-in-polygon (winding number) for points off the edges."""
    w = 0
    n = len(verts)
    for s in range(n):
        (ax, ay), (bx, by) = verts[s], verts[(s + 1) % n]
        cross = (bx - ax) * (y - ay) - (x - ax) * (by - ay)
        if ay <= y < by and cross > 0:
            w += 1
        elif by <= y < ay and cross < 0:
            w -= 1
    return w != 0


def test_point_in_polygon_against_winding_number():
    rng = np.random.default_rng(7)
    polys = [[(0.0, 0.0), (10.0, 0.0), (5.0, 10.0)],
             [(0.0, 0.0), (5.0, 0.0), (5.0, 5.0), (10.0, 5.0), (10.0, 10.0), (0.0, 10.0)],
             [(1.3, 0.7), (8.2, 2.1), (6.4, 9.3), (3.3, 4.4), (0.5, 8.8)]]
    for verts in polys:
        for x, y in rng.uniform(-1, 11, size=(400, 2)):
            assert O.point_in_polygon(x, y, verts) == _winding_inside(x, y, verts)


def test_boundary_links_brute_force():
    rng = np.random.default_rng(3)
    mask = rng.random((6, 5, 4)) < 0.4
    for d, (di, dj, dk) in zip("xyz", [(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
        got = set(O.identify_boundary_links(mask, d).tolist())
        want = set()
        for i in range(6 - di):
            for j in range(5 - dj):
                for k in range(4 - dk):
                    if mask[i, j, k] != mask[i + di, j + dj, k + dk]:
                        want.add(i + 6 * j + 30 * k)
        assert got == want


# ---------------------------------------------------------------------------
# eval_f: golden vectors, CSR vs stencil, closed forms
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("name", list(S.EVALF_CASES))
def test_eval_f_matches_reference(name):
    d = golden(f"evalf_{name}")
    f = oracle_film(S.EVALF_CASES[name])
    u = O.boundary_vectors(*d["field"], f.g, f.idx)
    F = O.eval_f_csr(d["X"], f.g, f.idx, u, f.material)
    assert np.array_equal(F, d["F"])  # bit-exact: same numpy/scipy operations
    Fs = O.eval_f_stencil(d["X"], f.g, f.idx, u, f.material)
    assert rel_err(Fs, d["F"]) < 1e-15


@pytest.mark.parametrize("name", list(S.EVALF_CASES))
def test_jvp_stencil_central_differences(name):
    """jvp_stencil is the exact derivative of eval_f (rhs.py:147-239) along P:
    the central difference (F(X + hP) - F(X - hP)) / 2h of the reference's own
    eval_f (bit-exact against tdgl3d above) converges to it at second order.
    A dropped or mis-signed term leaves an O(1) gap that does not shrink with h."""
    d = golden(f"evalf_{name}")
    f = oracle_film(S.EVALF_CASES[name])
    u = O.boundary_vectors(*d["field"], f.g, f.idx)
    X = d["X"]
    n = f.g.n_interior
    nphi = 3 if f.g.is_3d else 2
    rng = np.random.default_rng(100 + len(name))
    P = np.concatenate([rng.standard_normal(n) + 1j * rng.standard_normal(n),
                        rng.standard_normal(nphi * n).astype(np.complex128)])
    jp = O.jvp_stencil(X, P, f.g, f.idx, u, f.material)
    scale = float(np.max(np.abs(jp)))

    def gap(h):
        fd = (O.eval_f_csr(X + h * P, f.g, f.idx, u, f.material)
              - O.eval_f_csr(X - h * P, f.g, f.idx, u, f.material)) / (2.0 * h)
        return float(np.max(np.abs(fd - jp)))

    e1, e2 = gap(1e-3), gap(5e-4)
    assert e2 <= 1e-6 * scale, (e1, e2, scale)
    assert 3.5 <= e1 / e2 <= 4.5, (e1, e2)
    # the hole rows of the link blocks are exactly zero, as eval_f's are
    for q, ax in enumerate("xyz"):
        hm = f.idx[f"hole_{ax}_bc_interior"]
        if hm.size and q < nphi:
            assert np.all(jp[(q + 1) * n + hm] == 0.0)


def test_planted_jvp_mistake_is_caught(monkeypatch):
    """Dropping the conj(d psi) part of d[(1-|psi|^2) psi] in the oracle's JVP
    must fail the central-difference pin."""
    orig = O.jvp_stencil
    d = golden("evalf_c1_bz")
    f = oracle_film(S.EVALF_CASES["c1_bz"])
    u = O.boundary_vectors(*d["field"], f.g, f.idx)
    n = f.g.n_interior
    rng = np.random.default_rng(1)
    X = d["X"]
    P = np.concatenate([rng.standard_normal(n) + 1j * rng.standard_normal(n),
                        rng.standard_normal(3 * n).astype(np.complex128)])
    jp = orig(X, P, f.g, f.idx, u, f.material)
    c = X[:n]
    wrong = jp.copy()
    wrong[:n] += c * c * np.conj(P[:n])  # undo the -psi^2 conj(d psi) term
    h = 5e-4
    fd = (O.eval_f_csr(X + h * P, f.g, f.idx, u, f.material)
          - O.eval_f_csr(X - h * P, f.g, f.idx, u, f.material)) / (2.0 * h)
    assert np.max(np.abs(fd - jp)) <= 1e-6 * np.max(np.abs(jp))
    assert np.max(np.abs(fd - wrong)) > 1e-2 * np.max(np.abs(jp))


def test_fpsi_closed_form():
    g = O.Grid(6, 6, 1, kappa=2.0)
    idx = O.construct_indices(g)
    x = np.ones(g.dim_x, dtype=np.complex128)
    assert np.max(np.abs(O._fpsi(x, g, idx, None))) < 1e-14
    assert np.allclose(O._fpsi(0.5 * x, g, idx, None), 0.375, atol=1e-14)


def test_zero_gauge_laplacian_and_lphi_diagonal():
    g = O.Grid(5, 5, 5, kappa=3.0)
    idx = O.construct_indices(g)
    m = idx["interior_to_full"]
    L = O._lpsi(np.zeros(g.dim_x, np.complex128), "x", g, idx)
    assert np.allclose(np.asarray(L[m, m]).ravel(), -2.0)
    assert np.allclose(np.asarray(L[m, m - 1]).ravel(), 1.0)
    P = O._lphi("x", g, idx, None)
    assert np.allclose(np.asarray(P[m, m]).ravel(), -2.0 * (9.0 + 9.0))


def test_uniform_periodic_equilibrium():
    g = O.Grid(6, 6, 1, periodic_x=True, periodic_y=True, periodic_z=True)
    idx = O.construct_indices(g)
    X = np.zeros(g.n_state, np.complex128)
    X[:g.n_interior] = 1.0
    u = O.boundary_vectors(0, 0, 0, g, idx)
    assert np.max(np.abs(O.eval_f_csr(X, g, idx, u))) < 1e-12
    assert np.max(np.abs(O.eval_f_stencil(X, g, idx, u))) < 1e-12


def test_c4_antisymmetry_with_bz():
    g = O.Grid(8, 8, 1, 1.0, 1.0, 1.0, kappa=2.0)
    idx = O.construct_indices(g)
    X = np.zeros(g.n_state, np.complex128)
    X[:g.n_interior] = 1.0
    F = O.eval_f_csr(X, g, idx, O.boundary_vectors(0, 0, 0.5, g, idx))
    n = g.n_interior
    px = F[n:2 * n].reshape(7, 7)
    py = F[2 * n:3 * n].reshape(7, 7)
    assert np.allclose(px, -py.T, atol=1e-12)


def test_gauge_covariance_of_dpsi():
    """psi -> psi e^{i chi}, phi_a(m) -> phi_a(m) - (chi(m+a) - chi(m)) multiplies
    dpsi by e^{i chi} (the Peierls-phase Laplacian and (1-|psi|^2) psi are
    covariant); chi vanishes on the ghost shell."""
    g = O.Grid(9, 8, 7, 0.5, 0.5, 0.5, kappa=2.0)
    idx = O.construct_indices(g)
    rng = np.random.default_rng(2)
    n = g.n_interior
    X = rand_state(n, 3, 5)
    chi = np.zeros(g.full_shape)
    chi[1:-1, 1:-1, 1:-1] = rng.standard_normal((g.Nz - 1, g.Ny - 1, g.Nx - 1)) * 0.3
    chi = chi.reshape(-1)
    u = O.boundary_vectors(0, 0, 0, g, idx)
    m = idx["interior_to_full"]
    Y = X.copy()
    Y[:n] = X[:n] * np.exp(1j * chi[m])
    for q, s in enumerate((1, g.mj, g.mk)):
        Y[(q + 1) * n:(q + 2) * n] = X[(q + 1) * n:(q + 2) * n] - (chi[m + s] - chi[m])
    # zero-current ghosts copy psi across a zeroed normal link, so the boundary
    # rows are covariant too: every interior node transforms
    for ev in (O.eval_f_stencil, O.eval_f_csr):
        F0, F1 = ev(X, g, idx, u), ev(Y, g, idx, u)
        assert np.allclose(F1[:n], F0[:n] * np.exp(1j * chi[m]), atol=1e-12)
    # a hole's zeroed links break it next to the hole
    f = O.Film(g)
    f.add_hole([(1.2, 1.2), (2.8, 1.2), (2.8, 2.6), (1.2, 2.6)])
    F0 = O.eval_f_csr(X, g, f.idx, u, f.material)
    F1 = O.eval_f_csr(Y, g, f.idx, u, f.material)
    assert not np.allclose(F1[:n], F0[:n] * np.exp(1j * chi[m]), atol=1e-6)


def test_bfield_zero_for_zero_links():
    g = O.Grid(6, 6, 1)
    idx = O.construct_indices(g)
    X = np.zeros(g.n_state, np.complex128)
    X[:g.n_interior] = 1.0
    assert np.max(np.abs(O.bfield_full(X, g, idx)[2])) == 0.0


# ---------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------


def test_tgcr_and_newton_match_reference():
    case = S.EVALF_CASES["c1_bz"]
    f = oracle_film(case)
    u = O.boundary_vectors(0, 0, 0.3, f.g, f.idx)

    def F(Y):
        return O.eval_f_csr(Y, f.g, f.idx, u)

    d = golden("tgcr_trap")
    st = {}
    dx = O.tgcr_trap(F, d["x_lin"], d["b"], float(d["dt"]), tol=1e-8, eps_mf=1e-4, stats=st)
    assert np.array_equal(dx, d["dx"])
    assert st["tgcr_iters"] == int(d["n_iter"])
    d = golden("newton_trap")
    x, conv, it = O.newton_trap(F, d["x0"], d["gamma"], float(d["dt"]), 1e-3, 1e-3, 20, 1e-8, 1e-4)
    assert np.array_equal(x, d["x"]) and conv == bool(d["converged"]) and it == int(d["iters"])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tgcr_solves_dense_linear_systems(seed):
    """For a linear f(x) = A x the FD product is exact up to rounding, so TGCR
    must return (I - dt/2 A)^{-1} b."""
    rng = np.random.default_rng(seed)
    N = 12
    A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    dt = 0.3
    x = O.tgcr_trap(lambda v: A @ v, np.zeros(N, complex), b, dt, tol=1e-12, max_iter=400)
    # TGCR works in the real inner product, i.e. on R^{2N}: it converges within 2N steps
    want = np.linalg.solve(np.eye(N) - dt / 2 * A, b)
    assert np.allclose(x, want, rtol=1e-7, atol=1e-8)


def test_tgcr_reference_smoke_cases():
    b = np.array([1.0, 2.0], complex)
    assert np.allclose(O.tgcr_trap(lambda x: x, np.zeros(2, complex), b, dt=0.0, tol=1e-8), b, atol=1e-5)
    assert O.tgcr_trap(lambda x: x, np.zeros(2, complex), b, dt=0.1, max_iter=0).size == 0
    x0 = np.array([1.0, 2.0], complex)
    x, conv, _ = O.newton_trap(lambda x: np.zeros_like(x), x0, x0.copy(), 1.0, 1e-6, 1e-6)
    assert conv and np.allclose(x, x0)


def test_trapezoidal_hole_trajectory_matches_reference():
    d = golden("trap_hole")
    case = S.TRAP_HOLE
    f = oracle_film(case)
    af = case["field"]

    def field_at(t):
        return O.applied_field_at(t, float(d["t_stop"]), (0.0, 0.0, af["Bz"]))

    log = []
    times, hist = O.trapezoidal(d["x0"], f, field_at, 0.0, float(d["t_stop"]), case["dt"],
                                tol_gcr=case["tol_gcr"], log=log)
    assert np.array_equal(times, d["times"])
    assert np.array_equal(hist[:, d["sel"]], d["states"])
    assert [(c, it) for c, it, _, _ in log] == [(bool(c), int(it)) for c, it, _ in d["newton"]]
    tg = [k for *_, ks in log for k in ks]
    assert tg == [int(k) for k, _ in d["tgcr"]]


def test_extrapolating_predictor_moves_only_newtons_start():
    """predictor="extrapolate" (an extension with no counterpart in the
    reference) starts Newton from X + dt/dt_prev (X - X_prev).  The step
    equations and the convergence test are the reference's, so: the first step
    (no history yet) takes the reference's Newton iterations, every step is
    accepted as in the reference, the trajectory stays within 1e-7 of the
    reference's states (the loose Newton tolerance's reach), and the
    extrapolated start saves Newton iterations on this film."""
    d = golden("trap_hole")
    case = S.TRAP_HOLE
    f = oracle_film(case)
    af = case["field"]

    def field_at(t):
        return O.applied_field_at(t, float(d["t_stop"]), (0.0, 0.0, af["Bz"]))

    log = []
    times, hist = O.trapezoidal(d["x0"], f, field_at, 0.0, float(d["t_stop"]), case["dt"],
                                tol_gcr=case["tol_gcr"], log=log, predictor="extrapolate")
    ref = [(bool(c), int(it)) for c, it, _ in d["newton"]]
    assert np.array_equal(times, d["times"])
    assert [c for c, *_ in log] == [c for c, _ in ref]
    assert log[0][1] == ref[0][1]
    err = rel_err(hist[:, d["sel"]], d["states"])
    assert 0 < err < 1e-7
    assert sum(it for _, it, _, _ in log) < 0.8 * sum(it for _, it in ref)
    with pytest.raises(ValueError):
        from nanowire_tdgl_amd.solvers import integrators
        integrators.trapezoidal(d["x0"], None, None, None, 0.0, 1.0, 0.1, predictor="midpoint")


def test_euler_and_observables_match_reference():
    d = golden("euler_small")
    case = S.EULER_SMALL
    f = oracle_film(case)
    B = (0.0, 0.0, case["field"]["Bz"])
    times, hist = O.forward_euler(d["x0"], f, lambda t: O.applied_field_at(t, case["t_stop"], B),
                                  0.0, case["t_stop"], case["dt"], save_every=case["save_every"])
    assert np.array_equal(times, d["times"])
    assert np.array_equal(hist, d["states"])
    for s in range(hist.shape[1]):
        u = O.boundary_vectors(*O.applied_field_at(times[s], times[-1], B), f.g, f.idx)
        assert np.array_equal(np.stack(O.bfield_full(hist[:, s], f.g, f.idx, u)), d["bfield"][s])
        assert np.array_equal(np.stack(O.supercurrent(hist[:, s], f.g, f.idx)), d["supercurrent"][s])


def test_observables_aniso():
    d = golden("obs_aniso")
    case = S.OBS_ANISO
    f = oracle_film(case)
    B = (case["field"]["Bx"], case["field"]["By"], case["field"]["Bz"])
    for key, t in (("b_t0", 0.0), ("b_t1", 0.5), ("b_tend", 1.0)):
        u = O.boundary_vectors(*O.applied_field_at(t, 1.0, B), f.g, f.idx)
        assert np.array_equal(np.stack(O.bfield_full(d["X"], f.g, f.idx, u)), d[key]), key
    assert np.array_equal(np.stack(O.supercurrent(d["X"], f.g, f.idx)), d["js"])


# ---------------------------------------------------------------------------
# planted mistakes: each must make a golden comparison fail
# ---------------------------------------------------------------------------


def _evalf_case_fails(name) -> bool:
    d = golden(f"evalf_{name}")
    f = oracle_film(S.EVALF_CASES[name])
    u = O.boundary_vectors(*d["field"], f.g, f.idx)
    return rel_err(O.eval_f_csr(d["X"], f.g, f.idx, u, f.material), d["F"]) > 1e-12


def test_planted_bc_sign_error_is_caught(monkeypatch):
    orig = O.apply_bc

    def wrong(x, y1, y2, y3, g, idx, u):
        Bx, By, Bz = u
        orig(x, y1, y2, y3, g, idx, (Bx, By, -Bz))  # flux sign flipped
    monkeypatch.setattr(O, "apply_bc", wrong)
    assert _evalf_case_fails("c1_bz")


def test_planted_missing_doubling_is_caught(monkeypatch):
    orig = O._lphi

    def wrong(axis, g, idx, mat):  # drop one of the two along-link Laplacians
        return orig(axis, g, idx, mat) * 0.75
    monkeypatch.setattr(O, "_lphi", wrong)
    assert _evalf_case_fails("aniso_bxyz")


def test_planted_missing_interior_normal_zeroing_is_caught(monkeypatch):
    """tdgl3d zeroes phi_x on the interior plane i = Nx-1 as well as on the ghost
    plane (indices.py:392-430); zeroing only the ghost plane must be caught."""
    orig = O.construct_indices

    def wrong(g):
        idx = orig(g)
        for ax, N, st in (("x", g.Nx, 1), ("y", g.Ny, g.mj), ("z", g.Nz, g.mk)):
            m = idx[f"{ax}_normal_bc_mask"]
            if m.size:
                coord = (m // st) % (N + 1)
                idx[f"{ax}_normal_bc_mask"] = m[coord == 0]
        return idx
    monkeypatch.setattr(O, "construct_indices", wrong)
    assert _evalf_case_fails("c1_bz")


def test_planted_tau_error_is_caught(monkeypatch):
    monkeypatch.setattr(O, "TAU_RELAX", 0.2)
    assert _evalf_case_fails("trilayer")
