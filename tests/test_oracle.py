"""Pins of the CPU oracle (the parity checker) that do not depend on the
oracle's own formulas: closed forms, brute-force linear algebra, exact
rational arithmetic and analytic solutions on the sphere."""
import math
from fractions import Fraction
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.linalg


def phys_with(pkg, **over):
    P = pkg.discretization.Phys()
    base = dict(Grav=P.Grav, Omega=P.Omega, Rd=P.Rd, Cpd=P.Cpd, p0=P.p0, kappa=P.kappa)
    base.update(over)
    return SimpleNamespace(**base)


def node_lonlat(disc):
    X = disc.X
    idx = disc.glob.reshape(-1) - 1
    first = np.full(disc.num_g, -1)
    first[idx[::-1]] = np.arange(idx.size)[::-1]
    e, i = np.divmod(first, disc.N * disc.N)
    x, y, z = X[e, 0, 0, 0, i], X[e, 0, 1, 0, i], X[e, 0, 2, 0, i]
    r = np.sqrt(x * x + y * y + z * z)
    return np.arctan2(y, x), np.arcsin(z / r)


_DISC = {}


def disc_for(pkg, ne, nz):
    key = (ne, nz)
    if key not in _DISC:
        _DISC[key] = pkg.discretization.build(ne, 4, nz, 30000.0)
    return _DISC[key]


def uniform_state(disc, rho=1.0, theta=300.0):
    U = np.zeros((5, disc.num_g, disc.nz))
    U[0] = rho
    U[4] = rho * theta
    return U


# --------------------------------------------------------------------------
# equation of state and column Jacobian

def test_eos_closed_form(pkg, oracle_mod):
    d = disc_for(pkg, 2, 4)
    ph = phys_with(pkg)
    o = oracle_mod.Oracle(d, ph, (0.0, 0.0, 0.0))
    rng = np.random.default_rng(0)
    U = uniform_state(d)
    U[0] = rng.uniform(0.2, 1.3, U[0].shape)
    U[4] = U[0] * rng.uniform(280, 500, U[0].shape)
    A = o.prepare(U)
    p = ph.p0 * (ph.Rd * U[4] / ph.p0) ** (1.0 / (1.0 - ph.kappa))
    assert np.abs(A[0] / p - 1).max() < 1e-13
    assert np.abs(A[1] / (p / (ph.Rd * U[0])) - 1).max() < 1e-13
    # potential temperature recovered from p and T is rho*theta / rho
    assert np.abs(A[2] / (U[4] / U[0]) - 1).max() < 1e-13
    assert np.all(A[3] == 0)


def test_column_jacobian(pkg, oracle_mod):
    d = disc_for(pkg, 2, 6)
    ph = phys_with(pkg)
    o = oracle_mod.Oracle(d, ph, (0.0, 0.0, 0.0))
    U = pkg.state.jet_state(d, seed=1)
    A = o.prepare(U)
    JRhoW, JWRho, JWRhoTh, JRhoThW = o.jac(U, A)
    dz = d.dz
    dzL, dzR = dz[:, :-1], dz[:, 1:]
    rhoF = (U[0][:, :-1] * dzL + U[0][:, 1:] * dzR) / (dzL + dzR)
    assert np.allclose(JRhoW[..., 0], -rhoF / dzL, rtol=1e-14)
    assert np.allclose(JRhoW[..., 1], rhoF / dzR, rtol=1e-14)
    assert np.allclose(JWRho[..., 0], -ph.Grav * dzR / rhoF / (dzL + dzR), rtol=1e-14)

    # dp/d(rho theta) in the Jacobian is (1 - kappa) times a centred
    # difference of p: the reference's dPresdRhoTh (src/Models/Equation.jl:33-36)
    # drops the factor 1/(1-kappa) of the exact derivative, which the W-method
    # tolerates (the Jacobian is an approximation there)
    def pres(rt):
        return ph.p0 * (ph.Rd * rt / ph.p0) ** (1.0 / (1.0 - ph.kappa))

    dpR = -JWRhoTh[..., 0] * rhoF * 0.5 * (dzL + dzR)
    rtR = U[4][:, 1:]
    h = 1e-4 * rtR
    fd = (1 - ph.kappa) * (pres(rtR + h) - pres(rtR - h)) / (2 * h)
    assert np.abs(dpR / fd - 1).max() < 1e-7
    dpL = JWRhoTh[..., 1] * rhoF * 0.5 * (dzL + dzR)
    rtL = U[4][:, :-1]
    fdl = (1 - ph.kappa) * (pres(rtL + 1e-4 * rtL) - pres(rtL - 1e-4 * rtL)) / (2e-4 * rtL)
    assert np.abs(dpL / fdl - 1).max() < 1e-7


def dense_column_system(blocks, ig, nz, fac):
    """(I/fac - J) of one column with unknowns [rho(nz), w(nz-1), rhotheta(nz)]
    assembled entry by entry from the Jacobian's meaning:
    d rho_i/dt = JRhoW[1,i] w_i + JRhoW[2,i-1] w_{i-1} (same for rho theta),
    d w_i/dt = JWRho[2,i] rho_i + JWRho[1,i] rho_{i+1} + JWRhoTh[...] likewise."""
    JRhoW, JWRho, JWRhoTh, JRhoThW = (b[ig] for b in blocks)
    n = nz - 1
    N = 2 * nz + n
    r = lambda i: i  # noqa: E731
    w = lambda i: nz + i  # noqa: E731
    t = lambda i: nz + n + i  # noqa: E731
    Jm = np.zeros((N, N))
    for i in range(n):
        Jm[r(i), w(i)] += JRhoW[i, 0]
        Jm[r(i + 1), w(i)] += JRhoW[i, 1]
        Jm[t(i), w(i)] += JRhoThW[i, 0]
        Jm[t(i + 1), w(i)] += JRhoThW[i, 1]
        Jm[w(i), r(i)] += JWRho[i, 1]
        Jm[w(i), r(i + 1)] += JWRho[i, 0]
        Jm[w(i), t(i)] += JWRhoTh[i, 1]
        Jm[w(i), t(i + 1)] += JWRhoTh[i, 0]
    return np.eye(N) / fac - Jm, r, w, t


def test_schur_solve_equals_dense_solve(pkg, oracle_mod):
    """SchurSolve (tri build + F + Thomas + B) solves (I/fac - J) k = v."""
    nz = 7
    d = disc_for(pkg, 2, nz)
    ph = phys_with(pkg)
    o = oracle_mod.Oracle(d, ph, (0.0, 0.0, 0.0))
    U = pkg.state.jet_state(d, seed=2)
    A = o.prepare(U)
    blocks = o.jac(U, A)
    rng = np.random.default_rng(3)
    v0 = rng.standard_normal(U.shape)
    fac = 150.0
    k, tri = o.schur(v0.copy(), blocks, fac)
    for ig in (0, 5, d.num_g - 1):
        M, r, w, t = dense_column_system(blocks, ig, nz, fac)
        rhs = np.concatenate([v0[0, ig], v0[3, ig, :-1], v0[4, ig]])
        sol = np.linalg.solve(M, rhs)
        got = np.concatenate([k[0, ig], k[3, ig, :-1], k[4, ig]])
        assert np.abs(got - sol).max() <= 1e-12 * np.abs(sol).max()
        assert k[3, ig, -1] == 0.0
        assert np.allclose(k[1, ig], fac * v0[1, ig], rtol=1e-15)
        # the Schur tri matrix is the one Thomas solved: compare with LAPACK banded
        tr = tri[ig]  # (nz-1, 3): [A[i-1,i], A[i,i], A[i+1,i]] column storage
        ab = np.zeros((3, nz - 1))
        ab[0, 1:] = tr[1:, 0]
        ab[1] = tr[:, 1]
        ab[2, :-1] = tr[:-1, 2]
        vv = v0.copy()
        rhs_w = vv[3, ig, :-1] / fac
        for i in range(nz - 1):
            JB = [b[ig] for b in blocks]
            rhs_w[i] += JB[1][i, 1] * vv[0, ig, i] + JB[1][i, 0] * vv[0, ig, i + 1] + JB[2][i, 1] * vv[4, ig, i] \
                + JB[2][i, 0] * vv[4, ig, i + 1]
        wb = scipy.linalg.solve_banded((1, 1), ab, rhs_w)
        assert np.abs(wb - k[3, ig, :-1]).max() <= 1e-12 * np.abs(wb).max()


# --------------------------------------------------------------------------
# Rosenbrock-W tableau

def test_ssp_knoth_tableau_exact(oracle_mod, pkg):
    a, c, m, gamma = oracle_mod.Oracle.ssp_knoth()
    F = Fraction
    alpha = [[F(0)] * 3 for _ in range(3)]
    alpha[1][0], alpha[2][0], alpha[2][1] = F(1), F(1, 4), F(1, 4)
    G = [[F(1), F(0), F(0)], [F(0), F(1), F(0)], [F(-3, 4), F(-3, 4), F(1)]]
    b = [F(1, 6), F(1, 6), F(2, 3)]
    # inverse of the unit lower-triangular Gamma
    Gi = [[F(1), F(0), F(0)], [F(0), F(1), F(0)], [F(3, 4), F(3, 4), F(1)]]
    for i in range(3):
        for j in range(3):
            assert sum(G[i][q] * Gi[q][j] for q in range(3)) == (1 if i == j else 0)
    ax = [[sum(alpha[i][q] * Gi[q][j] for q in range(3)) for j in range(3)] for i in range(3)]
    # m = Gamma^T \ b
    mx = [None, None, None]
    mx[2] = b[2]
    mx[1] = b[1] + F(3, 4) * mx[2]
    mx[0] = b[0] + F(3, 4) * mx[2]
    for i in range(3):
        for j in range(i):
            assert a[i, j] == float(ax[i][j])
            assert c[i, j] == float(-Gi[i][j])
        assert m[i] == float(mx[i])
    assert gamma == 1.0
    host_a, host_c, host_m, host_g = pkg.host.ssp_knoth()
    assert np.array_equal(host_a[np.tril_indices(3, -1)], a[np.tril_indices(3, -1)])
    assert np.array_equal(host_c[np.tril_indices(3, -1)], c[np.tril_indices(3, -1)])
    assert np.array_equal(host_m, m) and host_g == gamma


def _ros_step(y, dt, f, jac, a, c, m, gamma):
    """One W-step with the stage structure of RosenbrockSchur.jl on a scalar ODE."""
    k = [0.0, 0.0, 0.0]
    J = jac(y)
    for i in range(3):
        V = y + sum(a[i, j] * k[j] for j in range(i))
        fv = f(V) + sum(c[i, j] / dt * k[j] for j in range(i))
        k[i] = fv / (1.0 / (dt * gamma) - J)
    return y + sum(m[i] * k[i] for i in range(3))


@pytest.mark.parametrize("jac_kind", ["exact", "scaled"])
def test_ssp_knoth_second_order(oracle_mod, jac_kind):
    """Order 2 on y' = -y^2 with the exact Jacobian and with an approximate one
    (three times the exact): a wrong tableau entry drops the order."""
    a, c, m, gamma = oracle_mod.Oracle.ssp_knoth()
    f = lambda y: -y * y  # noqa: E731
    jac = (lambda y: -2 * y) if jac_kind == "exact" else (lambda y: -6 * y)
    exact = 1.0 / (1.0 + 1.0)  # y(1) with y(0) = 1
    errs = []
    for nst in (80, 160, 320):
        y = 1.0
        for _ in range(nst):
            y = _ros_step(y, 1.0 / nst, f, jac, a, c, m, gamma)
        errs.append(abs(y - exact))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(1.85 < r < 2.2 for r in rates), rates


def test_ssp_knoth_without_jacobian_is_ssprk3(oracle_mod):
    """With J = 0 the W-method is the explicit Shu-Osher SSP-RK3 step."""
    a, c, m, gamma = oracle_mod.Oracle.ssp_knoth()
    f = lambda y: math.sin(y) - 0.3 * y  # noqa: E731
    for y0, dt in ((0.7, 0.1), (-1.3, 0.37)):
        got = _ros_step(y0, dt, f, lambda y: 0.0, a, c, m, gamma)
        y1 = y0 + dt * f(y0)
        y2 = 0.75 * y0 + 0.25 * (y1 + dt * f(y1))
        want = y0 / 3 + 2 / 3 * (y2 + dt * f(y2))
        assert abs(got - want) < 1e-15


# --------------------------------------------------------------------------
# tendency operators against analytic solutions


def test_zero_wind_continuity_exact(pkg, oracle_mod):
    """u = v = w = 0: the mass and theta fluxes vanish identically."""
    d = disc_for(pkg, 3, 5)
    o = oracle_mod.Oracle(d, phys_with(pkg), (0.0, 0.0, 0.0))
    U = pkg.state.jet_state(d, seed=4)
    U[1:4] = 0.0
    A = o.prepare(U)
    F, _ = o.fcn(U, A)
    assert np.all(F[0] == 0.0) and np.all(F[4] == 0.0)


def test_rest_state_hydrostatic_balance(pkg, oracle_mod):
    """A resting atmosphere in discrete hydrostatic balance,
    p_{k+1} - p_k = -g dz (rho_k + rho_{k+1}) / 2, has zero tendency: the
    pressure gradient and gravity in the w equation cancel (a sign error in
    either leaves 2 g)."""
    nz = 10
    d = disc_for(pkg, 3, nz)
    ph = phys_with(pkg)
    o = oracle_mod.Oracle(d, ph, (1e15, 1e15, 1e15))
    dz = d.H / nz
    zc = (np.arange(nz) + 0.5) * dz
    rho = 1.2 * np.exp(-zc / 8000.0)
    p = np.empty(nz)
    p[0] = 1.0e5
    for k in range(nz - 1):
        p[k + 1] = p[k] - ph.Grav * dz * (rho[k] + rho[k + 1]) / 2
    rt = ph.p0 / ph.Rd * (p / ph.p0) ** (1 - ph.kappa)
    U = np.zeros((5, d.num_g, nz))
    U[0] = rho
    U[4] = rt
    A = o.prepare(U)
    assert np.abs(A[0] / p - 1).max() < 1e-13
    F, _ = o.fcn(U, A)
    assert np.abs(F[3]).max() < 1e-9 * ph.Grav
    assert np.abs(F[1:3]).max() < 1e-8
    assert np.all(F[0] == 0) and np.abs(F[4]).max() < 1e-8
    # and the Rosenbrock-W step keeps it at rest
    V = U.copy()
    o.rw_step(V, 300.0, 1)
    assert np.abs(V[3]).max() < 1e-7
    assert np.abs(V[0] / U[0] - 1).max() < 1e-12


def test_coriolis_term(pkg, oracle_mod):
    """F(Omega) - F(0) = (v, -u) 2 Omega sin(lat) at every node."""
    d = disc_for(pkg, 4, 3)
    U = pkg.state.jet_state(d, seed=5)
    rng = np.random.default_rng(6)
    U[2] = 10 * rng.standard_normal(U[2].shape)
    lon, lat = node_lonlat(d)
    Om = 7.292e-5
    Fs = []
    for om in (Om, 0.0):
        o = oracle_mod.Oracle(d, phys_with(pkg, Omega=om), (0.0, 0.0, 0.0))
        A = o.prepare(U)
        Fs.append(o.fcn(U, A)[0])
    dF = Fs[0] - Fs[1]
    f = 2 * Om * np.sin(lat)[:, None]
    scale = np.abs(f * U[1]).max()
    assert np.abs(dF[1] - U[2] * f).max() < 1e-10 * scale
    assert np.abs(dF[2] + U[1] * f).max() < 1e-10 * scale
    assert np.abs(dF[3]).max() < 1e-12 * scale


def _converge(errs, lo=1.5):
    return all(errs[i + 1] < errs[i] / lo for i in range(len(errs) - 1))


def test_solid_body_rotation_momentum(pkg, oracle_mod):
    """u = u0 cos(lat), v = w = 0, uniform rho and theta, no rotation or
    gravity: the vector-invariant momentum plus kinetic-energy gradient give
    du/dt = 0 and dv/dt = -u0^2 sin(lat) cos(lat) / R."""
    u0 = 40.0
    errs = []
    for ne in (3, 6, 12):
        d = disc_for(pkg, ne, 3)
        o = oracle_mod.Oracle(d, phys_with(pkg, Grav=0.0, Omega=0.0), (0.0, 0.0, 0.0))
        lon, lat = node_lonlat(d)
        U = uniform_state(d)
        U[1] = (u0 * np.cos(lat))[:, None]
        A = o.prepare(U)
        F, _ = o.fcn(U, A)
        ana = -(u0 ** 2) * np.sin(lat) * np.cos(lat) / d.rad
        errs.append(max(np.abs(F[2][:, 1] - ana).max(), np.abs(F[1][:, 1]).max()) / np.abs(ana).max())
        assert np.abs(F[3]).max() < 1e-9 * np.abs(ana).max()
    assert errs[-1] < 1e-3 and _converge(errs, 4.0), errs


def test_divergence_of_meridional_flow(pkg, oracle_mod):
    """v = V0 cos(lat) (u = w = 0), uniform rho: the continuity tendency
    d rho/dt = -div(rho v) = 2 rho V0 sin(lat) / R."""
    V0 = 10.0
    errs = []
    for ne in (3, 6, 12):
        d = disc_for(pkg, ne, 3)
        o = oracle_mod.Oracle(d, phys_with(pkg, Grav=0.0, Omega=0.0), (0.0, 0.0, 0.0))
        lon, lat = node_lonlat(d)
        U = uniform_state(d, rho=1.1)
        U[2] = (V0 * np.cos(lat))[:, None]
        A = o.prepare(U)
        F, _ = o.fcn(U, A)
        ana = 2 * 1.1 * V0 * np.sin(lat) / d.rad
        errs.append(np.abs(F[0][:, 1] - ana).max() / np.abs(ana).max())
        # theta is uniform, so rho theta changes with rho: F5 = theta F1
        assert np.abs(F[4] - 300.0 * F[0]).max() < 1e-10 * np.abs(300 * ana).max()
    assert errs[-1] < 1e-3 and _converge(errs, 4.0), errs


@pytest.mark.parametrize("l", [2, 3])
def test_hyperviscosity_on_spherical_harmonics(pkg, oracle_mod, l):
    """Pass 1 gives the Laplacian of theta (-l(l+1)/R^2 Y); pass 2 damps it as
    -K (l(l+1)/R^2)^2 Y.  The pointwise bi-Laplacian carries the grid-scale
    error of the first Laplacian amplified by 1/h^2, so it is checked by its
    Galerkin projection onto Y in the DSS mass inner product and by an L2
    error that falls with resolution."""
    K = 1e15
    lap_err, proj_err, l2_err = [], [], []
    for ne in (4, 8, 16):
        d = disc_for(pkg, ne, 3)
        o = oracle_mod.Oracle(d, phys_with(pkg, Grav=0.0, Omega=0.0), (K, K, K))
        lon, lat = node_lonlat(d)
        Y = 0.5 * (3 * np.sin(lat) ** 2 - 1) if l == 2 else np.cos(lat) ** 3 * np.cos(3 * lon)
        U = uniform_state(d)
        U[4] = (300.0 + Y)[:, None]
        A = o.prepare(U)
        F, T1 = o.fcn(U, A)
        lap = -(l * (l + 1)) / d.rad ** 2
        lap_err.append(np.abs(T1[4][:, 1] - lap * Y).max() / np.abs(lap * Y).max())
        target = -K * lap * lap * Y
        Mw = d.M[0][:, 1] + d.M[1][:, 1]
        proj_err.append(abs((Mw * F[4][:, 1] * Y).sum() / (Mw * Y * Y).sum() / (-K * lap * lap) - 1))
        l2_err.append(np.linalg.norm(F[4][:, 1] - target) / np.linalg.norm(target))
    assert lap_err[-1] < 3e-4 and _converge(lap_err, 4.0), lap_err
    assert proj_err[-1] < 3e-3 and _converge(proj_err, 2.0), proj_err
    assert l2_err[-1] < 0.1 and _converge(l2_err, 2.0), l2_err


def test_vector_hyperviscosity_damps_rotational_flow(pkg, oracle_mod):
    """Solid-body rotation is an l = 1 eigenfunction of the Hodge vector
    Laplacian (grad div - curl curl, eigenvalue -2/R^2) with zero divergence:
    the curl hyperviscosity returns -K (2/R^2)^2 (u, v) and the gradient part
    nothing.  Checked by Galerkin projection, which converges to 1 and 0."""
    K = 1e15
    curl_err, grad_err = [], []
    for ne in (4, 8, 16):
        d = disc_for(pkg, ne, 3)
        lon, lat = node_lonlat(d)
        U = uniform_state(d)
        U[1] = (20.0 * np.cos(lat))[:, None]
        Fs = {}
        for name, ko in (("curl", (K, 0.0, 0.0)), ("none", (0.0, 0.0, 0.0)), ("grad", (0.0, K, 0.0))):
            o = oracle_mod.Oracle(d, phys_with(pkg, Grav=0.0, Omega=0.0), ko)
            A = o.prepare(U)
            Fs[name] = o.fcn(U, A)[0]
        target = -K * (2.0 / d.rad ** 2) ** 2 * 20.0 * np.cos(lat)
        Mw = d.M[0][:, 1] + d.M[1][:, 1]
        nrm = (Mw * target * target).sum()
        curl_err.append(abs((Mw * (Fs["curl"][1][:, 1] - Fs["none"][1][:, 1]) * target).sum() / nrm - 1))
        grad_err.append(abs((Mw * (Fs["grad"][1][:, 1] - Fs["none"][1][:, 1]) * target).sum() / nrm))
    assert curl_err[-1] < 0.06 and _converge(curl_err, 3.0), curl_err
    assert grad_err[-1] < 1e-3 and _converge(grad_err, 10.0), grad_err


# ---- tracers (OperatorKernel.jl:722-759, ViscKernel.jl:169-218,277-327) ----

def _tracer_setup(pkg, oracle_mod, ntr, ne=3, nz=8, seed=11):
    disc = disc_for(pkg, ne, nz)
    phys = pkg.discretization.Phys()
    hd = pkg.host.hyperdiffusion_for(ne) if hasattr(pkg, "host") else 5e14
    U0 = pkg.state.jet_state(disc, seed=seed)
    rng = np.random.default_rng(seed)
    U0[1] += 3.0 * rng.standard_normal(U0[1].shape)
    U0[2] += 3.0 * rng.standard_normal(U0[2].shape)
    U0[3] += 0.3 * rng.standard_normal(U0[3].shape)
    U0[3][:, -1] = 0.0
    orc = oracle_mod.Oracle(disc, phys, (hd, hd, hd), NumTr=ntr)
    orc0 = oracle_mod.Oracle(disc, phys, (hd, hd, hd))
    return disc, U0, orc, orc0


def test_tracer_equal_to_rhotheta_reproduces_theta(pkg, oracle_mod):
    """A tracer holding rho*theta has mixing ratio theta; its hyperviscosity
    cache and tendency are then the rho*theta ones bit for bit, and the
    dynamics are those of the run without tracers."""
    disc, U0, orc, orc0 = _tracer_setup(pkg, oracle_mod, 1)
    U = np.concatenate([U0, U0[4:5]], axis=0)
    A = orc.prepare(U.copy())
    F, T1 = orc.fcn(U.copy(), A)
    assert np.array_equal(F[5], F[4]) and np.array_equal(T1[5], T1[4])
    F0, T10 = orc0.fcn(U0.copy(), orc0.prepare(U0.copy()))
    assert np.array_equal(F[:5], F0) and np.array_equal(T1[:5], T10)


def test_constant_mixing_ratio(pkg, oracle_mod):
    """Tracer = c0 * rho: its hyperviscosity is that of a constant (zero up
    to the rounding of D applied to a constant), and its transport is c0
    times the mass tendency."""
    disc, U0, orc, _ = _tracer_setup(pkg, oracle_mod, 1)
    c0 = 0.37
    U = np.concatenate([U0, c0 * U0[0:1]], axis=0)
    F, T1 = orc.fcn(U.copy(), orc.prepare(U.copy()))
    assert np.abs(T1[5]).max() <= 1e-13 * np.abs(T1[4]).max()
    err = np.linalg.norm(F[5] - c0 * F[0]) / np.linalg.norm(c0 * F[0])
    assert err <= 1e-10, err


def test_tracer_tendency_scales_linearly(pkg, oracle_mod):
    """The tracer tendency is linear in the tracer at fixed density and
    winds (the upwind limiter depends on ratios of mixing-ratio
    differences only), and each tracer is independent of the others."""
    disc, U0, orc, _ = _tracer_setup(pkg, oracle_mod, 2)
    rng = np.random.default_rng(3)
    q = U0[0] * (1.0 + 0.1 * rng.standard_normal(U0[0].shape))
    U = np.concatenate([U0, q[None], 4.0 * q[None]], axis=0)
    F, T1 = orc.fcn(U.copy(), orc.prepare(U.copy()))
    assert np.allclose(F[6], 4.0 * F[5], rtol=1e-12, atol=1e-12 * np.abs(F[6]).max())
    assert np.allclose(T1[6], 4.0 * T1[5], rtol=1e-12, atol=1e-12 * np.abs(T1[6]).max())


def test_tracers_are_passive_in_the_step(pkg, oracle_mod):
    """Tracers do not enter the Schur complement (k = fac v for them,
    SchurSolve.jl:106-109) nor any dynamical tendency: two Rosenbrock-W
    steps with two tracers leave rho, u, v, w, rho*theta bit-identical to
    the run without tracers."""
    disc, U0, orc, orc0 = _tracer_setup(pkg, oracle_mod, 2)
    rng = np.random.default_rng(5)
    q = U0[0] * (1.0 + 0.1 * rng.standard_normal(U0[0].shape))
    V = np.concatenate([U0, q[None], U0[4:5]], axis=0)
    V0 = U0.copy()
    orc.rw_step(V, 300.0, 2)
    orc0.rw_step(V0, 300.0, 2)
    assert np.array_equal(V[:5], V0)
    assert np.isfinite(V).all() and not np.array_equal(V[5], q)


# ---- physics hooks (FcnGPU.jl:953-966) ----

def _hooks_case(pkg, oracle_mod, damping=None, forcing=None):
    disc = disc_for(pkg, 3, 12)
    phys = pkg.discretization.Phys()
    U0 = pkg.state.jet_state(disc, seed=9)
    rng = np.random.default_rng(9)
    U0[3] += 0.3 * rng.standard_normal(U0[3].shape)
    U0[3][:, -1] = 0.0
    hd = 5e14 * (30 / 3) ** 3.2
    base = oracle_mod.Oracle(disc, phys, (hd, hd, hd))
    hooked = oracle_mod.Oracle(disc, phys, (hd, hd, hd), damping=damping, forcing=forcing)
    A = base.prepare(U0.copy())
    F0, _ = base.fcn(U0.copy(), A)
    F1, _ = hooked.fcn(U0.copy(), A)
    return disc, U0, A, F1 - F0, F0


def _close(dF, want, F0, rtol=1e-13):
    """dF = (F0 + term) - F0 carries the rounding of the sum: 2 ulp of |F0|."""
    return np.all(np.abs(dF - want) <= rtol * np.abs(want) + 4.5e-16 * np.abs(F0) + 1e-300)


def test_damping_sponge_profile(pkg, oracle_mod):
    """DampingW: the w tendency gains -Relax sin^2(pi/2 (1 - (H - z)/S)) w in
    the top S metres of the layer centres zP and nothing below; the other
    components are untouched."""
    H, S, relax = 30000.0, 10000.0, 1.0 / 3600.0
    disc, U0, A, dF, F0 = _hooks_case(pkg, oracle_mod, damping=(H, S, relax))
    zP = oracle_mod.layer_center_height(disc)  # (NumG, Nz)
    assert np.all(dF[[0, 1, 2, 4]] == 0.0)
    below = zP < H - S
    assert np.all(dF[3][below] == 0.0)
    coef = relax * np.sin(0.5 * np.pi * (1.0 - (H - zP) / S)) ** 2
    want = np.where(below, 0.0, -coef * U0[3])
    assert _close(dF[3], want, F0[3])
    top = zP[:, -1]
    assert np.all(coef[:, -1] <= relax) and np.all(coef[:, -1] > 0.9 * relax)  # near full strength at the lid


def test_held_suarez_forcing(pkg, oracle_mod):
    """HeldSuarezForcing with the reference's branch on sigma < 0.7: winds are
    damped by k_f max(0, (sigma - sigma_b)/(1 - sigma_b)) (zero above
    sigma_b), and rho*theta relaxes towards T_eq(lat, sigma) where
    sigma < 0.7: dF5 = -rho k_T (T - T_eq) / sigma^kappa with T = p/(rho Rd)."""
    hs = pkg.host.held_suarez_param()
    disc, U0, A, dF, F0 = _hooks_case(pkg, oracle_mod, forcing=hs)
    phys = pkg.discretization.Phys()
    lat = oracle_mod.node_lonlat(disc)[:, 1][:, None]
    p = A[0]
    sigma = p / phys.p0
    hf = np.maximum(0.0, (sigma - hs["sigma_b"]) / (1.0 - hs["sigma_b"]))
    assert _close(dF[1], -hs["k_f"] * hf * U0[1], F0[1])
    assert _close(dF[2], -hs["k_f"] * hf * U0[2], F0[2])
    assert np.all(dF[1][sigma < hs["sigma_b"]] == 0.0) and np.any(dF[1] != 0.0)
    kT = np.where(sigma < 0.7, hs["k_a"] + (hs["k_s"] - hs["k_a"]) * hf * np.cos(lat) ** 4, 0.0)
    Teq = np.maximum(hs["T_min"], (hs["T_equator"] - hs["DeltaT_y"] * np.sin(lat) ** 2
                                   - hs["DeltaTh_z"] * np.log(sigma) * np.cos(lat) ** 2) * sigma ** phys.kappa)
    T = p / (U0[0] * phys.Rd)
    want = -U0[0] * kT * (T - Teq) / sigma ** phys.kappa
    assert _close(dF[4], want, F0[4], rtol=1e-11)
    # relaxation towards T_eq: the theta tendency opposes T - T_eq
    m = (kT > 0) & (np.abs(T - Teq) > 1e-6)
    assert np.all(np.sign(dF[4][m]) == -np.sign((T - Teq)[m]))
    assert np.all(dF[[0, 3]] == 0.0)


# ---- moist equation of state (Equation.jl:56-75) ----

def _moist_state(pkg, disc, seed=13):
    U0 = pkg.state.jet_state(disc, seed=seed)
    rng = np.random.default_rng(seed)
    qv = 0.012 * (1.0 + 0.3 * rng.random(U0[0].shape))
    qc = 2e-4 * rng.random(U0[0].shape)
    return np.concatenate([U0, (qv * U0[0])[None], (qc * U0[0])[None]], axis=0)


def test_moist_eos(pkg, oracle_mod):
    """p of Models.Moist inverts Rd rho*theta = p^(1-kappa_m) p0^kappa_m with
    kappa_m = R_m / c_pml of the dry air, vapour and cloud densities;
    T = p / R_m; theta = (p0/p)^kappa_m T.  Without vapour and cloud it is the
    dry equation of state."""
    disc = disc_for(pkg, 3, 8)
    P = pkg.discretization.Phys()
    U = _moist_state(pkg, disc)
    orc = oracle_mod.Oracle(disc, P, NumTr=2, moist=(6, 7, P.Cpv, P.Cpl, P.Rv))
    A = orc.prepare(U.copy())
    rho, rth, rv, rc = U[0], U[4], U[5], U[6]
    rd = rho - rv - rc
    Rm = P.Rd * rd + P.Rv * rv
    km = Rm / (P.Cpd * rd + P.Cpv * rv + P.Cpl * rc)
    p = A[0]
    assert np.allclose(p ** (1.0 - km) * P.p0 ** km, P.Rd * rth, rtol=1e-13, atol=0)
    assert np.allclose(A[1], p / Rm, rtol=1e-14, atol=0)
    assert np.allclose(A[2], (P.p0 / p) ** km * A[1], rtol=1e-13, atol=0)
    Ud = U.copy()
    Ud[5:] = 0.0
    Am = orc.prepare(Ud.copy())
    Ad = oracle_mod.Oracle(disc, P).prepare(Ud[:5].copy())
    assert np.allclose(Am, Ad, rtol=1e-13, atol=0)
    # vapour lowers kappa_m below kappa: the moist pressure differs
    assert np.abs(A[0] / oracle_mod.Oracle(disc, P).prepare(U[:5].copy())[0] - 1.0).max() > 1e-4


# ---- HorLimit: LimitKernel! + DivRhoTrUpwind3LimKernel! (HorLimiterKernel.jl:2-196)
def _lim_state(pkg, ne=4, nz=8):
    D = pkg.discretization
    disc = D.build(ne, 4, nz, 30000.0)
    U = pkg.state.jet_state(disc, seed=0)
    rng = np.random.default_rng(1)
    U[1] += 5 * rng.standard_normal(U[1].shape)
    U[3] += 0.2 * rng.standard_normal(U[3].shape)
    U[3][:, -1] = 0.0
    return disc, D.Phys(), U


def test_horlimit_inactive_and_constant(pkg, oracle_mod):
    """A horizontally uniform mixing ratio q(z) leaves nothing to limit: the
    limited tendency equals the unlimited transport (whose tracer
    hyperviscosity vanishes for it); a constant mixing ratio c0 gives
    c0 * F_rho; the dynamics are untouched."""
    disc, phys, U = _lim_state(pkg)
    K = pkg.host.hyperdiffusion_for(4)
    ol = oracle_mod.Oracle(disc, phys, (K, K, K), NumTr=1, horlimit=True)
    ol.dtau = 150.0
    ou = oracle_mod.Oracle(disc, phys, (K, K, K), NumTr=1)
    qz = 1e-3 * (1 + np.linspace(0, 1, disc.nz))[None, :]
    U2 = np.concatenate([U, (qz * U[0])[None]], 0)
    Fl, _ = ol.fcn(U2.copy(), ol.prepare(U2.copy()))
    Fu, _ = ou.fcn(U2.copy(), ou.prepare(U2.copy()))
    assert np.linalg.norm(Fl[5] - Fu[5]) <= 1e-12 * np.linalg.norm(Fu[5])
    assert np.array_equal(Fl[:5], Fu[:5])
    U3 = np.concatenate([U, (1e-3 * U[0])[None]], 0)
    Fl, _ = ol.fcn(U3.copy(), ol.prepare(U3.copy()))
    assert np.linalg.norm(Fl[5] - 1e-3 * Fl[0]) <= 1e-9 * np.linalg.norm(1e-3 * Fl[0])


def test_horlimit_bounds_a_step(pkg, oracle_mod):
    """A step in the mixing ratio: an explicit update with the limited
    tendency over dtau keeps q inside the initial range (to rounding) at
    time steps where the unlimited transport overshoots."""
    disc, phys, U = _lim_state(pkg)
    K = pkg.host.hyperdiffusion_for(4)
    ol = oracle_mod.Oracle(disc, phys, (K, K, K), NumTr=1, horlimit=True)
    ou = oracle_mod.Oracle(disc, phys, (K, K, K), NumTr=1)
    ll = oracle_mod.node_lonlat(disc)
    q = np.where(np.abs(ll[:, 0]) < 0.8, 2e-3, 1e-3)[:, None] * np.ones((1, disc.nz))
    U3 = np.concatenate([U, (q * U[0])[None]], 0)
    over = []
    for dt in (150.0, 600.0, 1200.0):
        ol.dtau = dt
        Fl, _ = ol.fcn(U3.copy(), ol.prepare(U3.copy()))
        Fu, _ = ou.fcn(U3.copy(), ou.prepare(U3.copy()))
        ql = (U3[5] + dt * Fl[5]) / (U3[0] + dt * Fl[0])
        qu = (U3[5] + dt * Fu[5]) / (U3[0] + dt * Fu[0])
        assert ql.min() >= 1e-3 * (1 - 1e-12) and ql.max() <= 2e-3 * (1 + 1e-12), dt
        over.append(qu.min() < 1e-3 * (1 - 1e-6) or qu.max() > 2e-3 * (1 + 1e-6))
    assert all(over)


def test_oracle_threads_do_not_change_results(pkg, oracle_mod):
    """The DSS adds run in a fixed (colour, element) order, so the tendency
    (with tracers and HorLimit) and the steps are bit-identical at 1 and at
    4 threads."""
    disc, phys, U = _lim_state(pkg, ne=3, nz=6)
    U2 = np.concatenate([U, (1e-3 * (1 + 0.3 * np.sign(U[1])) * U[0])[None]], 0)
    K = pkg.host.hyperdiffusion_for(3)
    prev = oracle_mod.set_threads(0)
    try:
        out = []
        for nt in (1, 4):
            assert oracle_mod.set_threads(nt) in (1, nt)
            orc = oracle_mod.Oracle(disc, phys, (K, K, K), NumTr=1, horlimit=True)
            orc.dtau = 300.0
            F, T1 = orc.fcn(U2.copy(), orc.prepare(U2.copy()))
            V = U2.copy()
            orc.rw_step(V, 300.0, 2)
            out.append((F, T1, V))
        for a, b in zip(*out):
            assert np.array_equal(a, b)
    finally:
        oracle_mod.set_threads(prev)
