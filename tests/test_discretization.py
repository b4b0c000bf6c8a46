"""Static inputs: GLL basis, derivative matrices, sphere metric, DSS mass."""
import math

import numpy as np
import pytest


def test_gll_closed_form(pkg):
    x, w = pkg.discretization.gausslobatto(5)
    s = math.sqrt(3.0 / 7.0)
    assert np.allclose(x, [-1, -s, 0, s, 1], atol=1e-15, rtol=0)
    assert np.allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], atol=1e-15, rtol=0)
    x4, w4 = pkg.discretization.gausslobatto(4)
    s5 = math.sqrt(1.0 / 5.0)
    assert np.allclose(x4, [-1, -s5, s5, 1], atol=1e-15) and np.allclose(w4, [1 / 6, 5 / 6, 5 / 6, 1 / 6])


@pytest.mark.parametrize("p", [3, 4])
def test_derivative_matrices(pkg, p):
    DW, DS, x, w = pkg.discretization.derivative_matrices(p)
    n = p + 1
    # exact on polynomials of degree <= p
    for k in range(n):
        f = x ** k
        df = k * x ** (k - 1) if k else np.zeros(n)
        assert np.allclose(DS @ f, df, atol=1e-13)
    # summation by parts: W DS + DS' W = diag(-1, 0, ..., 0, 1)
    W = np.diag(w)
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1, 1
    assert np.abs(W @ DS + DS.T @ W - B).max() < 1e-13
    # weak derivative DW = -W^-1 DS' W
    assert np.abs(DW - (-np.linalg.inv(W) @ DS.T @ W)).max() < 1e-13


@pytest.fixture(scope="module")
def disc6(pkg):
    return pkg.discretization.build(6, 4, 5, 30000.0)


def test_metric_structure(disc6):
    d = disc6
    assert (d.J > 0).all()
    # dXdxI is the adjugate of the local Jacobian: det(adj L) = det(L)^2
    A = np.swapaxes(d.dXdxI, -1, -2)  # [..., i, j] = dXdxI[i, j]
    det = np.linalg.det(A)  # (NF, Nz, N^2, 2)
    J = np.transpose(d.J, (0, 1, 3, 2))
    assert np.abs(det / J ** 2 - 1).max() < 1e-10
    # structural zeros of the shallow metric (JacobiSphere3GPU.jl:174-175)
    assert np.all(d.dXdxI[..., 2, 0] == 0) and np.all(d.dXdxI[..., 2, 1] == 0)
    # flat topography: the terrain rows vanish to roundoff
    assert np.abs(d.dXdxI[..., 0, 2]).max() < 1e-12 * np.abs(d.dXdxI[..., 2, 2]).max()


@pytest.mark.parametrize("ne,tol", [(4, 1e-7), (8, 1e-9)])
def test_shell_volume(pkg, ne, tol):
    """sum w_i w_j (J1 + J2) over elements and layers -> 4 pi R^2 H."""
    d = pkg.discretization.build(ne, 4, 4, 30000.0)
    wq = np.outer(d.w, d.w).reshape(-1)
    vol = (d.J * wq).sum()
    assert abs(vol / (4 * math.pi * d.rad ** 2 * d.H) - 1) < tol


def test_mass_is_sum_of_jacobians(disc6):
    d = disc6
    M = np.zeros_like(d.M)
    nf, nz, _, nn = d.J.shape
    for e in range(nf):
        for i in range(nn):
            g = d.glob[e, i] - 1
            M[0, g] += d.J[e, :, 0, i]
            M[1, g] += d.J[e, :, 1, i]
    assert np.abs(M - d.M).max() <= 1e-12 * np.abs(d.M).max()


def test_layer_thickness(disc6):
    assert np.allclose(disc6.dz, disc6.H / disc6.nz, rtol=1e-10)


def test_seeded_state(pkg, disc6):
    st = pkg.state
    n1 = st.noise(0, disc6.num_g, disc6.nz)
    n2 = st.noise(0, disc6.num_g, disc6.nz)
    n3 = st.noise(1, disc6.num_g, disc6.nz)
    assert np.array_equal(n1, n2) and not np.array_equal(n1, n3)
    assert n1.min() >= -1 and n1.max() < 1 and abs(n1.mean()) < 0.02
    U = st.jet_state(disc6, seed=0)
    phys = pkg.discretization.Phys()
    # theta recovered from the state deviates from the analytic profile by the noise only
    zm = (np.arange(disc6.nz) + 0.5) * disc6.H / disc6.nz
    T = 288 - 0.005 * zm
    p = 1e5 * (T / 288) ** (phys.Grav / (phys.Rd * 0.005))
    theta = T * (phys.p0 / p) ** phys.kappa
    assert np.abs(U[4] / U[0] - theta - n1).max() < 1e-10
    assert np.all(U[2] == 0) and np.all(U[3] == 0)
    assert 0 <= U[1].min() and U[1].max() <= 35.0
