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


def mountain(X):
    """A smooth surface height (m) from the node directions of a metric's X."""
    x, y, z = X[:, 0, 0, 0, :], X[:, 0, 1, 0, :], X[:, 0, 2, 0, :]
    r = np.sqrt(x * x + y * y + z * z)
    return 800.0 * (1.0 + x / r) + 200.0 * (1.0 + 2.0 * y * z / (r * r))  # (NF, N^2), >= 0


@pytest.fixture(scope="module")
def topo(pkg):
    D, G = pkg.discretization, pkg.grid
    ne, p, nz, H, rad = 3, 4, 6, 30000.0, 6.37122e6
    g = G.cubed_sphere(ne, rad)
    DW, DS, xw, w = D.derivative_matrices(p)
    z = D.vertical_grid(nz, H)
    flat = D.sphere_metric(g.face_corners(), xw, DS, z, rad, H)
    n = p + 1
    zs = mountain(flat[0]).reshape(-1, n, n)  # (NF, J, I)
    mt = D.sphere_metric(g.face_corners(), xw, DS, z, rad, H, zs=zs)
    return dict(D=D, g=g, DS=DS, xw=xw, z=z, H=H, rad=rad, n=n, zs=zs, flat=flat, topo=mt)


def test_topography_flat_limit(topo):
    """zs = 0 gives the flat metric bit for bit."""
    T = topo
    zero = np.zeros_like(T["zs"])
    m0 = T["D"].sphere_metric(T["g"].face_corners(), T["xw"], T["DS"], T["z"], T["rad"], T["H"], zs=zero)
    for a, b in zip(m0, T["flat"]):
        assert np.array_equal(a, b)


def test_topography_column_volume(topo):
    """J = DzhR * Jhor, and DzhR telescopes over the layers: the sum of J
    over both vertical nodes and all layers is Jhor (H - zs) at every node
    (Jhor = the same sum of the flat metric over H)."""
    T = topo
    Jt = T["topo"][2].sum(axis=(1, 2))  # (NF, N^2)
    Jhor = T["flat"][2].sum(axis=(1, 2)) / T["H"]
    zs = T["zs"].reshape(Jt.shape)
    assert np.all(T["topo"][2] > 0)
    assert np.allclose(Jt, Jhor * (T["H"] - zs), rtol=1e-12, atol=0)
    assert not np.allclose(Jt, Jhor * T["H"], rtol=1e-6)


def test_topography_metric_inverts_to_height_derivatives(topo):
    """dXdxI is the adjugate of dXdxLoc, so dXdxLoc = J inv(dXdxI).  Its
    horizontal block is the flat one, its third row holds the derivatives
    of the node heights h = |X| - Rad along the reference coordinates
    (D applied across the element row / column) and half the layer height."""
    T = topo
    X, dXdxI, J = T["topo"]
    n, DS, rad = T["n"], T["DS"], T["rad"]
    A = np.swapaxes(dXdxI, -1, -2)  # [..., i, j] = dXdxI[i, j]
    Lloc = J[..., None, None].transpose(0, 1, 3, 2, 4, 5) * np.linalg.inv(A)  # (NF, Nz, N^2, 2, 3, 3)
    Af = np.swapaxes(T["flat"][1], -1, -2)
    Lflat = T["flat"][2][..., None, None].transpose(0, 1, 3, 2, 4, 5) * np.linalg.inv(Af)
    assert np.allclose(Lloc[..., :2, :2], Lflat[..., :2, :2], rtol=0, atol=1e-9 * np.abs(Lflat[..., :2, :2]).max())
    assert np.all(np.abs(Lloc[..., :2, 2]) <= 1e-12 * np.abs(Lloc).max())
    h = np.sqrt((X * X).sum(axis=2)) - rad  # (NF, Nz, 2, N^2)
    hg = h.reshape(h.shape[:3] + (n, n))  # [..., J, I]
    dx = np.einsum("ik,fzKjk->fzKji", DS, hg).reshape(h.shape)
    dy = np.einsum("jk,fzKki->fzKji", DS, hg).reshape(h.shape)
    dzh = 0.5 * (h[:, :, 1] - h[:, :, 0])
    L31 = np.transpose(Lloc[..., 2, 0], (0, 1, 3, 2))
    L32 = np.transpose(Lloc[..., 2, 1], (0, 1, 3, 2))
    L33 = np.transpose(Lloc[..., 2, 2], (0, 1, 3, 2))
    scale = np.abs(dx).max()
    assert scale > 1.0  # the terrain row is not roundoff
    assert np.allclose(L31, dx, rtol=0, atol=1e-6 * scale)
    assert np.allclose(L32, dy, rtol=0, atol=1e-6 * scale)
    assert np.allclose(L33, dzh[:, :, None, :], rtol=1e-8, atol=0)


def test_mmass_and_layer_center_brute_force(pkg):
    """MMass (MassCG.jl:73) and zP (DiscretizationCG.jl:130-134) by plain
    loops over (IF, ID) in the reference's order."""
    D = pkg.discretization
    d = D.build(2, 4, 3, 30000.0)
    n = d.N
    MM = D.mass_mmass(d.J, d.glob, d.num_g, d.w)
    zP = D.layer_center_height(d.X, d.glob, d.num_g, d.rad)
    MMb = np.zeros((d.num_g, d.nz))
    zPb = np.zeros((d.num_g, d.nz))
    for IF in range(d.nf):
        for j in range(n):
            for i in range(n):
                ID = i + n * j
                g = d.glob[IF, ID] - 1
                for iz in range(d.nz):
                    MMb[g, iz] += 0.5 * (d.J[IF, iz, 0, ID] + d.J[IF, iz, 1, ID]) * d.w[i] * d.w[j]
                    xm = [0.5 * (d.X[IF, iz, c, 0, ID] + d.X[IF, iz, c, 1, ID]) for c in range(3)]
                    zPb[g, iz] = max(math.sqrt(xm[0] * xm[0] + xm[1] * xm[1] + xm[2] * xm[2]) - d.rad, 0.0)
    assert np.array_equal(MM, MMb)
    assert np.array_equal(zP, zPb)
    # 0.5 (J1 + J2) with the unit weights of the two vertical nodes: MMass
    # sums to half the shell volume 4 pi R^2 H (1.2e-5 relative at ne=2)
    vol = 4.0 * math.pi * d.rad ** 2 * d.H
    assert abs(2.0 * MM.sum() / vol - 1.0) < 2e-5
