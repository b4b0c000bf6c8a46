"""Output and restart on the host (no GPU): the sub-cell interpolation
matrix, the VTK points and file layout, and the restart file format."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def out(pkg):
    return pkg.output


def test_vtk_inter_is_exact_for_polynomials(pkg, out):
    """vtkInter (vtkSphere.jl:193-210) evaluates the element polynomial at
    the sub-cell centres: exact for x^a y^b, a, b <= p."""
    D = pkg.discretization
    _, _, xw, _ = D.derivative_matrices(4)
    P = 3
    inter = out.vtk_inter(xw, P)  # [j, i, jr, ir]
    assert np.allclose(inter.sum(axis=(0, 1)), 1.0, atol=1e-14, rtol=0)
    centres = -1.0 + (2.0 / P) * (np.arange(P) + 0.5)
    for a in range(5):
        for b in range(5):
            f = (xw[None, :] ** a) * (xw[:, None] ** b)  # f[j, i] = x_i^a y_j^b
            got = np.einsum("jipq,ji->pq", inter, f)
            want = (centres[None, :] ** a) * (centres[:, None] ** b)  # [jr, ir]
            assert np.allclose(got, want, atol=1e-13, rtol=0), (a, b)


def test_vtk_points_at_nodes_follow_the_layers(pkg, out):
    """With two sub-cells per direction every corner is a GLL node, where the
    Lagrange map of TransSphereX is exact: corners sit at radius R + z of
    their layer interface."""
    D = pkg.discretization
    d = D.build(2, 4, 3, 30000.0)
    pts = out.vtk_points(d.X, d.xw, 2, d.rad).reshape(d.nf, d.nz, 2, 2, 8, 3)
    r = np.linalg.norm(pts, axis=-1)
    for iz in range(d.nz):
        assert np.allclose(r[:, iz, :, :, :4], d.rad + d.z[iz], rtol=1e-13, atol=1e-6)
        assert np.allclose(r[:, iz, :, :, 4:], d.rad + d.z[iz + 1], rtol=1e-13, atol=1e-6)


def test_vtk_file_round_trip(out, tmp_path):
    rng = np.random.default_rng(0)
    pts = rng.standard_normal((16 * 8, 3))
    rho = rng.standard_normal(16)
    u = rng.standard_normal(16)
    path = str(tmp_path / "t.vtk")
    out.write_vtk(path, pts, {"rho": rho, "u": u})
    h = out.read_vtk_header(path)
    assert h["points"] == 128 and h["cells"] == 16 and h["connectivity_ok"]
    assert h["fields"] == ["rho", "u"]
    assert np.array_equal(h["rho"], rho) and np.array_equal(h["u"], u)


def _glob(NumG=7, Nz=5):
    from types import SimpleNamespace

    return SimpleNamespace(NumG=NumG, Nz=Nz, NF=3, N=5, Precision="fp64")


def test_restart_round_trip_is_bit_exact(pkg, out, tmp_path):
    G = _glob()
    rng = np.random.default_rng(1)
    V = rng.standard_normal((7, G.NumG, G.Nz))
    V[3, 0, 0] = -0.0
    V[2, 1, 1] = np.nextafter(0.0, 1.0)  # a subnormal survives
    path = str(tmp_path / "r.bin")
    out.write_restart(path, V, G, step=12, time=3600.0, extra={"seed": 0})
    W, head = out.read_restart(path, G)
    assert W.tobytes() == V.tobytes()
    assert head["NumTr"] == 2 and head["step"] == 12 and head["time"] == 3600.0 and head["extra"] == {"seed": 0}
    # the data are the array's bytes in the reference layout after the header
    raw = open(path, "rb").read()
    assert raw[:8] == b"CGDRST1\n" and raw.endswith(V.astype("<f8").tobytes())


def test_restart_rejects_bad_files(pkg, out, tmp_path):
    G = _glob()
    V = np.zeros((5, G.NumG, G.Nz))
    path = str(tmp_path / "r.bin")
    out.write_restart(path, V, G)
    with pytest.raises(pkg.host.CGDError, match="does not match"):
        out.read_restart(path, _glob(NumG=8))
    open(str(tmp_path / "bad.bin"), "wb").write(b"not a restart")
    with pytest.raises(pkg.host.CGDError, match="not a restart"):
        out.read_restart(str(tmp_path / "bad.bin"))
    data = open(path, "rb").read()
    open(str(tmp_path / "short.bin"), "wb").write(data[:-8])
    with pytest.raises(pkg.host.CGDError, match="truncated"):
        out.read_restart(str(tmp_path / "short.bin"))
    with pytest.raises(pkg.host.CGDError):
        out.write_restart(path, V.astype(np.float32), G)
