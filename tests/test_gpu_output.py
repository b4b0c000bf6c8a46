"""Output interpolation on the device and restarts (GPU, C ABI).

The interpolation kernels are checked against the same sums in numpy
(vtkOutputKernel.jl:49-158); a run resumed from a restart file equals the
uninterrupted run bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def np_interp(inter, glob, c):
    """cCell[IF, Iz, J, I] = sum_{jP, iP} Inter[I,J,iP,jP] c[Iz, Glob[iP + N jP, IF]]."""
    nf, nn = glob.shape
    n = inter.shape[0]
    g = glob.reshape(nf, n, n) - 1  # [IF, jP, iP]
    vals = c[g]  # (NF, jP, iP, Nz)
    return np.einsum("jipq,fjiz->fzpq", inter, vals)


def test_interpolation_kernels(pkg, torch_cuda):
    torch = torch_cuda
    host, out, D = pkg.host, pkg.output, pkg.discretization
    disc = D.build(3, 4, 6, 30000.0)
    phys = D.Phys()
    FE, Metric, Global, Cache, JCache = host.device_setup(disc, phys, device="cuda")
    U0 = pkg.state.jet_state(disc, seed=2)
    rng = np.random.default_rng(2)
    U0[3] += 0.5 * rng.standard_normal(U0[3].shape)
    U = torch.as_tensor(U0, device="cuda")
    P = 3
    inter = out.vtk_inter(disc.xw, P)
    Inter = torch.as_tensor(inter, device="cuda")
    cCell = torch.empty((disc.nf, disc.nz, P, P), dtype=torch.float64, device="cuda")
    out.InterpolateGPU(cCell, U[1], Inter, FE.Glob)
    torch.cuda.synchronize()
    ref = np_interp(inter, disc.glob, U0[1])
    assert np.allclose(cCell.cpu().numpy(), ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    out.InterpolateRhoGPU(cCell, U[4], U[0], Inter, FE.Glob)
    torch.cuda.synchronize()
    nf, n = disc.nf, disc.N
    g = disc.glob.reshape(nf, n, n) - 1
    ref = np.einsum("jipq,fjiz->fzpq", inter, U0[4][g] / U0[0][g])
    assert np.allclose(cCell.cpu().numpy(), ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    out.InterpolateWBGPU(cCell, U[1], U[2], U[3], Inter, Metric.dXdxI, FE.Glob)
    torch.cuda.synchronize()
    w = U0[3][g]  # (NF, jP, iP, Nz)
    dX = disc.dXdxI  # (NF, Nz, NN, 2, 3, 3) with [..., j, i] = dXdxI[i, j]
    d31 = dX[:, 0, :, 0, 0, 2].reshape(nf, n, n)
    d32 = dX[:, 0, :, 0, 1, 2].reshape(nf, n, n)
    d33 = dX[:, 0, :, 0, 2, 2].reshape(nf, n, n)
    w0 = -(U0[1][g][..., 0] * d31 + U0[2][g][..., 0] * d32) / d33
    wb = np.concatenate([w0[..., None], w[..., :-1]], axis=-1)
    ref = np.einsum("jipq,fjiz->fzpq", inter, 0.5 * (w + wb))
    assert np.allclose(cCell.cpu().numpy(), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_vtk_file_from_device(pkg, torch_cuda, tmp_path):
    torch = torch_cuda
    host, out, D = pkg.host, pkg.output, pkg.discretization
    disc = D.build(2, 4, 4, 30000.0)
    phys = D.Phys()
    FE, Metric, Global, Cache, JCache = host.device_setup(disc, phys, device="cuda", NumTr=1)
    U0 = pkg.state.jet_state(disc, seed=0)
    U0 = np.concatenate([U0, 0.01 * U0[0:1]], axis=0)
    U = torch.as_tensor(U0, device="cuda")
    path = str(tmp_path / "out.vtk")
    out.unstructured_vtkSphere(path, U, FE, Metric, Global, disc.xw, disc.rad, ord_print=2)
    h = out.read_vtk_header(path)
    assert h["cells"] == disc.nf * disc.nz * 4 and h["points"] == 8 * h["cells"] and h["connectivity_ok"]
    assert h["fields"] == ["rho", "u", "v", "w", "Th", "Tr1"]
    assert np.allclose(h["Tr1"], 0.01, rtol=1e-12)  # constant mixing ratio
    ref = np_interp(out.vtk_inter(disc.xw, 2), disc.glob, U0[0])
    assert np.allclose(h["rho"], ref.reshape(-1), rtol=1e-14)


@pytest.mark.parametrize("ntr", [0, 2])
def test_restart_resumes_bit_for_bit(pkg, torch_cuda, tmp_path, ntr):
    torch = torch_cuda
    host, out, D = pkg.host, pkg.output, pkg.discretization
    disc = D.build(3, 4, 10, 30000.0)
    phys = D.Phys()
    FE, Metric, Global, Cache, JCache = host.device_setup(disc, phys, device="cuda", NumTr=ntr)
    U0 = pkg.state.jet_state(disc, seed=4)
    if ntr:
        U0 = np.concatenate([U0] + [0.01 * (k + 1) * U0[0:1] for k in range(ntr)], axis=0)
    A = torch.as_tensor(U0, device="cuda").clone()
    host.TimeStepper(A, 300.0, 4, FE, Metric, phys, Cache, JCache, Global)
    B = torch.as_tensor(U0, device="cuda").clone()
    host.TimeStepper(B, 300.0, 2, FE, Metric, phys, Cache, JCache, Global)
    path = str(tmp_path / "restart.bin")
    out.write_restart(path, B, Global, step=2, time=600.0)
    del B
    C, head = out.read_restart(path, Global, device="cuda")
    assert head["step"] == 2 and head["NumTr"] == ntr
    host.TimeStepper(C, 300.0, 2, FE, Metric, phys, Cache, JCache, Global)
    torch.cuda.synchronize()
    assert torch.equal(A, C)
