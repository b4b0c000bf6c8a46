"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Bar: relative L2 error per prognostic variable <= 1e-12 for the FP64 path
(the tolerance of the plain FP64 drop-in), on the seeded jet state.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_l2(a, b):
    den = np.linalg.norm(np.ravel(b))
    return float(np.linalg.norm(np.ravel(a - b)) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _setup(pkg, oracle_mod, ne, nz, p=4, seed=0):
    D = pkg.discretization
    host = pkg.host
    disc = D.build(ne, p, nz, 30000.0)
    phys = D.Phys()
    U0 = pkg.state.jet_state(disc, seed=seed)
    FE, Metric, Global, Cache, JCache = host.device_setup(disc, phys, device="cuda")
    orc = oracle_mod.Oracle(disc, phys, (Global.Model.HyperDCurl, Global.Model.HyperDGrad, Global.Model.HyperDDiv))
    return disc, phys, U0, (FE, Metric, Global, Cache, JCache), orc


@pytest.mark.parametrize("ne,nz", [(3, 8), (4, 10), (6, 10), (4, 64), (5, 41), (3, 3)])
def test_fcn_prepare_parity(pkg, oracle_mod, torch_cuda, ne, nz):
    torch = torch_cuda
    host = pkg.host
    disc, phys, U0, (FE, Metric, Global, Cache, JCache), orc = _setup(pkg, oracle_mod, ne, nz)
    print(f"fcn parity ne={ne} nz={nz}", flush=True)
    U = torch.as_tensor(U0, device="cuda")
    F = torch.full_like(U, 7.0)  # FcnGPU! must zero-and-overwrite
    eq = host.CompressibleShallow()
    host.FcnPrepareGPU(U, FE, Metric, phys, Cache, None, Global, None, eq)
    host.FcnGPU(F, U, FE, Metric, phys, Cache, None, Global, None, eq)
    torch.cuda.synchronize()
    A = orc.prepare(U0.copy())
    Fo, T1o = orc.fcn(U0.copy(), A)
    Ag = Cache.AuxG.cpu().numpy()
    for c in range(3):
        assert rel_l2(Ag[c], A[c]) <= TOL, f"AuxG[{c}]"
    T1g = Cache.Temp1.cpu().numpy()
    for v in range(5):
        assert rel_l2(T1g[v], T1o[v]) <= TOL, f"Temp1[{v}] {rel_l2(T1g[v], T1o[v]):.3e}"
    Fg = F.cpu().numpy()
    for v in range(5):
        e = rel_l2(Fg[v], Fo[v])
        print(f"  F{v + 1} rel L2 {e:.3e}")
        assert e <= TOL, f"F{v + 1} rel L2 {e:.3e}"
    assert np.all(Fg[3][:, -1] == 0.0)  # the w lid receives no tendency


@pytest.mark.parametrize("ne,nz", [(3, 8), (4, 64)])
def test_jac_schur_parity(pkg, oracle_mod, torch_cuda, ne, nz):
    torch = torch_cuda
    host = pkg.host
    disc, phys, U0, (FE, Metric, Global, Cache, JCache), orc = _setup(pkg, oracle_mod, ne, nz)
    print(f"jac/schur parity ne={ne} nz={nz}", flush=True)
    U = torch.as_tensor(U0, device="cuda")
    eq = host.CompressibleShallow()
    host.FcnPrepareGPU(U, FE, Metric, phys, Cache, None, Global, None, eq)
    host.JacSchurGPU(JCache, U, FE, Metric, phys, Cache, Global, None, eq)
    A = orc.prepare(U0.copy())
    blocks = orc.jac(U0.copy(), A)
    for name, ob in zip(("JRhoW", "JWRho", "JWRhoTh", "JRhoThW"), blocks):
        assert rel_l2(getattr(JCache, name).cpu().numpy(), ob) <= TOL, name
    rng = np.random.default_rng(1)
    v0 = rng.standard_normal(U0.shape)
    v = torch.as_tensor(v0, device="cuda")
    k = torch.zeros_like(v)
    JCache.CompTri = True
    host.SchurSolveGPU(k, v, JCache, 150.0, Cache, Global)
    torch.cuda.synchronize()
    assert JCache.CompTri is False
    vo = v0.copy()
    ko, trio = orc.schur(vo, blocks, 150.0)
    assert rel_l2(JCache.tri.cpu().numpy(), trio) <= TOL
    kg, vg = k.cpu().numpy(), v.cpu().numpy()
    for c in range(5):
        assert rel_l2(kg[c], ko[c]) <= TOL, f"k[{c}] {rel_l2(kg[c], ko[c]):.3e}"
        assert rel_l2(vg[c], vo[c]) <= TOL, f"v[{c}]"
    assert np.all(kg[3][:, -1] == 0.0)


@pytest.mark.parametrize("ne,nz,dt,nsteps", [(3, 8, 300.0, 2), (6, 10, 300.0, 3)])
def test_rw_step_parity(pkg, oracle_mod, torch_cuda, ne, nz, dt, nsteps):
    torch = torch_cuda
    host = pkg.host
    disc, phys, U0, (FE, Metric, Global, Cache, JCache), orc = _setup(pkg, oracle_mod, ne, nz)
    print(f"rw parity ne={ne} nz={nz} steps={nsteps}", flush=True)
    V = torch.as_tensor(U0, device="cuda").clone()
    host.TimeStepper(V, dt, nsteps, FE, Metric, phys, Cache, JCache, Global)
    torch.cuda.synchronize()
    Vo = U0.copy()
    orc.rw_step(Vo, dt, nsteps)
    Vg = V.cpu().numpy()
    assert np.isfinite(Vg).all()
    for c in range(5):
        e = rel_l2(Vg[c], Vo[c])
        de = rel_l2(Vg[c] - U0[c], Vo[c] - U0[c])
        print(f"  V{c + 1} rel L2 {e:.3e} (increment {de:.3e})")
        assert e <= TOL, f"V{c + 1} rel L2 {e:.3e}"


def test_rw_composed_equals_fused(pkg, torch_cuda):
    """RosenbrockSchur with non-native callables (stage-by-stage composition in
    Python) equals the single cgd_rw_step call."""
    torch = torch_cuda
    host = pkg.host
    D = pkg.discretization
    disc = D.build(3, 4, 8, 30000.0)
    phys = D.Phys()
    U0 = pkg.state.jet_state(disc, seed=3)
    FE, Metric, Global, Cache, JCache = host.device_setup(disc, phys, device="cuda")
    eq = host.CompressibleShallow()
    V1 = torch.as_tensor(U0, device="cuda").clone()
    host.RosenbrockSchur(V1, 300.0, host.FcnGPU, host.FcnPrepareGPU, host.JacSchurGPU, FE, Metric, phys, Cache,
                         JCache, None, Global, None, eq)

    def fcn(*a):
        return host.FcnGPU(*a)

    V2 = torch.as_tensor(U0, device="cuda").clone()
    host.RosenbrockSchur(V2, 300.0, fcn, host.FcnPrepareGPU, host.JacSchurGPU, FE, Metric, phys, Cache, JCache,
                         None, Global, None, eq)
    torch.cuda.synchronize()
    for c in range(5):
        assert rel_l2(V2[c].cpu().numpy(), V1[c].cpu().numpy()) <= 1e-14
