"""The moist equation of state (Models.Moist) on the GPU against the oracle:
FcnPrepareGPU's p, T, theta, and Rosenbrock-W steps (the fused stage kernels
recompute the moist pressure of every new state), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    den = np.linalg.norm(np.ravel(b))
    return float(np.linalg.norm(np.ravel(a - b)) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def moist_case(pkg, oracle_mod, ne, nz, seed=13):
    D, host = pkg.discretization, pkg.host
    disc = D.build(ne, 4, nz, 30000.0)
    P = D.Phys()
    FE, Metric, Global, Cache, JCache = host.device_setup(disc, P, device="cuda", NumTr=2, moist=(6, 7))
    hd = host.hyperdiffusion_for(ne)
    orc = oracle_mod.Oracle(disc, P, (hd, hd, hd), NumTr=2, moist=(6, 7, P.Cpv, P.Cpl, P.Rv))
    U0 = pkg.state.jet_state(disc, seed=seed)
    rng = np.random.default_rng(seed)
    qv = 0.012 * (1.0 + 0.3 * rng.random(U0[0].shape))
    qc = 2e-4 * rng.random(U0[0].shape)
    U0 = np.concatenate([U0, (qv * U0[0])[None], (qc * U0[0])[None]], axis=0)
    return disc, P, (FE, Metric, Global, Cache, JCache), orc, U0


def test_moist_prepare(pkg, oracle_mod, torch_cuda):
    torch = torch_cuda
    host = pkg.host
    disc, P, (FE, Metric, Global, Cache, JCache), orc, U0 = moist_case(pkg, oracle_mod, 3, 10)
    U = torch.as_tensor(U0, device="cuda")
    host.FcnPrepareGPU(U, FE, Metric, P, Cache, None, Global, None, host.CompressibleShallow())
    torch.cuda.synchronize()
    Ao = orc.prepare(U0.copy())
    Ag = Cache.AuxG.cpu().numpy()
    for c in range(3):
        assert rel_l2(Ag[c], Ao[c]) <= 1e-14, f"AuxG[{c + 1}]"


@pytest.mark.parametrize("ne,nz,nsteps", [(3, 10, 2), (4, 64, 2), (2, 70, 1)])
def test_moist_steps(pkg, oracle_mod, torch_cuda, ne, nz, nsteps):
    torch = torch_cuda
    host = pkg.host
    disc, P, (FE, Metric, Global, Cache, JCache), orc, U0 = moist_case(pkg, oracle_mod, ne, nz)
    print(f"moist steps ne={ne} nz={nz}", flush=True)
    V = torch.as_tensor(U0, device="cuda").clone()
    host.TimeStepper(V, 300.0, nsteps, FE, Metric, P, Cache, JCache, Global)
    torch.cuda.synchronize()
    Vo = U0.copy()
    orc.rw_step(Vo, 300.0, nsteps)
    Vg = V.cpu().numpy()
    for c in range(7):
        assert rel_l2(Vg[c], Vo[c]) <= 1e-12, f"V{c + 1}: {rel_l2(Vg[c], Vo[c]):.3e}"
    # the moist pressure matters: a dry run of the same state differs
    FE0, Metric0, Global0, Cache0, JCache0 = host.device_setup(disc, P, device="cuda", NumTr=2)
    W = torch.as_tensor(U0, device="cuda").clone()
    host.TimeStepper(W, 300.0, nsteps, FE0, Metric0, P, Cache0, JCache0, Global0)
    torch.cuda.synchronize()
    assert rel_l2(W[3].cpu().numpy(), Vg[3]) > 1e-6
