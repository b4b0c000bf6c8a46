"""HorLimit on the GPU (limited horizontal tracer transport, FP64) against the
oracle, through the C ABI.

LimitKernel! and DivRhoTrUpwind3LimKernel! (HorLimiterKernel.jl:2-196) replace
the tracer transport and its hyperviscosity when Model.HorLimit is set: the
tracer tendencies and the Rosenbrock-W steps match the oracle to 1e-12
relative L2 per component; the dynamics are those of a run without tracers,
bit for bit; and at a size beyond the oracle's reach an explicit update over
dtau with the limited tendency keeps a step in the mixing ratio inside its
initial range (to rounding).
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


def lim_state(pkg, disc, seed, ntr):
    """Perturbed jet plus tracers: a noisy mixing ratio, a step in longitude,
    a constant mixing ratio."""
    D = pkg.discretization
    U0 = pkg.state.jet_state(disc, seed=seed)
    rng = np.random.default_rng(seed)
    U0[1] += 3.0 * rng.standard_normal(U0[1].shape)
    U0[2] += 3.0 * rng.standard_normal(U0[2].shape)
    U0[3] += 0.3 * rng.standard_normal(U0[3].shape)
    U0[3][:, -1] = 0.0
    lon = D.node_lonlat(disc)[:, 0]
    step = np.where(np.abs(lon) < 0.8, 2e-3, 1e-3)[:, None] * np.ones((1, disc.nz))
    tr = [U0[0] * (0.01 + 0.002 * rng.standard_normal(U0[0].shape)), step * U0[0], 0.5 * U0[0]]
    return np.concatenate([U0] + [t[None] for t in tr[:ntr]], axis=0)


def setup(pkg, oracle_mod, ne, nz, ntr, seed=1, oracle=True):
    D, host = pkg.discretization, pkg.host
    disc = D.build(ne, 4, nz, 30000.0)
    phys = D.Phys()
    dev = host.device_setup(disc, phys, device="cuda", NumTr=ntr, nT=8 + ntr, horlimit=True)
    orc = None
    if oracle:
        hd = host.hyperdiffusion_for(ne)
        orc = oracle_mod.Oracle(disc, phys, (hd, hd, hd), NumTr=ntr, horlimit=True)
    return disc, phys, dev, orc, lim_state(pkg, disc, seed, ntr)


@pytest.mark.parametrize("ne,nz,ntr,dtau", [(3, 10, 3, 300.0), (4, 20, 2, 1200.0), (2, 70, 1, 150.0)])
def test_horlimit_tendency(pkg, oracle_mod, torch_cuda, ne, nz, ntr, dtau):
    """nz = 70 runs the 16-level chunks."""
    torch = torch_cuda
    host = pkg.host
    disc, phys, (FE, Metric, Global, Cache, JCache), orc, U0 = setup(pkg, oracle_mod, ne, nz, ntr)
    print(f"horlimit tendency ne={ne} nz={nz} NumTr={ntr} dtau={dtau}", flush=True)
    U = torch.as_tensor(U0, device="cuda")
    F = torch.empty_like(U)
    eq = host.CompressibleShallow()
    Global.dtauStage = dtau
    host.FcnPrepareGPU(U, FE, Metric, phys, Cache, None, Global, None, eq)
    host.FcnGPU(F, U, FE, Metric, phys, Cache, None, Global, None, eq)
    torch.cuda.synchronize()
    orc.dtau = dtau
    Fo, _ = orc.fcn(U0.copy(), orc.prepare(U0.copy()))
    Fg = F.cpu().numpy()
    for v in range(5 + ntr):
        assert rel_l2(Fg[v], Fo[v]) <= (1e-11 if v == 0 else TOL), f"F{v + 1}: {rel_l2(Fg[v], Fo[v]):.3e}"


def test_horlimit_unset_dtau_is_an_error(pkg, oracle_mod, torch_cuda):
    torch = torch_cuda
    host = pkg.host
    disc, phys, (FE, Metric, Global, Cache, JCache), _, U0 = setup(pkg, oracle_mod, 2, 8, 1, oracle=False)
    U = torch.as_tensor(U0, device="cuda")
    F = torch.empty_like(U)
    eq = host.CompressibleShallow()
    host.FcnPrepareGPU(U, FE, Metric, phys, Cache, None, Global, None, eq)
    with pytest.raises(host.CGDError, match="dtau must be positive"):
        host.FcnGPU(F, U, FE, Metric, phys, Cache, None, Global, None, eq)
    torch.cuda.synchronize()


@pytest.mark.parametrize("ne,nz,ntr,nsteps", [(3, 10, 3, 2), (4, 40, 2, 2)])
def test_horlimit_steps(pkg, oracle_mod, torch_cuda, ne, nz, ntr, nsteps):
    torch = torch_cuda
    host = pkg.host
    disc, phys, (FE, Metric, Global, Cache, JCache), orc, U0 = setup(pkg, oracle_mod, ne, nz, ntr)
    print(f"horlimit steps ne={ne} nz={nz} NumTr={ntr}", flush=True)
    V = torch.as_tensor(U0, device="cuda").clone()
    host.TimeStepper(V, 300.0, nsteps, FE, Metric, phys, Cache, JCache, Global)
    torch.cuda.synchronize()
    Vo = U0.copy()
    orc.rw_step(Vo, 300.0, nsteps)
    Vg = V.cpu().numpy()
    assert np.isfinite(Vg).all()
    for c in range(5 + ntr):
        e = rel_l2(Vg[c], Vo[c])
        assert e <= TOL, f"V{c + 1} rel L2 {e:.3e}"
    # passive: the dynamics equal a run without tracers bit for bit
    FE0, Metric0, Global0, Cache0, JCache0 = host.device_setup(disc, phys, device="cuda")
    W = torch.as_tensor(U0[:5], device="cuda").clone()
    host.TimeStepper(W, 300.0, nsteps, FE0, Metric0, phys, Cache0, JCache0, Global0)
    torch.cuda.synchronize()
    assert torch.equal(V[:5], W)


@pytest.mark.parametrize("ne,nz,dtaus", [(16, 32, (150.0, 300.0)), (30, 64, (150.0,))])
def test_horlimit_bounds_at_size(pkg, oracle_mod, torch_cuda, ne, nz, dtaus):
    """The step tracer after an explicit update over dtau with the limited
    tendency stays in [1e-3, 2e-3] (relative rounding 1e-12); the noisy and
    the constant mixing ratios stay finite and the constant one stays 0.5 (to
    1e-10).  The limiter bounds the horizontal transport only: dtau stays
    where the vertical Courant number of the w noise (0.3 m/s over layers of
    30 km / nz) is below 0.1, as the oracle requires too (at ne = 30, nz = 64
    and dtau = 300 its update leaves the range through the vertical flux)."""
    torch = torch_cuda
    host = pkg.host
    disc, phys, (FE, Metric, Global, Cache, JCache), _, U0 = setup(pkg, oracle_mod, ne, nz, 3, oracle=False)
    print(f"horlimit bounds ne={ne} nz={nz}", flush=True)
    U = torch.as_tensor(U0, device="cuda")
    F = torch.empty_like(U)
    eq = host.CompressibleShallow()
    for dtau in dtaus:
        Global.dtauStage = dtau
        host.FcnPrepareGPU(U, FE, Metric, phys, Cache, None, Global, None, eq)
        host.FcnGPU(F, U, FE, Metric, phys, Cache, None, Global, None, eq)
        rho = U[0] + dtau * F[0]
        q = (U[6] + dtau * F[6]) / rho
        assert bool(torch.isfinite(F).all())
        assert float(q.min()) >= 1e-3 * (1 - 1e-12), (dtau, float(q.min()))
        assert float(q.max()) <= 2e-3 * (1 + 1e-12), (dtau, float(q.max()))
        qc = (U[7] + dtau * F[7]) / rho
        assert float((qc - 0.5).abs().max()) <= 1e-10, dtau
    torch.cuda.synchronize()


def test_horlimit_composed_equals_fused(pkg, oracle_mod, torch_cuda):
    """host.RosenbrockSchur through a Python tendency callback (the composed
    call sequence, dtauStage = dt set by the stepper) equals the fused step."""
    torch = torch_cuda
    host = pkg.host
    disc, phys, (FE, Metric, Global, Cache, JCache), _, U0 = setup(pkg, oracle_mod, 3, 20, 2, oracle=False)
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
    assert Global.dtauStage == 300.0
    for c in range(7):
        assert rel_l2(V2[c].cpu().numpy(), V1[c].cpu().numpy()) <= 1e-14, f"V{c + 1}"


def test_horlimit_graph_replays_equal_eager_steps(pkg, oracle_mod, torch_cuda):
    """Two captured steps with tracers and HorLimit replayed three times, with
    eager steps and a synchronisation in between; each replay equals two
    eager steps bit for bit."""
    torch = torch_cuda
    host = pkg.host
    disc, phys, (FE, Metric, Global, Cache, JCache), _, U0 = setup(pkg, oracle_mod, 3, 20, 2, oracle=False)
    ctx = host.init_context(FE, Metric, phys, Global)
    G = torch.as_tensor(U0, device="cuda").clone()
    E = G.clone()
    ctx.graph_create(G, 150.0, Cache, JCache, nsteps=2)
    torch.cuda.synchronize()
    assert torch.equal(G, E), "capturing must not run the steps"
    for rep in range(3):
        ctx.graph_launch()
        torch.cuda.synchronize()
        Gr = G.clone()
        host.TimeStepper(E, 150.0, 2, FE, Metric, phys, Cache, JCache, Global)
        torch.cuda.synchronize()
        assert torch.equal(Gr, E), f"replay {rep} differs from the eager steps"
    assert torch.isfinite(G).all()
