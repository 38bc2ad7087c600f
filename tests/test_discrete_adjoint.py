"""The discrete adjoint on the GPU (SP_OPT_ADJOINT 1: the exact transpose
of the forward time stepping, edge bands and sponge included): kernels
against the oracle's discrete adjoint (oracle.orc_adjoint_exact, itself
pinned by finite differences beside a free surface in
tests/test_oracle_pins.py), and the autograd gradient against central
differences of the float32 solver with the source and the receivers two
cells below a free top and the model perturbed everywhere, sponges
included -- the case where the reference-style backward is not the
gradient."""
import numpy as np
import pytest

from oracle import oracle as O
from seispie_amd import _lib, configs
from tests.helpers import oracle_model, rel_l2

pytestmark = pytest.mark.gpu


def _surface_problem(physics, nx=70, nz=61, nt=300, nshots=2, seed=0):
    """Smooth random model, top free, sponges on the other sides; sources
    and receivers two cells below the surface (two receivers share a cell)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    sr = lambda: (lambda a: (a / np.abs(a).max()).reshape(-1))(gaussian_filter(rng.standard_normal((nx, nz)), 3))
    vp = (3000 * (1 + 0.05 * sr())).astype(np.float32)
    vs = (1730 * (1 + 0.05 * sr())).astype(np.float32)
    rho = (2600 * (1 + 0.05 * sr())).astype(np.float32)
    comps = configs.ricker_stf(nt, 5e-4, 30.0, ang=30.0)
    stf = configs.shot_stf(physics, comps)
    rec = np.array([i * nz + 2 for i in range(6, nx - 6, 4)] + [30 * nz + 2], np.int32)
    shots = [configs.Shot(np.array([(20 + 25 * s) * nz + 2], np.int32), stf, rec) for s in range(nshots)]
    g = configs.Grid(nx, nz, 5.0, 5.0, [vp, vs, rho], 0, sponge=(12, 0.03, 1, 1, 1, 0), shots=shots)
    return configs.Problem(physics, nt, 5e-4, [g], sae=1)


def _spin_surface_problem(nx=70, nz=61, nt=300, nshots=2, seed=0):
    """The spin counterpart of _surface_problem: smooth random lam, mu, nu,
    j, mu_c, nu_c, rho (lam_c unused by the equations)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    sr = lambda: (lambda a: (a / np.abs(a).max()).reshape(-1))(gaussian_filter(rng.standard_normal((nx, nz)), 3))
    n = nx * nz
    f = lambda v, a: (v * (1 + a * sr())).astype(np.float32)
    fields = [f(1.0e10, 0.05), f(8.0e9, 0.05), f(1.0e9, 0.1), f(2700.0, 0.1), np.zeros(n, np.float32),
              f(1.5e8, 0.1), f(3e8, 0.1), f(2600.0, 0.05)]
    comps = configs.ricker_stf(nt, 5e-4, 30.0, ang=30.0)
    stf = configs.shot_stf(_lib.SPIN, comps)
    rec = np.array([i * nz + 2 for i in range(6, nx - 6, 4)] + [30 * nz + 2], np.int32)
    shots = [configs.Shot(np.array([(20 + 25 * s) * nz + 2], np.int32), stf, rec) for s in range(nshots)]
    g = configs.Grid(nx, nz, 5.0, 5.0, fields, 1, sponge=(12, 0.03, 1, 1, 1, 0), shots=shots)
    return configs.Problem(_lib.SPIN, nt, 5e-4, [g], sae=1)


@pytest.mark.parametrize("physics", [_lib.SH, _lib.PSV, _lib.SPIN])
def test_discrete_adjoint_kernels_match_the_oracle(ctx_factory, physics):
    prob = _spin_surface_problem() if physics == _lib.SPIN else _surface_problem(physics)
    g = prob.grids[0]
    comps = ["y"] if physics == _lib.SH else ["x", "z"]
    rng = np.random.default_rng(7)
    W = [rng.standard_normal((len(comps), s.rec_k.size, prob.nt)).astype(np.float32) for s in g.shots]
    print("case: discrete adjoint vs oracle, physics", physics, flush=True)
    ctx = ctx_factory()
    ctx.set_option(_lib.OPT_ADJOINT, 1)
    ids = configs.load_problem(ctx, prob)
    ctx.forward_ckpt(20)
    for i, w in zip(ids, W):
        ctx.set_adjoint_source(i, -w[:, :, ::-1])   # dL/d records, time-reversed and negated
    ctx.adjoint_ckpt(1)
    kids = [("lam", _lib.K_LAM), ("mu", _lib.K_MU), ("rho", _lib.K_RHO)]
    if physics == _lib.SPIN:
        kids += [("nu", _lib.K_NU), ("j", _lib.K_J), ("muc", _lib.K_MUC)]
    got = {nm: ctx.kernel(0, k) for nm, k in kids}
    K = None
    from tests.helpers import oracle_sim, oracle_stf
    for s, w in zip(g.shots, W):
        sim = oracle_sim(prob, g, s)
        _, snaps = sim.forward(oracle_stf(prob, s), sae=1, precision="f32")
        k1 = sim.adjoint_exact({c: w[n] for n, c in enumerate(comps)}, snaps, "f32")
        K = k1 if K is None else {nm: K[nm] + k1[nm] for nm in K}
    for nm in (["mu", "rho"] if physics == _lib.SH else [k for k, _ in kids]):
        assert np.abs(K[nm]).max() > 0
        err = rel_l2(got[nm], K[nm])
        print("%s rel L2 %.2e" % (nm, err))
        assert err <= 1e-4, nm
    with pytest.raises(_lib.SeispieError, match="adjoint_interval must be 1"):
        ctx.adjoint_ckpt(5)
    with pytest.raises(_lib.SeispieError, match="sp_forward_ckpt"):
        ctx.gradient(1)


@pytest.mark.parametrize("physics", [_lib.SH, _lib.PSV])
def test_discrete_adjoint_autograd_matches_finite_differences_beside_a_free_surface(physics):
    """L linear in the records; directional derivatives along smooth
    perturbations of the whole model (edge rows and sponges included) from
    loss.backward() against central differences of the float32 solver at
    eps and 2 eps combined by Richardson extrapolation."""
    import torch
    from scipy.ndimage import gaussian_filter
    from seispie_amd.autograd import Simulation
    prob = _surface_problem(physics, nshots=1)
    g = prob.grids[0]
    sim = Simulation(physics, g.nx, g.nz, g.dx, g.dz, prob.nt, prob.dt,
                     [(s.src_k, s.stf, s.rec_k) for s in g.shots], sponge=g.sponge, interval=30, adjoint="discrete")
    dev = sim.device
    rng = np.random.default_rng(5)
    m0 = [torch.tensor(f, device=dev, dtype=torch.float64) for f in g.fields]
    r = torch.tensor(rng.standard_normal((1, sim.ncomp, g.shots[0].rec_k.size, prob.nt)), device=dev)

    def loss_at(m):
        return (sim(*[t.float() for t in m]).double() * r).sum()

    leaves = [t.float().clone().requires_grad_(True) for t in m0]
    loss_at(leaves).backward()
    grads = [t.grad.double() for t in leaves]
    eps = 5e-3
    for f in ([1, 2] if physics == _lib.SH else [0, 1, 2]):
        print("case: discrete-adjoint fd, physics", physics, "field", f, flush=True)
        d = gaussian_filter(rng.standard_normal((g.nx, g.nz)), 4).reshape(-1)
        d = torch.tensor(d / np.abs(d).max(), device=dev)
        quot = []
        with torch.no_grad():
            for h in (eps, 2 * eps):
                vals = []
                for sgn in (1, -1):
                    m = [t.clone() for t in m0]
                    m[f] = m0[f] * (1 + sgn * h * d)
                    vals.append(float(loss_at(m)))
                quot.append((vals[0] - vals[1]) / (2 * h))
        fd = (4 * quot[0] - quot[1]) / 3
        adj = float((grads[f] * m0[f] * d).sum())
        print("fd(eps) %.6e fd(2 eps) %.6e extrapolated %.6e adj %.6e" % (quot[0], quot[1], fd, adj))
        assert adj == pytest.approx(fd, rel=1e-3)
    sim.close()


def test_discrete_adjoint_spin_autograd_matches_finite_differences_beside_a_free_surface():
    """Spin through SpinRecords with adjoint="discrete": loss.backward()
    against Richardson-extrapolated central differences of the float32
    solver for mu and rho, sources and receivers beside a free top, the
    whole model perturbed.  The nu, j and mu_c derivatives of this loss are
    four orders of magnitude smaller, below what float32 central
    differences resolve (at eps and 2 eps they differ by 20%); those kernels
    are pinned by test_discrete_adjoint_kernels_match_the_oracle (GPU equal
    to the oracle's discrete adjoint within 1e-6) and by the oracle's own
    float64 finite differences (tests/test_oracle_pins.py)."""
    import torch
    from scipy.ndimage import gaussian_filter
    from seispie_amd.autograd import Simulation
    prob = _spin_surface_problem(nshots=1)
    g = prob.grids[0]
    sim = Simulation(_lib.SPIN, g.nx, g.nz, g.dx, g.dz, prob.nt, prob.dt,
                     [(s.src_k, s.stf, s.rec_k) for s in g.shots], sponge=g.sponge, interval=30, adjoint="discrete")
    dev = sim.device
    rng = np.random.default_rng(5)
    m0 = [torch.tensor(f, device=dev, dtype=torch.float64) for f in g.fields]
    r = torch.tensor(rng.standard_normal((1, 2, g.shots[0].rec_k.size, prob.nt)), device=dev)

    def loss_at(m):
        return (sim(*[t.float() for t in m])[:, :2].double() * r).sum()

    leaves = [t.float().clone().requires_grad_(True) for t in m0]
    loss_at(leaves).backward()
    grads = [t.grad.double() for t in leaves]
    eps = 5e-3
    for f in (1, 7):   # mu, rho
        print("case: spin discrete-adjoint fd, field", f, flush=True)
        d = gaussian_filter(rng.standard_normal((g.nx, g.nz)), 4).reshape(-1)
        d = torch.tensor(d / np.abs(d).max(), device=dev)
        quot = []
        with torch.no_grad():
            for h in (eps, 2 * eps):
                vals = []
                for sgn in (1, -1):
                    m = [t.clone() for t in m0]
                    m[f] = m0[f] * (1 + sgn * h * d)
                    vals.append(float(loss_at(m)))
                quot.append((vals[0] - vals[1]) / (2 * h))
        fd = (4 * quot[0] - quot[1]) / 3
        adj = float((grads[f] * m0[f] * d).sum())
        print("fd(eps) %.6e fd(2 eps) %.6e extrapolated %.6e adj %.6e" % (quot[0], quot[1], fd, adj))
        assert adj == pytest.approx(fd, rel=1e-3)
    sim.close()
