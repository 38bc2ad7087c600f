"""Checkpointed adjoint and the torch.autograd.Function over it.

* sp_forward_ckpt + sp_adjoint_ckpt (forward wavefield recomputed from
  checkpoints) gives the same records and kernels, bit for bit, as the
  stored-snapshot sp_gradient, for every physics and for checkpoint
  intervals that do and do not divide by the correlation interval;
* the autograd gradient of the reference's waveform misfit equals the
  oracle's kernels chained to vp/vs/rho (K / dt, vps2lm chain rule);
* the autograd gradient of an arbitrary (linear) loss matches central
  finite differences of that loss through the forward solver.
"""
import numpy as np
import pytest

from seispie_amd import _lib, configs
from tests.helpers import oracle_gradient, rel_l2
from tests.test_gpu_parity import KERN_TOL, _data_for, _small

pytestmark = pytest.mark.gpu


def _ctx_with(ctx_factory, prob, data):
    ctx = ctx_factory()
    ids = configs.load_problem(ctx, prob)
    for i, d in zip(ids, data):
        ctx.set_data(i, d)
    return ctx, ids


@pytest.mark.parametrize("physics", [_lib.SH, _lib.PSV, _lib.SPIN])
@pytest.mark.parametrize("sae", [10, 5])
def test_checkpointed_adjoint_equals_stored_snapshots(ctx_factory, physics, sae):
    prob = _small(physics, nt=200, nshots=2)
    prob.sae = sae
    data = _data_for(prob)
    print("case: ckpt", physics, "sae", sae, flush=True)
    ctx, ids = _ctx_with(ctx_factory, prob, data)
    norms = ctx.gradient(prob.sae)
    ref_k = [ctx.kernel(0, w) for w in range(_lib.K_VP)] + [ctx.kernel(0, _lib.K_HESS)]
    ref_r = [ctx.records(i) for i in ids]
    assert np.abs(ref_k[_lib.K_HESS - 3]).max() > 0
    for C in (10, 6, 200, 2):
        print("interval", C, flush=True)
        ctx.forward_ckpt(C)
        for i in ids:
            np.testing.assert_array_equal(ctx.records(i), ref_r[i])
        n2 = ctx.misfit()
        np.testing.assert_array_equal(n2, norms)
        ctx.adjoint_ckpt(prob.sae)
        got = [ctx.kernel(0, w) for w in range(_lib.K_VP)] + [ctx.kernel(0, _lib.K_HESS)]
        for w, (a, b) in enumerate(zip(got, ref_k)):
            np.testing.assert_array_equal(a, b, err_msg="kernel %d interval %d" % (w, C))


@pytest.mark.parametrize("physics", [_lib.SH, _lib.PSV])
def test_checkpointed_adjoint_with_two_cells_per_lane(ctx_factory, physics):
    """The checkpointed forward and adjoint with SP_OPT_FUSE 2 (two cells per lane for
    the two-step launches, including the recomputed segments) give the
    stored-snapshot gradient of one cell per lane bit for bit."""
    prob = _small(physics, 60, 130, 200, 2)
    prob.sae = 10
    data = _data_for(prob)
    print("case: ckpt two cells per lane", physics, flush=True)
    ctx, ids = _ctx_with(ctx_factory, prob, data)
    ctx.set_option(_lib.OPT_CHUNK, 16)
    norms = ctx.gradient(prob.sae)
    ref_k = [ctx.kernel(0, w) for w in range(_lib.K_VP)] + [ctx.kernel(0, _lib.K_HESS)]
    ref_r = [ctx.records(i) for i in ids]
    ctx2, ids2 = _ctx_with(ctx_factory, prob, data)
    ctx2.set_option(_lib.OPT_FUSE, 2)
    ctx2.set_option(_lib.OPT_CHUNK, 16)
    for C in (10, 6):
        print("interval", C, flush=True)
        ctx2.forward_ckpt(C)
        for i in ids2:
            np.testing.assert_array_equal(ctx2.records(i), ref_r[i])
        np.testing.assert_array_equal(ctx2.misfit(), norms)
        ctx2.adjoint_ckpt(prob.sae)
        got = [ctx2.kernel(0, w) for w in range(_lib.K_VP)] + [ctx2.kernel(0, _lib.K_HESS)]
        for w, (x, y) in enumerate(zip(got, ref_k)):
            np.testing.assert_array_equal(x, y, err_msg="kernel %d interval %d" % (w, C))


def test_checkpointed_adjoint_under_the_graph_option(ctx_factory):
    """SP_OPT_GRAPH 1 leaves the checkpointed forward and adjoint eager: the
    same records and kernels as without it."""
    prob = _small(_lib.PSV, nt=200, nshots=2)
    prob.sae = 10
    data = _data_for(prob)
    print("case: ckpt with the graph option", flush=True)
    out = []
    for graph in (0, 1):
        ctx, ids = _ctx_with(ctx_factory, prob, data)
        ctx.set_option(_lib.OPT_GRAPH, graph)
        ctx.forward_ckpt(10)
        recs = [ctx.records(i) for i in ids]
        ctx.misfit()
        ctx.adjoint_ckpt(prob.sae)
        out.append((recs, [ctx.kernel(0, w) for w in (_lib.K_LAM, _lib.K_MU, _lib.K_RHO, _lib.K_HESS)]))
    for a, b in zip(out[0][0], out[1][0]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(out[0][1], out[1][1]):
        assert np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)


def test_checkpoint_errors(ctx_factory):
    prob = _small(_lib.PSV, nt=40, nshots=1)
    ctx = ctx_factory()
    configs.load_problem(ctx, prob)
    with pytest.raises(_lib.SeispieError, match="sp_forward_ckpt"):
        ctx.adjoint_ckpt(10)
    with pytest.raises(_lib.SeispieError, match="even"):
        ctx.forward_ckpt(3)
    ctx.forward_ckpt(8)
    with pytest.raises(_lib.SeispieError, match="multiple"):
        ctx.adjoint_ckpt(7)
    ctx.set_field(0, _lib.F_MU, ctx.get_field(0, _lib.F_MU))  # a model change drops the checkpoints
    with pytest.raises(_lib.SeispieError, match="sp_forward_ckpt"):
        ctx.adjoint_ckpt(10)


def _sim_from(prob, sae, interval=None):
    from seispie_amd.autograd import Simulation
    g = prob.grids[0]
    shots = [(s.src_k, s.stf, s.rec_k) for s in g.shots]
    return Simulation(prob.physics, g.nx, g.nz, g.dx, g.dz, prob.nt, prob.dt, shots, sponge=g.sponge,
                      interval=interval, sae=sae)


@pytest.mark.parametrize("physics", [_lib.SH, _lib.PSV])
def test_autograd_waveform_gradient_matches_oracle(physics):
    import torch
    from seispie_amd.autograd import waveform_loss
    prob = _small(physics, nt=200, nshots=2)
    prob.sae = 10
    data = _data_for(prob)
    print("case: autograd vs oracle", physics, flush=True)
    sim = _sim_from(prob, prob.sae, interval=20)
    dev = sim.device
    vp, vs, rho = (torch.tensor(f, device=dev, requires_grad=True) for f in prob.grids[0].fields)
    rec = sim(vp, vs, rho)
    d = torch.tensor(np.stack(data), device=dev)
    loss = waveform_loss(rec, d, prob.dt)
    loss.backward()
    kref, nref = oracle_gradient(prob, data)
    assert float(loss.detach()) == pytest.approx(float((nref ** 2).sum()), rel=1e-4)
    vp64, vs64, rho64 = (f.astype(np.float64) for f in prob.grids[0].fields)
    kl, km, kr = (kref[0][k].astype(np.float64) / prob.dt for k in ("lam", "mu", "rho"))
    want = {"vs": 2 * rho64 * vs64 * (km - 2 * kl), "rho": (vp64 ** 2 - 2 * vs64 ** 2) * kl + vs64 ** 2 * km + kr}
    got = {"vs": vs.grad, "rho": rho.grad}
    if physics == _lib.PSV:
        want["vp"] = 2 * rho64 * vp64 * kl
        got["vp"] = vp.grad
    else:
        assert float(vp.grad.abs().max()) == 0.0
    for k in want:
        err = rel_l2(got[k].cpu().numpy(), want[k])
        print(k, "gradient rel L2 %.3e" % err)
        assert err <= KERN_TOL, (k, err)
    h = sim.hessian().cpu().numpy()
    assert h.min() >= 0 and h.max() > 0
    sim.close()


def test_spin_autograd_gradient_matches_oracle():
    """Spin records through SpinRecords: the gradients of the waveform loss
    on the x and z records are the oracle's spin kernels / dt (mu_c and nu_c
    both the (mu_c + nu_c) kernel, lam_c zero); a loss on the rotation
    records raises."""
    import torch
    from seispie_amd.autograd import waveform_loss
    prob = _small(_lib.SPIN, nt=200, nshots=2)
    prob.sae = 10
    data = _data_for(prob)
    print("case: spin autograd vs oracle", flush=True)
    sim = _sim_from(prob, prob.sae, interval=20)
    dev = sim.device
    fields = [torch.tensor(f, device=dev, requires_grad=True) for f in prob.grids[0].fields]
    rec = sim(*fields)
    assert rec.shape[1] == 3
    d = torch.tensor(np.stack(data), device=dev)
    loss = waveform_loss(rec[:, :2], d[:, :2], prob.dt)
    loss.backward()
    kref, nref = oracle_gradient(prob, data)
    assert float(loss.detach()) == pytest.approx(float((nref ** 2).sum()), rel=1e-4)
    k = {n: kref[0][n].astype(np.float64) / prob.dt for n in ("lam", "mu", "nu", "j", "muc", "rho")}
    want = [k["lam"], k["mu"], k["nu"], k["j"], None, k["muc"], k["muc"], k["rho"]]
    for name, f, w in zip(("lam", "mu", "nu", "j", "lam_c", "mu_c", "nu_c", "rho"), fields, want):
        g = f.grad.cpu().numpy()
        if w is None:
            assert np.all(g == 0), name
            continue
        err = rel_l2(g, w)
        print(name, "gradient rel L2 %.3e" % err)
        assert err <= KERN_TOL, (name, err)
    rot = sim(*[f.detach().clone().requires_grad_(True) for f in fields])
    with pytest.raises(NotImplementedError, match="rotation"):
        rot[:, 2].sum().backward()
    sim.close()


def test_autograd_matches_finite_differences():
    """A loss linear in the records (white noise in time), sae = 1, sponges
    on all sides, interior receivers, model perturbations outside the
    sponge layers: directional derivatives from the adjoint match central
    differences of the float32 solver to 1e-3.  The central differences at
    eps and 2 eps are combined by Richardson extrapolation, which removes
    their O(eps^2) truncation error: for the vs direction that error is
    1.3e-3 at eps = 5e-3 (on the oracle in float64 the quotient moves from
    -6.2931 to -6.3937 between eps = 2e-3 and 2e-2; extrapolated to 0 it is
    -6.29207).  (Inside a sponge layer the
    velocities are damped every step, so there v_adj . v_fwd is not the
    density derivative: on the oracle a whole-grid density perturbation
    is off by 0.7-3%, an interior one by < 1e-4.)"""
    import torch
    from scipy.ndimage import gaussian_filter
    prob = _small(_lib.PSV, nx=90, nz=80, nt=300, nshots=1, sponge=(20, 0.03, 1, 1, 1, 1))
    g = prob.grids[0]
    # receivers well inside the sponges
    g.shots[0].rec_k = np.array([i * g.nz + 40 for i in range(25, 66, 5)], np.int32)
    g.shots[0].src_k = np.array([45 * g.nz + 55], np.int32)
    sim = _sim_from(prob, 1, interval=30)
    dev = sim.device
    rng = np.random.default_rng(5)
    m0 = [torch.tensor(f, device=dev, dtype=torch.float64) for f in g.fields]
    r = torch.tensor(rng.standard_normal((1, 2, g.shots[0].rec_k.size, prob.nt)), device=dev)

    def loss_at(m):
        rec = sim(*[t.float() for t in m])
        return (rec.double() * r).sum()

    leaves = [t.float().clone().requires_grad_(True) for t in m0]
    loss_at(leaves).backward()
    grads = [t.grad.double() for t in leaves]
    eps = 5e-3
    mask = np.zeros((g.nx, g.nz))
    mask[22:g.nx - 22, 22:g.nz - 22] = 1
    mask = gaussian_filter(mask, 2).reshape(-1)
    for f in range(3):
        print("case: fd field", f, flush=True)
        d = gaussian_filter(rng.standard_normal((g.nx, g.nz)), 4).reshape(-1) * mask
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


def test_autograd_backward_after_another_forward():
    """A second forward run replaces the checkpoints; the first graph's
    backward rebuilds them from its saved model."""
    import torch
    prob = _small(_lib.PSV, nt=100, nshots=1)
    sim = _sim_from(prob, 5, interval=10)
    dev = sim.device
    g = prob.grids[0]
    a = [torch.tensor(f, device=dev, requires_grad=True) for f in g.fields]
    b = [torch.tensor(f * 1.01, device=dev, requires_grad=True) for f in g.fields]
    ra = sim(*a)
    rb = sim(*b)
    ra.square().sum().backward()
    ga = [t.grad.clone() for t in a]
    a2 = [torch.tensor(f, device=dev, requires_grad=True) for f in g.fields]
    sim(*a2).square().sum().backward()
    for x, y in zip(ga, a2):
        torch.testing.assert_close(x, y.grad, rtol=0, atol=0)
    rb.square().sum().backward()
    assert all(torch.isfinite(t.grad).all() for t in b)
    sim.close()


@pytest.mark.parametrize("physics", [_lib.PSV, _lib.SPIN])
def test_autograd_accepts_host_and_float64_leaves(physics):
    """Model leaves on the host (float64 here) get gradients of their own
    device and dtype, equal to the float32 GPU leaves' gradients."""
    import torch
    prob = _small(physics, nt=100, nshots=1)
    sim = _sim_from(prob, 5, interval=10)
    g = prob.grids[0]
    gpu = [torch.tensor(f, device=sim.device, requires_grad=True) for f in g.fields]
    host = [torch.tensor(f.astype(np.float64), requires_grad=True) for f in g.fields]
    w = None
    for leaves in (gpu, host):
        rec = sim(*leaves)
        if w is None:
            w = torch.randn(rec.shape, generator=torch.Generator().manual_seed(3)).to(rec.device)
            w[:, sim.nmis:] = 0
        (rec * w).sum().backward()
    for a, b in zip(gpu, host):
        assert b.grad.device.type == "cpu" and b.grad.dtype == torch.float64
        torch.testing.assert_close(b.grad, a.grad.cpu().double(), rtol=0, atol=0)
    assert any(float(b.grad.abs().max()) > 0 for b in host)
    sim.close()
