"""The gradient (forward with snapshots, misfit, adjoint, kernel
accumulation: solver.py:290-404) at the benchmark workloads' full sizes.

* P-SV 1000^2, nt 4000, sae 50: shot 0 alone gives the oracle's kernels bit
  for bit (one shot per grid: the sums follow the oracle's order), with one
  and with two cells per lane, in strips and in column groups; its misfit
  norms agree to 1e-12 (float64 partial sums of the same float32 residuals,
  added in another order).
* The 128-shot batch (the timed workload: time-major kernel sums over
  128 shots, snapshot pool filled in sub-batches) equals, within the
  kernel tolerance, the 124-shot batch without shots {0, 41, 86, 127} plus
  those four shots' single-shot gradients; each shot's misfit equals its
  single-shot misfit.
* Micropolar spin 600^2, nt 3000: shot 0 alone equals the oracle's spin
  kernels bit for bit.
* P-SV 400^2 x 16 shots, nt 3000, through the torch.autograd.Function with
  sae 50: loss.backward() equals the oracle's kernels / dt chained to
  vp, vs, rho within the kernel tolerance.

Observed data are records of the workload's true model (bench.py's
pairing: smooth random or checkerboard true model, homogeneous initial
model).  The oracle runs on the host's cores (about 10-40 s per case).
"""
import numpy as np
import pytest

from seispie_amd import _lib, configs
from tests.helpers import oracle_gradient, rel_l2

pytestmark = pytest.mark.gpu

KERN_TOL = 1e-4


def _homogeneous(prob):
    for g in prob.grids:
        g.fields = [np.full_like(f, f.mean()) for f in g.fields]
    return prob


def _records(ctx_factory, prob):
    ctx = ctx_factory()
    ids = configs.load_problem(ctx, prob)
    ctx.forward(0)
    out = [ctx.records(i) for i in ids]
    ctx.close()
    return out


def _gradient(ctx_factory, prob, data, which, **opts):
    ctx = ctx_factory()
    for k, v in opts.items():
        ctx.set_option(getattr(_lib, "OPT_" + k.upper()), v)
    ids = configs.load_problem(ctx, prob)
    for i, d in zip(ids, data):
        ctx.set_data(i, d)
    norms = ctx.gradient(prob.sae)
    K = {nm: ctx.kernel(0, w) for nm, w in which}
    ctx.close()
    return K, norms


def _subset(prob, keep):
    g = prob.grids[0]
    return configs.Problem(prob.physics, prob.nt, prob.dt,
                           [configs.Grid(g.nx, g.nz, g.dx, g.dz, g.fields, g.param, g.sponge,
                                         [g.shots[i] for i in keep])], sae=prob.sae, name=prob.name)


PSV_K = [("lam", _lib.K_LAM), ("mu", _lib.K_MU), ("rho", _lib.K_RHO)]
SPIN_K = PSV_K + [("nu", _lib.K_NU), ("j", _lib.K_J), ("muc", _lib.K_MUC)]


@pytest.mark.timeout(600)
def test_benchmark_shot0_gradient_equals_the_oracle(ctx_factory):
    true = configs.psv_config3()
    init = _homogeneous(configs.psv_config3())
    one_t, one_i = _subset(true, [0]), _subset(init, [0])
    print("case: 1000^2 x 4000 shot 0 gradient vs oracle", flush=True)
    data = _records(ctx_factory, one_t)
    kref, nref = oracle_gradient(one_i, data)
    for opts in ({}, {"fuse": 2, "chunk": 128}, {"fuse": 4, "chunk": 64}):
        K, norms = _gradient(ctx_factory, one_i, data, PSV_K, **opts)
        np.testing.assert_allclose(norms, nref, rtol=1e-12)  # float64 partial sums in another order
        for nm, _ in PSV_K:
            assert np.abs(kref[0][nm]).max() > 0, nm
            np.testing.assert_array_equal(K[nm], kref[0][nm], err_msg="%s %s" % (nm, opts))
        print("options", opts, "kernels equal the oracle's", flush=True)


@pytest.mark.timeout(600)
def test_benchmark_batch_gradient_is_the_sum_of_its_shots(ctx_factory):
    true = configs.psv_config3()
    init = _homogeneous(configs.psv_config3())
    print("case: 1000^2 x 4000, 128-shot gradient vs 124-shot batch + 4 single shots", flush=True)
    data = _records(ctx_factory, true)
    K_all, n_all = _gradient(ctx_factory, init, data, PSV_K)
    print("128-shot batch done", flush=True)
    few = [0, 41, 86, 127]
    rest = [i for i in range(128) if i not in few]
    K_rest, n_rest = _gradient(ctx_factory, _subset(init, rest), [data[i] for i in rest], PSV_K)
    np.testing.assert_allclose(n_rest, n_all[rest], rtol=1e-12)
    total = {nm: K_rest[nm].astype(np.float64) for nm, _ in PSV_K}
    for i in few:
        K1, n1 = _gradient(ctx_factory, _subset(init, [i]), [data[i]], PSV_K)
        np.testing.assert_allclose(n1[0], n_all[i], rtol=1e-12)
        for nm, _ in PSV_K:
            total[nm] += K1[nm]
    for nm, _ in PSV_K:
        assert np.abs(K_all[nm]).max() > 0
        err = rel_l2(K_all[nm], total[nm])
        share = rel_l2(K_all[nm], K_rest[nm])  # what the four shots contribute
        print("%s: batch vs parts rel L2 %.2e (the four shots' share %.2e)" % (nm, err, share), flush=True)
        assert err <= KERN_TOL and share > 100 * err, nm


@pytest.mark.timeout(600)
def test_spin_shot0_gradient_equals_the_oracle(ctx_factory):
    true = configs.spin_config4()
    init = _homogeneous(configs.spin_config4())
    one_t, one_i = _subset(true, [0]), _subset(init, [0])
    print("case: spin 600^2 x 3000 shot 0 gradient vs oracle", flush=True)
    data = _records(ctx_factory, one_t)
    kref, nref = oracle_gradient(one_i, data)
    K, norms = _gradient(ctx_factory, one_i, data, SPIN_K)
    np.testing.assert_allclose(norms, nref, rtol=1e-12)  # float64 partial sums in another order
    for nm, _ in SPIN_K:
        assert np.abs(kref[0][nm]).max() > 0, nm
        np.testing.assert_array_equal(K[nm], kref[0][nm], err_msg=nm)


@pytest.mark.timeout(900)
def test_config2_autograd_gradient_matches_the_oracle(ctx_factory):
    import torch
    from seispie_amd.autograd import Simulation, waveform_loss
    true = configs.psv_config2(true_model=True)
    init = configs.psv_config2(true_model=False)
    print("case: 400^2 x 16 shots x 3000 through autograd vs oracle", flush=True)
    data = _records(ctx_factory, true)
    g = init.grids[0]
    sim = Simulation(init.physics, g.nx, g.nz, g.dx, g.dz, init.nt, init.dt,
                     [(s.src_k, s.stf, s.rec_k) for s in g.shots], sponge=g.sponge, sae=init.sae)
    dev = sim.device
    vp, vs, rho = (torch.tensor(f, device=dev, requires_grad=True) for f in g.fields)
    d = torch.tensor(np.stack([x[:sim.nmis] for x in data]), device=dev)
    loss = waveform_loss(sim(vp, vs, rho), d, init.dt)
    loss.backward()
    sim.close()
    print("autograd done; oracle gradient of 16 shots", flush=True)
    kref, nref = oracle_gradient(init, data)
    assert float(loss.detach()) == pytest.approx(float((nref ** 2).sum()), rel=1e-5)
    vp64, vs64, rho64 = (f.astype(np.float64) for f in g.fields)
    kl, km, kr = (kref[0][k].astype(np.float64) / init.dt for k in ("lam", "mu", "rho"))
    want = {"vp": 2 * rho64 * vp64 * kl, "vs": 2 * rho64 * vs64 * (km - 2 * kl),
            "rho": (vp64 ** 2 - 2 * vs64 ** 2) * kl + vs64 ** 2 * km + kr}
    for name, t in (("vp", vp), ("vs", vs), ("rho", rho)):
        err = rel_l2(t.grad.cpu().numpy(), want[name])
        print(name, "gradient rel L2 %.3e" % err, flush=True)
        assert err <= KERN_TOL, (name, err)
