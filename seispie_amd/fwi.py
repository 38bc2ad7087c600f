"""Full-waveform inversion in vp on the batched solver, with the optional
learned step conditioner.

`VpSolver` is a solver plugin for seispie's optimizers
(seispie/optimize/base.py:17-69 calls compute_gradient, compute_misfit,
update_model, export_field, mu.copy_to_host and stream.synchronize).  The
reference's fd2d plugin inverts mu (fd2d.py:442-443); this one inverts P
velocity with vs = vp / sqrt(3) and density fixed, every shot in one
batched gradient pass:
  * misfit  = sum over shots and components of the waveform norm
    (misfit/waveform.py:28-34, summed as solver.py:362-371 does);
  * gradient = K_vp + K_vs / sqrt(3) (chain rule through vps2lm), smoothed
    with the reference's Gaussian (fd2d.py:349-352);
  * the pseudo-Hessian diagonal of the same pass (forward particle-velocity
    energy summed over shots and time) is kept in `hess` for the
    conditioner.

`LearnedPrecond` is the step conditioner: a small convolutional network
(seispie_amd.cnn, HIP kernels) maps the gradient and the pseudo-Hessian of
an iteration to a positive per-cell scale s, and the optimizers use
diag(s) as the initial inverse Hessian (L-BFGS) or step conditioner (CG).
It is trained on the solver's own pseudo-Hessians: the target scale is the
inverse of the source-side pseudo-Hessian times the receiver-side
illumination (the energy of one simulation that fires every receiver),
water-levelled and smoothed -- the illumination compensation the
unconditioned gradient lacks.  The receiver-side run is needed only to
make training targets; at inference the network sees what every gradient
pass already produces.
"""
import math
from time import perf_counter

import numpy as np

from . import _lib, cnn, configs

SQ3 = math.sqrt(3.0)


class _Field:
    """The `mu` attribute of the plugin contract: copy_to_host(dst, stream)."""

    def __init__(self, solver):
        self.solver = solver

    def copy_to_host(self, dst, stream=None):
        dst[...] = self.solver.vp.reshape(np.shape(dst))


class _Stream:
    def synchronize(self):
        pass


class VpSolver:
    """Batched P-SV solver inverting vp (vs = vp / sqrt(3), rho fixed)."""

    def __init__(self, prob, data, sigma=5, device=0, chunk=0, autotune=True, adjoint="reference", mute_top=0):
        if prob.physics != _lib.PSV or len(prob.grids) != 1 or prob.grids[0].param != 0:
            raise ValueError("VpSolver needs a one-grid P-SV problem in vp/vs/rho")
        if adjoint not in ("reference", "discrete"):
            raise ValueError("adjoint must be 'reference' or 'discrete'")
        g = prob.grids[0]
        self.prob = prob
        self.nx, self.nz = g.nx, g.nz
        self.sae = prob.sae
        self.sigma = int(sigma)
        self.adjoint = adjoint
        # gradient mute below a free surface: zero in the top mute_top rows,
        # a cosine ramp over the next mute_top rows (sources and receivers
        # there make the misfit far from linear in the model at their cells)
        self.mute = None
        if mute_top > 0:
            j = np.arange(self.nz, dtype=np.float64)
            ramp = np.clip((j - mute_top) / mute_top, 0.0, 1.0)
            self.mute = np.tile(0.5 - 0.5 * np.cos(np.pi * ramp), self.nx).astype(np.float32)
        self.ctx = _lib.Context(device)
        if adjoint == "discrete":
            # the exact transpose of the time stepping (edge bands, sponge):
            # the gradient also with sources and receivers beside a free surface
            self.ctx.set_option(_lib.OPT_ADJOINT, 1)
            c = max(2, int(round(math.sqrt(prob.nt))))
            self.interval = c + (c & 1)
        if chunk:
            self.ctx.set_option(_lib.OPT_CHUNK, chunk)
        self.ids = configs.load_problem(self.ctx, prob)
        for i, d in zip(self.ids, data):
            self.ctx.set_data(i, d)
        if autotune and not chunk:
            # work decomposition for this batch shape (cached per process)
            self.ctx.autotune(20)
        self.vp = np.ascontiguousarray(g.fields[0], np.float32).copy()
        self.rho = np.ascontiguousarray(g.fields[2], np.float32).copy()
        self.mu = _Field(self)
        self.stream = _Stream()
        self.hess = None
        self.n_gradients = 0
        self.n_misfits = 0
        self.exported = {}

    def _set(self, vp):
        self.vp = np.ascontiguousarray(vp, np.float32).reshape(-1).copy()
        self.ctx.set_model_vps(0, self.vp, (self.vp / np.float32(SQ3)).astype(np.float32), self.rho)

    def update_model(self, vp):
        self._set(vp)

    def compute_misfit(self):
        self.ctx.forward(0)
        self.n_misfits += 1
        return float(self.ctx.misfit().sum())

    def compute_gradient(self):
        if self.adjoint == "discrete":
            self.ctx.forward_ckpt(self.interval)
            norms = self.ctx.misfit()
            self.ctx.adjoint_ckpt(1)
        else:
            norms = self.ctx.gradient(self.sae)
        self.ctx.kernels_vps(0)
        g = self.ctx.kernel(0, _lib.K_VP).astype(np.float64) + self.ctx.kernel(0, _lib.K_VS) / SQ3
        g = g.astype(np.float32)
        if self.sigma > 0:
            g = self.ctx.smooth(g, self.nx, self.nz, self.sigma)
        if self.mute is not None:
            g = g * self.mute
        self.hess = self.ctx.kernel(0, _lib.K_HESS)
        self.n_gradients += 1
        return g, float(norms.sum()), self.vp.copy()

    def export_field(self, field, name, it):
        self.exported[(name, it)] = np.array(field, copy=True)

    def close(self):
        self.ctx.close()


def receiver_illumination(prob, vp, rho, device=0):
    """Pseudo-Hessian of one simulation that fires every receiver of shot 0
    at once with the shots' source function: the receiver-side illumination
    (forward particle-velocity energy summed over time), [nx*nz]."""
    g = prob.grids[0]
    s0 = g.shots[0]
    nrec = s0.rec_k.size
    stf = np.repeat(np.asarray(s0.stf, np.float32)[:, :1, :], nrec, axis=1)
    one = configs.Problem(prob.physics, prob.nt, prob.dt,
                          [configs.Grid(g.nx, g.nz, g.dx, g.dz, [vp, (vp / np.float32(SQ3)).astype(np.float32), rho],
                                        0, sponge=g.sponge,
                                        shots=[configs.Shot(np.asarray(s0.rec_k, np.int32), stf, s0.rec_k)])],
                          sae=prob.sae)
    ctx = _lib.Context(device)
    try:
        ids = configs.load_problem(ctx, one)
        ctx.set_data(ids[0], np.zeros((_lib.N_REC[prob.physics], nrec, prob.nt), np.float32))
        ctx.gradient(prob.sae)
        return ctx.kernel(0, _lib.K_HESS)
    finally:
        ctx.close()


class LearnedPrecond:
    """Learned step conditioner: s = exp(L * net([g / max|g|, log-pseudo-Hessian, depth])),
    L = ln(1 / water), the network's output scaled to about [-1, 0]."""

    def __init__(self, solver, channels=(3, 32, 32, 32, 1), device=0, seed=0, water=1e-4, smooth=5, smax=1e4,
                 train_steps=1500, lr=3e-3):
        self.solver = solver
        self.device = device
        self.nx, self.nz = solver.nx, solver.nz
        self.net = cnn.Net(list(channels), device=device, seed=seed)
        self.water = water
        self.smooth = smooth
        self.smax = smax
        self.train_steps = train_steps
        self.lr = lr
        self.trained = False
        self.scale = None
        self.train_seconds = 0.0
        self.prep_seconds = 0.0
        self.train_loss = None
        self.fallbacks = 0

    # ----------------------------------------------------------- features
    def _smooth(self, a):
        """The solver's Gaussian smoothing (fd2d.py:71-108) on the device."""
        out = self.solver.ctx.smooth(np.ascontiguousarray(a, np.float32).reshape(-1), self.nx, self.nz, self.smooth)
        return out.reshape(self.nx, self.nz)

    def features(self, g, hess):
        g = np.asarray(g, np.float64).reshape(self.nx, self.nz)
        h = np.asarray(hess, np.float64).reshape(self.nx, self.nz)
        x = np.empty((self.net.channels[0], self.nx, self.nz), np.float32)
        x[0] = g / max(np.abs(g).max(), 1e-30)
        hn = h / max(h.max(), 1e-30)
        # smoothed like the target: the network's 9 x 9 receptive field
        # could not reproduce a Gaussian of this width itself
        x[1] = self._smooth(np.log10(hn + 1e-6) / 6.0 + 1.0)
        if x.shape[0] > 2:
            x[2] = (np.arange(self.nz) / max(self.nz - 1, 1))[None, :]
        return x

    def target(self, hess, hrec):
        """log of the water-levelled inverse of the source x receiver
        illumination, smoothed, with its maximum at 0."""
        return illumination_target(self, hess, hrec)

    # ----------------------------------------------------------- training
    def train(self, samples, steps=300, lr=2e-3, crop=128, seed=0, log_every=0):
        """samples: list of (g, hess, hrec).  Adam on random crops (and
        their mirror images in x) of each sample; returns the loss history."""
        rng = np.random.default_rng(seed)
        data = [(self.features(g, h), self.target(h, r)) for g, h, r in samples]
        t0 = perf_counter()
        hist = []
        self.trained = True
        cx, cz = min(crop, self.nx), min(crop, self.nz)
        for it in range(steps):
            x, t = data[it % len(data)]
            i0 = rng.integers(0, self.nx - cx + 1)
            j0 = rng.integers(0, self.nz - cz + 1)
            xc = x[:, i0:i0 + cx, j0:j0 + cz]
            tc = t[i0:i0 + cx, j0:j0 + cz]
            if rng.random() < 0.5:
                xc, tc = xc[:, ::-1], tc[::-1]
            loss = self.net.loss_grad(np.ascontiguousarray(xc), np.ascontiguousarray(tc))
            self.net.adam_step(lr)
            hist.append(loss)
            if log_every and it % log_every == 0:
                print("  precond training step %d loss %.4e" % (it, loss))
        self.net.sync()
        self.train_seconds += perf_counter() - t0
        return hist

    def evaluate(self, g, hess, hrec, border=0):
        """Mean squared error of the network's full-grid output against the
        target, in natural-log units of the scale (cells within `border` of
        the edges left out)."""
        y = self.net.forward(self.features(g, hess))[0]
        t = self.target(hess, hrec)
        b = border
        d = (y - t)[b:self.nx - b, b:self.nz - b] if b else (y - t)
        return float(np.mean(d.astype(np.float64) ** 2) * math.log(1.0 / self.water) ** 2)

    # ---------------------------------------------------------- inference
    def prepare(self, g):
        """Train on the first gradient pass: its gradient and pseudo-Hessian,
        and one receiver-side illumination run in the same model."""
        t0 = perf_counter()
        hrec = receiver_illumination(self.solver.prob, self.solver.vp, self.solver.rho, self.device)
        hist = self.train([(g, self.solver.hess, hrec)], steps=self.train_steps, lr=self.lr)
        self.train_loss = (hist[0], hist[-1])
        self.trained = True
        self.prep_seconds += perf_counter() - t0

    def refresh(self, g):
        """Scale field of the current iteration from its gradient and the
        solver's pseudo-Hessian of the same pass (training first, on the
        first call, unless train() was called)."""
        if not self.trained:
            self.prepare(g)
        y = self.net.forward(self.features(g, self.solver.hess))[0].astype(np.float64) * math.log(1.0 / self.water)
        s = np.exp(np.clip(y, -math.log(self.smax), 0.0)).reshape(-1)
        self.scale = s if np.all(np.isfinite(s)) else None
        if self.scale is None:
            self.fallbacks += 1

    def __call__(self, q):
        if self.scale is None:
            return q
        return self.scale * q

    def close(self):
        self.net.close()


def illumination_target(pre, hess, hrec):
    """log of the water-levelled inverse of the source x receiver
    illumination (hrec None: the source side alone), smoothed on the device,
    scaled by ln(1/water) with its maximum at 0."""
    h = np.asarray(hess, np.float64).reshape(pre.nx, pre.nz)
    hr = h / h.max()
    if hrec is not None:
        r = np.asarray(hrec, np.float64).reshape(pre.nx, pre.nz)
        hr = hr * (r / r.max())
    t = pre._smooth(-np.log(hr + pre.water)).astype(np.float64)
    return ((t - t.max()) / math.log(1.0 / pre.water)).astype(np.float32)


class IlluminationPrecond:
    """Non-learned step conditioners for the ablation of LearnedPrecond:
    the same scale, computed directly each iteration instead of predicted.
      mode "exact":  the network's training target, recomputed in the
                     current model every iteration (one receiver-side
                     illumination simulation per iteration);
      mode "source": the source-side pseudo-Hessian alone (no extra run)."""

    def __init__(self, solver, mode="exact", device=0, water=1e-4, smooth=5, smax=1e4):
        if mode not in ("exact", "source"):
            raise ValueError("mode must be 'exact' or 'source'")
        self.solver, self.mode, self.device = solver, mode, device
        self.nx, self.nz = solver.nx, solver.nz
        self.water, self.smooth, self.smax = water, smooth, smax
        self.scale = None
        self.prep_seconds = 0.0
        self.train_seconds = 0.0
        self.train_loss = (0.0, 0.0)
        self.fallbacks = 0

    _smooth = LearnedPrecond._smooth

    def refresh(self, g):
        t0 = perf_counter()
        hrec = None
        if self.mode == "exact":
            hrec = receiver_illumination(self.solver.prob, self.solver.vp, self.solver.rho, self.device)
        y = illumination_target(self, self.solver.hess, hrec).astype(np.float64) * math.log(1.0 / self.water)
        s = np.exp(np.clip(y, -math.log(self.smax), 0.0)).reshape(-1)
        self.scale = s if np.all(np.isfinite(s)) else None
        if self.scale is None:
            self.fallbacks += 1
        self.prep_seconds += perf_counter() - t0

    def __call__(self, q):
        return q if self.scale is None else self.scale * q

    def close(self):
        pass


def make_optimizer(solver, method="lbfgs", niter=20, precond=None, quiet=True):
    """An optimizer of seispie_amd.optimize with the reference example's
    settings (examples/inverse-checker/config.ini:14-21)."""
    from importlib import import_module
    opt = getattr(import_module("seispie_amd.optimize." + method), method)()
    opt.config = {"niter": str(niter), "lbfgs_mem": "5", "ls_step": "10", "ls_step_max": "0.5",
                  "ls_step_init": "0.05", "ls_thresh": "1.2"}
    opt.path = {}
    opt.mpi = None

    class _W:
        pass

    w = _W()
    w.solver = solver
    opt.setup(w)
    opt.precond = precond
    if quiet:
        opt.bracket.head = 0
        opt.quiet = True
    return opt


def observed(prob_true, device=0):
    """Records of every shot in the true model."""
    ctx = _lib.Context(device)
    try:
        ids = configs.load_problem(ctx, prob_true)
        ctx.forward(0)
        return [ctx.records(i) for i in ids]
    finally:
        ctx.close()


def time_to_misfit(prob_true, prob_init, niter=20, train_steps=1500, method="lbfgs", device=0, sigma=5, log=print,
                   data=None, adjoint="reference", mute_top=0):
    """Plain L-BFGS for niter iterations, then L-BFGS with the learned
    conditioner: the wall time the conditioned run needs to reach the plain
    run's final misfit, as a fraction of the plain run's wall time.  The
    conditioner trains inside its run, on the first gradient pass (plus one
    receiver-side illumination run), so `time_to_misfit` includes training;
    `time_to_misfit_without_training` subtracts it."""
    data = observed(prob_true, device) if data is None else data
    plain = VpSolver(prob_init, data, sigma=sigma, device=device, adjoint=adjoint, mute_top=mute_top)
    opt = make_optimizer(plain, method, niter)
    opt.on_iteration = lambda i, t, f: log("  plain iteration %d: %.2f s, misfit %.5e" % (i, t, f))
    t0 = perf_counter()
    opt.run()
    t_plain = perf_counter() - t0
    hist_plain = list(opt.history)
    f_target = hist_plain[-1][1]
    log("plain %s: %d iterations, misfit %.4e -> %.4e in %.2f s (%d gradients, %d misfit runs)"
        % (method, niter, hist_plain[0][1], f_target, t_plain, plain.n_gradients, plain.n_misfits))
    plain.close()

    solver = VpSolver(prob_init, data, sigma=sigma, device=device, adjoint=adjoint, mute_top=mute_top)
    pre = LearnedPrecond(solver, device=device, train_steps=train_steps)
    opt2 = make_optimizer(solver, method, niter, precond=pre)
    opt2.on_iteration = lambda i, t, f: log("  conditioned iteration %d: %.2f s, misfit %.5e" % (i, t, f))
    t2 = perf_counter()
    opt2.run()
    t_learned = perf_counter() - t2
    hist = list(opt2.history)
    log("conditioner: prepared in %.2f s inside the run (illumination run + %d training steps in %.2f s, "
        "loss %.3e -> %.3e)" % (pre.prep_seconds, train_steps, pre.train_seconds, *pre.train_loss))
    reach = next((t for t, f in hist if f <= f_target), None)
    log("learned %s: misfit %.4e -> %.4e in %.2f s; reaches %.4e at %s s"
        % (method, hist[0][1], hist[-1][1], t_learned, f_target, "%.2f" % reach if reach is not None else "never"))
    solver.close()
    pre.close()
    return {
        "plain_seconds": t_plain, "plain_final_misfit": f_target, "plain_history": hist_plain,
        "learned_seconds": t_learned, "learned_history": hist, "learned_final_misfit": hist[-1][1],
        "prep_seconds": pre.prep_seconds, "train_seconds": pre.train_seconds,
        "train_loss": pre.train_loss,
        "time_to_misfit": None if reach is None else reach / t_plain,
        "time_to_misfit_without_training": None if reach is None else (reach - pre.prep_seconds) / t_plain,
        "fallbacks": pre.fallbacks,
    }
