"""Differentiable forward modelling: torch.autograd over the HIP solver.

    sim = Simulation(PSV, nx, nz, dx, dz, nt, dt, shots, sponge=(20, 0.015, 1, 1, 1, 0))
    rec = sim(vp, vs, rho)            # [nshot, ncomp, nrec, nt] on the GPU
    loss = some_function(rec)
    loss.backward()                   # vp.grad, vs.grad, rho.grad

The forward pass is the library's batched time loop over every shot
(sp_forward_ckpt), which keeps the full wavefield state every `interval`
steps.  The backward pass is the hand-written adjoint (sp_adjoint_ckpt): the
adjoint sources are -dL/d(records), time-reversed, injected at the receivers
as the reference injects its waveform adjoint sources (solver.py:290-321,
misfit/waveform.py:10,25); the forward wavefield it correlates with is
recomputed segment by segment from the checkpoints, last segment first, so
device memory holds nt/interval states and interval/sae snapshots per shot
instead of the nt/sae snapshots of the reference's stored-wavefield adjoint
(solver.py:193-203).

The accumulated kernels satisfy K = dt * dL/dm (lam, mu, rho), as the
reference's k_mu does for its waveform misfit, so the gradient is K / dt,
chained to vp, vs, rho through vps2lm (fd2d.py:21-33):
    dL/dvp  = 2 rho vp K_lam
    dL/dvs  = 2 rho vs (K_mu - 2 K_lam)
    dL/drho = (vp^2 - 2 vs^2) K_lam + vs^2 K_mu + K_rho.
Sampling the correlation every `sae` steps (the reference's
adjoint_interval) approximates the time integral; sae = 1 is the default.
Inside sponge layers the velocities are damped every step, so there the
density kernel (adjoint velocity times forward velocity) is approximate;
lam and mu are not affected.  That backward reuses the forward operator, as
the reference's adjoint does (solver.py:308-321), so next to a free surface,
where the difference operators' edge bands make D- differ from -(D+)^T, it
is not the gradient of L either.

adjoint="discrete" (SH, P-SV and spin) makes the backward the library's
discrete adjoint instead: the exact transpose of the forward time stepping, edge
bands and sponge included, correlated every step (sae = 1), so K / dt is
dL/dm up to float32 rounding anywhere in the grid, free surface and sponges
included.

SH uses mu and rho only (its vp gradient is zero); P-SV uses all three.

Micropolar-spin shots take the reference's spin fields instead (fd2d.py:172):
    rec = sim(lam, mu, nu, j, lam_c, mu_c, nu_c, rho)   # [nshot, 3, nrec, nt]
with gradients K/dt from the library's spin kernels: lam, mu, nu, j and rho
from their own kernels; mu_c and nu_c both from the (mu_c + nu_c) kernel,
since they enter the equations only as that sum (solver.py:140-141); lam_c
is unused by the equations (fd2d.py:210), so its gradient is zero.  Only
the x and z records can carry a loss: the adjoint injects x and z sources,
as the misfit does, and a nonzero gradient on the rotation records raises.
"""
import math

import numpy as np
import torch

from . import _lib


def default_interval(nt):
    """Even checkpoint interval near sqrt(nt): memory ~ 2 sqrt(nt) fields."""
    c = max(2, int(round(math.sqrt(nt))))
    return c + (c & 1)


class Simulation:
    """A grid with a fixed set of shots whose records are differentiable in
    the model.  shots: list of (src_k [nsrc], stf [n_inj, nsrc, nt], rec_k
    [nrec]) with grid indices k = i*nz + j; every shot has the same nrec."""

    def __init__(self, physics, nx, nz, dx, dz, nt, dt, shots, sponge=None, interval=None, sae=1, device=0,
                 init_model=None, autotune=True, adjoint="reference"):
        if physics not in (_lib.SH, _lib.PSV, _lib.SPIN):
            raise ValueError("unknown physics %r" % (physics,))
        if adjoint not in ("reference", "discrete"):
            raise ValueError("adjoint must be 'reference' or 'discrete'")
        if adjoint == "discrete":
            if sae != 1:
                raise ValueError("the discrete adjoint correlates every step: sae must be 1")
        self.adjoint = adjoint
        if nt % sae:
            raise ValueError("nt must be a multiple of sae")
        nrec = {len(np.asarray(s[2]).reshape(-1)) for s in shots}
        if len(nrec) != 1:
            raise ValueError("every shot needs the same number of receivers")
        self.physics, self.nx, self.nz, self.nt, self.dt = physics, nx, nz, nt, float(dt)
        self.sae = int(sae)
        self.interval = int(interval or default_interval(nt))
        if self.interval < 2 or self.interval % 2:
            raise ValueError("interval must be even and >= 2")
        self.nrec = nrec.pop()
        self.device = torch.device("cuda", device)
        self.ctx = _lib.Context(device)
        self.ctx.set_physics(physics, nt, dt)
        self.ctx.set_option(_lib.OPT_ADJOINT, 1 if adjoint == "discrete" else 0)
        self.param = 1 if physics == _lib.SPIN else 0
        if init_model is None:
            ones = np.ones(nx * nz, "f4")
            init_model = ((3000 * ones, 1700 * ones, 2500 * ones) if self.param == 0 else
                          (1e10 * ones, 8e9 * ones, 1e9 * ones, 2700 * ones, 0 * ones, 1.5e8 * ones, 3e8 * ones,
                           2600 * ones))
        self.grid = self.ctx.add_grid(nx, nz, dx, dz, init_model, self.param)
        if sponge is not None:
            self.ctx.set_sponge(self.grid, *sponge)
        self.nshot = len(shots)
        for src_k, stf, rec_k in shots:
            self.ctx.add_shot(self.grid, src_k, stf, rec_k)
        if autotune:
            # work decomposition and steps per launch for this batch shape
            # (cached per process); results do not depend on the choice
            self.ctx.autotune(20)
        self.ncomp = _lib.N_REC[physics]
        self.nmis = _lib.N_MIS[physics]
        self._gen = 0      # forward runs so far: backward re-runs the forward if another came between
        self._last = None  # model of the last forward run

    def __call__(self, *fields):
        """(vp, vs, rho) for SH and P-SV; (lam, mu, nu, j, lam_c, mu_c,
        nu_c, rho) for spin.  Returns the records [nshot, ncomp, nrec, nt]."""
        if self.param == 1:
            if len(fields) != len(SPIN_FIELDS):
                raise ValueError("spin records take the 8 fields %s" % (SPIN_FIELDS,))
            return SpinRecords.apply(self, *fields)
        if len(fields) != 3:
            raise ValueError("SH / P-SV records take (vp, vs, rho)")
        return SeismicRecords.apply(*fields, self)

    # ------------------------------------------------------------ plumbing
    def _bind_stream(self):
        self.ctx.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def _set_model(self, *fields):
        self._bind_stream()
        flat = [t.detach().to(self.device, torch.float32).contiguous().reshape(-1) for t in fields]
        for t in flat:
            if t.numel() != self.nx * self.nz:
                raise ValueError("model fields must have nx*nz = %d cells" % (self.nx * self.nz))
        if self.param == 1:
            for f, t in zip(SPIN_FIELD_IDS, flat):
                self.ctx.set_field(self.grid, f, t)
        else:
            self.ctx.set_model_vps(self.grid, *flat)
        return flat

    def run_forward(self, *fields):
        flat = self._set_model(*fields)
        self.ctx.forward_ckpt(self.interval)
        self._gen += 1
        self._last = flat
        out = torch.empty((self.nshot, self.ncomp, self.nrec, self.nt), device=self.device, dtype=torch.float32)
        for i in range(self.nshot):
            self.ctx.records(i, out[i])
        return out, self._gen

    def run_adjoint(self, grad_records, which=(_lib.K_LAM, _lib.K_MU, _lib.K_RHO)):
        """Kernels `which` of L for dL/d(records) = grad_records."""
        self._bind_stream()
        ad = (-grad_records[:, :self.nmis]).flip(-1).to(torch.float32).contiguous()
        for i in range(self.nshot):
            self.ctx.set_adjoint_source(i, ad[i])
        self.ctx.adjoint_ckpt(self.sae)
        n = self.nx * self.nz
        ks = [torch.empty(n, device=self.device, dtype=torch.float32) for _ in which]
        for t, w in zip(ks, which):
            self.ctx.kernel(self.grid, w, t)
        return ks

    def _refresh(self, fn_ctx, fields):
        """Rebuild the checkpoints if another forward run replaced them."""
        if self._gen != fn_ctx.gen:
            self._set_model(*fields)
            self.ctx.forward_ckpt(self.interval)
            self._gen += 1
            fn_ctx.gen = self._gen

    def hessian(self):
        """Pseudo-Hessian diagonal of the last backward pass: dt * sum over
        shots and time of the forward particle-velocity energy, [nx*nz]."""
        t = torch.empty(self.nx * self.nz, device=self.device, dtype=torch.float32)
        self._bind_stream()
        self.ctx.kernel(self.grid, _lib.K_HESS, t)
        return t

    def close(self):
        self.ctx.close()


# the spin model fields in the reference's order (fd2d.py:172) and their ids
SPIN_FIELDS = ("lam", "mu", "nu", "j", "lam_c", "mu_c", "nu_c", "rho")
SPIN_FIELD_IDS = (_lib.F_LAM, _lib.F_MU, _lib.F_NU, _lib.F_J, _lib.F_LAMC, _lib.F_MUC, _lib.F_NUC, _lib.F_RHO)


class SeismicRecords(torch.autograd.Function):
    """records = F(vp, vs, rho); backward = the hand-written adjoint."""

    @staticmethod
    def forward(ctx, vp, vs, rho, sim):
        out, gen = sim.run_forward(vp, vs, rho)
        ctx.sim = sim
        ctx.gen = gen
        ctx.shape = vp.shape
        ctx.like = [(t.device, t.dtype) for t in (vp, vs, rho)]
        ctx.save_for_backward(*sim._last)
        return out

    @staticmethod
    def backward(ctx, grad):
        sim = ctx.sim
        vp, vs, rho = ctx.saved_tensors
        sim._refresh(ctx, (vp, vs, rho))
        kl, km, kr = sim.run_adjoint(grad.contiguous())
        inv = 1.0 / sim.dt
        vp64, vs64, rho64 = (t.double() for t in (vp, vs, rho))
        kl, km, kr = (k.double() * inv for k in (kl, km, kr))
        live = vp64 > vs64  # vps2lm sets lambda to 0 where vp <= vs
        kl_eff = torch.where(live, kl, torch.zeros_like(kl))
        gvp = 2.0 * rho64 * vp64 * kl_eff
        gvs = 2.0 * rho64 * vs64 * (km - 2.0 * kl_eff)
        grho = torch.where(live, vp64 * vp64 - 2.0 * vs64 * vs64, torch.zeros_like(vp64)) * kl + vs64 * vs64 * km + kr
        shape = ctx.shape
        g = [x.float().reshape(shape) for x in (gvp, gvs, grho)]
        # each gradient on its input's device and in its dtype (inputs may be
        # host or float64 tensors; the solver itself runs on the GPU in float32)
        g = [x.to(device=d, dtype=t) for x, (d, t) in zip(g, ctx.like)]
        return (g[0], g[1], g[2], None)


class SpinRecords(torch.autograd.Function):
    """records = F(lam, mu, nu, j, lam_c, mu_c, nu_c, rho) for spin shots;
    backward = the hand-written adjoint with the spin kernels."""

    @staticmethod
    def forward(ctx, sim, *fields):
        out, gen = sim.run_forward(*fields)
        ctx.sim = sim
        ctx.gen = gen
        ctx.shape = fields[0].shape
        ctx.like = [(t.device, t.dtype) for t in fields]
        ctx.save_for_backward(*sim._last)
        return out

    @staticmethod
    def backward(ctx, grad):
        sim = ctx.sim
        if bool((grad[:, sim.nmis:] != 0).any()):
            # the adjoint injects x and z sources only, as the misfit does
            raise NotImplementedError("spin: losses on the rotation (y) records cannot be differentiated; "
                                      "use the x and z components")
        sim._refresh(ctx, ctx.saved_tensors)
        kl, km, knu, kj, kc, kr = sim.run_adjoint(
            grad.contiguous(), (_lib.K_LAM, _lib.K_MU, _lib.K_NU, _lib.K_J, _lib.K_MUC, _lib.K_RHO))
        inv = 1.0 / sim.dt
        g = [(k.double() * inv).float().reshape(ctx.shape) for k in (kl, km, knu, kj)]
        gc = (kc.double() * inv).float().reshape(ctx.shape)
        gr = (kr.double() * inv).float().reshape(ctx.shape)
        out = [g[0], g[1], g[2], g[3], torch.zeros_like(gc), gc, gc.clone(), gr]
        return (None,) + tuple(x.to(device=d, dtype=t) for x, (d, t) in zip(out, ctx.like))


def waveform_loss(records, data, dt):
    """The reference's waveform misfit (misfit/waveform.py:5-34) in torch:
    sum over shots, misfit components and receivers of the squared tapered
    residual; records/data [nshot, ncomp, nrec, nt].  Its gradient drives
    the same adjoint sources as sp_misfit."""
    nt = records.shape[-1]
    t = torch.arange(nt, device=records.device, dtype=torch.float64) * dt
    t_end = (nt - 1) * dt
    tw = t_end / 10.0
    t_min, t_max = tw, t_end - tw
    taper = torch.ones_like(t)
    lo, hi = t <= t_min, t >= t_max
    taper[lo] = 0.5 + 0.5 * torch.cos(math.pi * (t_min - t[lo]) / tw)
    taper[hi] = 0.5 + 0.5 * torch.cos(math.pi * (t_max - t[hi]) / tw)
    res = (data - records).double() * taper
    return (res * res).sum()
