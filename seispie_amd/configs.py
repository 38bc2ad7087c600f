"""Synthetic workloads: the five benchmark configurations of this project
(SH 201^2; P-SV 400^2 x 16 shots; P-SV 1000^2 x 128 shots plus a mixed
batch; micropolar spin 600^2 x 32 shots; the learned-preconditioner FWI
model on 500^2), built with the reference's own conventions:

  * grid index k = i*nz + j, dx/dz float32 scalars (fd2d.py:189-196);
  * Ricker sources (source/ricker.py:7-11) with t0 = 1.2/f0, amp = 1e10,
    sources at z = 10 m, receivers at z = 2 dz (grid index rounding of
    fd2d.py:132-134, 157-159).  Receivers sit two cells below the surface
    because the reference's difference operators are zero in their edge
    bands (solver.py:22-51): rows j = 0 (all physics) and j = 1 (SH) are
    decoupled from the interior, so a receiver there records exactly zero;
  * sponge width 20, alpha 0.015 on the left, right and bottom sides, top
    free (examples/forward/config.ini:20-25).

Every builder takes a `scale` so tests can run the same construction on a
small grid.
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np

from seispie_amd import _lib
from seispie_amd.solver.source.ricker import ricker


@dataclass
class Shot:
    src_k: np.ndarray      # [nsrc] int32
    stf: np.ndarray        # [n_inj, nsrc, nt] float32
    rec_k: np.ndarray      # [nrec] int32


@dataclass
class Grid:
    nx: int
    nz: int
    dx: float
    dz: float
    fields: list           # float32 [nx*nz] arrays in `param` order
    param: int             # 0: vp, vs, rho; 1: spin list (fd2d.py:172)
    sponge: tuple = (20, 0.015, 1, 1, 1, 0)
    shots: List[Shot] = field(default_factory=list)


@dataclass
class Problem:
    physics: int
    nt: int
    dt: float
    grids: List[Grid]
    sae: int = 50
    name: str = ""

    @property
    def nshots(self):
        return sum(len(g.shots) for g in self.grids)

    def cell_updates(self, passes=1):
        return passes * self.nt * sum(g.nx * g.nz * len(g.shots) for g in self.grids)

    def alg_bytes_per_step(self):
        """8 B per state field per shot cell + 4 B per material field per grid cell."""
        nst = {_lib.SH: 4, _lib.PSV: 7, _lib.SPIN: 12}[self.physics]
        nmat = {_lib.SH: 2, _lib.PSV: 3, _lib.SPIN: 7}[self.physics]
        return sum(8 * nst * g.nx * g.nz * len(g.shots) + 4 * nmat * g.nx * g.nz for g in self.grids)


def ricker_stf(nt, dt, f0, ang=0.0, amp=1e10, t0=None):
    """float32 (x, y, z) Ricker traces of length nt."""
    t0 = 1.2 / f0 if t0 is None else t0
    t = np.arange(nt) * dt
    sx, sy, sz = ricker(t, f0, t0, ang, amp)
    return sx.astype(np.float32), sy.astype(np.float32), sz.astype(np.float32)


def shot_stf(physics, comps):
    """[n_inj, 1, nt] from (x, y, z) traces."""
    sx, sy, sz = comps
    if physics == _lib.SH:
        return sy[None, None, :]
    return np.stack([sx[None, :], sz[None, :]])


def gindex(x, z, dx, dz, nz):
    return int(np.round(x / dx)) * nz + int(np.round(z / dz))


def smooth_random(nx, nz, seed, sigma_cells, amp):
    """Gaussian-filtered N(0,1) noise scaled to max |delta| = amp."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    a = gaussian_filter(rng.standard_normal((nx, nz)), sigma_cells, mode="wrap")
    return (amp * a / np.abs(a).max()).reshape(-1)


def checkerboard(nx, nz, cell, amp):
    i = np.arange(nx)[:, None] // cell
    j = np.arange(nz)[None, :] // cell
    return (amp * (-1.0) ** (i + j)).reshape(-1)


REC_DEPTH_CELLS = 2  # shallowest row the stencils couple to the interior


def _geometry(physics, nx, nz, dx, dz, nt, dt, nshots, nrec, f0, ang=0.0, src_z=10.0, rec_z=None):
    """Evenly spaced sources at depth src_z (one per shot), receivers at
    rec_z (default 2 dz)."""
    dx32, dz32 = np.float32(dx), np.float32(dz)
    lx = (nx - 1) * dx
    rec_x = np.linspace(0.0, lx, nrec)
    rec_z = REC_DEPTH_CELLS * dz if rec_z is None else rec_z
    rec_k = np.array([gindex(x, rec_z, dx32, dz32, nz) for x in rec_x], dtype=np.int32)
    comps = ricker_stf(nt, dt, f0, ang)
    stf = shot_stf(physics, comps)
    shots = []
    for s in range(nshots):
        x = lx * (s + 0.5) / nshots
        shots.append(Shot(np.array([gindex(x, src_z, dx32, dz32, nz)], np.int32), stf, rec_k))
    return shots


def sh_config1(scale=1.0, nt=None):
    """SH forward: 201^2, dx 10 m, dt 1 ms, nt 2000, vs 3000 with a -10%
    Gaussian anomaly (sigma 200 m), rho 2700, 10 Hz Ricker at 100 m depth,
    101 surface receivers."""
    n = max(int(round(201 * scale)), 24)
    dx = 10.0 * 201 / n
    nt = nt or 2000
    dt = 1e-3
    x = np.arange(n)[:, None] * dx
    z = np.arange(n)[None, :] * dx
    c = (n - 1) * dx / 2
    r2 = (x - c) ** 2 + (z - c) ** 2
    vs = (3000.0 * (1 - 0.10 * np.exp(-r2 / (2 * 200.0 ** 2)))).reshape(-1).astype(np.float32)
    vp = (np.sqrt(3.0) * vs).astype(np.float32)
    rho = np.full(n * n, 2700.0, np.float32)
    dx32 = np.float32(dx)
    nrec = 101 if n >= 101 else n
    rec_k = np.array([gindex(xx, REC_DEPTH_CELLS * dx, dx32, dx32, n)
                      for xx in np.linspace(0, (n - 1) * dx, nrec)], np.int32)
    stf = shot_stf(_lib.SH, ricker_stf(nt, dt, 10.0))
    g = Grid(n, n, dx, dx, [vp, vs, rho], 0, shots=[Shot(np.array([gindex(c, 100.0, dx32, dx32, n)], np.int32), stf, rec_k)])
    return Problem(_lib.SH, nt, dt, [g], sae=50, name="sh_201x201_nt2000")


def psv_config2(scale=1.0, nt=None, nshots=16, true_model=True):
    """P-SV forward+adjoint: 400^2, dx 5 m, dt 0.5 ms, nt 3000, vp/vs/rho
    3000/1730/2600 with a +-5% checkerboard of 40x40 cells (true model) or
    homogeneous (initial), 16 shots of 15 Hz, 200 receivers, sae 50."""
    n = max(int(round(400 * scale)), 24)
    nt = nt or 3000
    dt = 5e-4
    cb = checkerboard(n, n, max(int(round(40 * scale)), 2), 0.05) if true_model else 0.0
    vp = (3000.0 * (1 + cb) * np.ones(n * n)).astype(np.float32)
    vs = (1730.0 * (1 + cb) * np.ones(n * n)).astype(np.float32)
    rho = (2600.0 * (1 + cb) * np.ones(n * n)).astype(np.float32)
    shots = _geometry(_lib.PSV, n, n, 5.0, 5.0, nt, dt, nshots, min(200, n), 15.0, ang=30.0)
    return Problem(_lib.PSV, nt, dt, [Grid(n, n, 5.0, 5.0, [vp, vs, rho], 0, shots=shots)], sae=50,
                   name="psv_%dx%d_%dshots_nt%d" % (n, n, nshots, nt))


def _smooth_model(nx, nz, seeds, amp, sigma_cells, bg=(3000.0, 1730.0, 2600.0)):
    return [(b * (1 + smooth_random(nx, nz, s, sigma_cells, amp))).astype(np.float32) for b, s in zip(bg, seeds)]


def psv_config3(nshots=128, n=1000, nt=4000, sigma_cells=60):
    """Shot-batched P-SV gradient: 1000^2, dx 5 m, nt 4000, seeded smooth
    random vp/vs/rho (+-8%, seeds 1/2/3, 300 m correlation length),
    128 shots, 200 receivers."""
    dt = 5e-4
    fields = _smooth_model(n, n, (1, 2, 3), 0.08, sigma_cells)
    shots = _geometry(_lib.PSV, n, n, 5.0, 5.0, nt, dt, nshots, 200, 15.0, ang=30.0)
    return Problem(_lib.PSV, nt, dt, [Grid(n, n, 5.0, 5.0, fields, 0, shots=shots)], sae=50,
                   name="psv_%dx%d_%dshots_nt%d" % (n, n, nshots, nt))


def psv_config3_mixed(scale=1.0, nt=4000, counts=(22, 21, 21)):
    """Mixed batch: 22/21/21 shots on 300x600 / 500x500 / 800x400 grids."""
    dt = 5e-4
    grids = []
    for (nx, nz), ns, seeds in zip(((300, 600), (500, 500), (800, 400)), counts, ((1, 2, 3), (4, 5, 6), (7, 8, 9))):
        nx, nz = max(int(nx * scale), 24), max(int(nz * scale), 24)
        fields = _smooth_model(nx, nz, seeds, 0.08, max(60 * scale, 2))
        shots = _geometry(_lib.PSV, nx, nz, 5.0, 5.0, nt, dt, ns, min(200, nx), 15.0, ang=30.0)
        grids.append(Grid(nx, nz, 5.0, 5.0, fields, 0, shots=shots))
    return Problem(_lib.PSV, nt, dt, grids, sae=50, name="psv_mixed_%d_%d_%d" % counts)


def spin_config4(scale=1.0, nt=None, nshots=32):
    """Micropolar spin P-SV: 600^2, dx 5 m, nt 3000, lambda/mu from
    vp/vs/rho 3200/1850/2650, nu 1.005e9 and j 2700 (+-10% smooth random,
    seeds 4/5), mu_c 1.5e8, nu_c 3e8 (examples/spin/generate_model.py:25-29),
    32 shots of 10 Hz, 300 receivers."""
    n = max(int(round(600 * scale)), 24)
    nt = nt or 3000
    dt = 5e-4
    N = n * n
    rho = np.full(N, 2650.0, np.float32)
    vp, vs = np.float32(3200.0), np.float32(1850.0)
    lam = np.full(N, np.float32(rho[0] * (vp * vp - 2 * vs * vs)), np.float32)
    mu = np.full(N, np.float32(rho[0] * vs * vs), np.float32)
    sig = max(60 * scale, 2)
    nu = (1.005e9 * (1 + smooth_random(n, n, 4, sig, 0.10))).astype(np.float32)
    jj = (2700.0 * (1 + smooth_random(n, n, 5, sig, 0.10))).astype(np.float32)
    lam_c = np.full(N, 7.75e8, np.float32)
    mu_c = np.full(N, 1.5e8, np.float32)
    nu_c = np.full(N, 3.0e8, np.float32)
    shots = _geometry(_lib.SPIN, n, n, 5.0, 5.0, nt, dt, nshots, min(300, n), 10.0, ang=30.0)
    g = Grid(n, n, 5.0, 5.0, [lam, mu, nu, jj, lam_c, mu_c, nu_c, rho], 1, shots=shots)
    return Problem(_lib.SPIN, nt, dt, [g], sae=50, name="spin_%dx%d_%dshots_nt%d" % (n, n, nshots, nt))


def fwi_config5_model(n=500, dx=5.0, true_model=True, variant=0):
    """vp 3000 background with two +8% Gaussian blobs (sigma 60 m) over a
    curved -8% arc (variant 0, the benchmark's), or (variant 1, a second
    model for the conditioner's ablation) a -8% and a +8% blob at other
    places over a dipping +6% layer; vs = vp/sqrt(3), rho 2600."""
    x = np.arange(n)[:, None] * dx
    z = np.arange(n)[None, :] * dx
    L = (n - 1) * dx
    vp = np.full((n, n), 3000.0)
    if true_model and variant == 1:
        vp -= 240.0 * np.exp(-((x - 0.30 * L) ** 2 + (z - 0.55 * L) ** 2) / (2 * 60.0 ** 2))
        vp += 240.0 * np.exp(-((x - 0.70 * L) ** 2 + (z - 0.30 * L) ** 2) / (2 * 50.0 ** 2))
        dip = z - (0.65 * L + 0.15 * (x - 0.5 * L))
        vp += 180.0 / (1.0 + np.exp(-dip / 20.0))
    elif true_model:
        for cx, cz in ((0.35 * L, 0.40 * L), (0.65 * L, 0.40 * L)):
            vp += 240.0 * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * 60.0 ** 2))
        r = np.sqrt((x - 0.5 * L) ** 2 + (z - 0.30 * L) ** 2)
        arc = np.exp(-((r - 0.35 * L) ** 2) / (2 * 40.0 ** 2)) * (z > 0.45 * L)
        vp -= 240.0 * arc
    vp = vp.reshape(-1).astype(np.float32)
    vs = (vp / np.sqrt(3.0)).astype(np.float32)
    rho = np.full(n * n, 2600.0, np.float32)
    return [vp, vs, rho]


def fwi_config5(n=500, nt=3000, nshots=64, true_model=True, surface=False, variant=0):
    """Learned-conditioner FWI: 500^2, dx 5 m, dt 0.5 ms, nt 3000, 64 shots of
    a 15 Hz Ricker, 200 receivers.  Default: sponges on all four sides,
    sources at 125 m and receivers at 130 m depth, just below the 20-cell
    sponge: the reference-style adjoint is the gradient only away from the
    stencils' edge bands and the sponges (with sources and receivers at 10 m
    beside a free surface the finite-difference slope of the misfit and the
    reference-style kernels disagree even in sign).  surface=True: top free,
    sources at 10 m, receivers at 2 dz, for the discrete adjoint."""
    dt = 5e-4
    if surface:
        # the acquisition of the other workloads: top free, sources at 10 m,
        # receivers two cells down; needs the discrete adjoint for a gradient
        shots = _geometry(_lib.PSV, n, n, 5.0, 5.0, nt, dt, nshots, min(200, n), 15.0, ang=30.0)
        sponge = (20, 0.015, 1, 1, 1, 0)
    else:
        shots = _geometry(_lib.PSV, n, n, 5.0, 5.0, nt, dt, nshots, min(200, n), 15.0, ang=30.0, src_z=125.0,
                          rec_z=130.0)
        sponge = (20, 0.015, 1, 1, 1, 1)
    g = Grid(n, n, 5.0, 5.0, fwi_config5_model(n, 5.0, true_model, variant), 0, sponge=sponge, shots=shots)
    return Problem(_lib.PSV, nt, dt, [g], sae=50,
                   name="fwi_%dx%d_%dshots_nt%d%s%s" % (n, n, nshots, nt, "_surface" if surface else "",
                                                       "_model%d" % variant if variant else ""))


def load_problem(ctx, prob):
    """Create the grids and shots of a Problem in a library Context."""
    ctx.set_physics(prob.physics, prob.nt, prob.dt)
    ids = []
    for g in prob.grids:
        gid = ctx.add_grid(g.nx, g.nz, g.dx, g.dz, g.fields, g.param)
        ctx.set_sponge(gid, *g.sponge)
        for s in g.shots:
            ids.append(ctx.add_shot(gid, s.src_k, s.stf, s.rec_k))
    return ids
