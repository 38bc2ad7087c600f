"""Test helpers: run a seispie_amd.configs.Problem through the oracle."""
import numpy as np

from oracle import oracle as O
from seispie_amd import _lib

COMP_NAMES = {_lib.SH: ["y"], _lib.PSV: ["x", "z"], _lib.SPIN: ["x", "z", "y"]}
MIS_NAMES = {_lib.SH: ["y"], _lib.PSV: ["x", "z"], _lib.SPIN: ["x", "z"]}


def oracle_model(g):
    if g.param == 0:
        vp, vs, rho = g.fields
        lam, mu = O.vps2lm(vp, vs, rho)
        return {"rho": rho, "lam": lam, "mu": mu}
    lam, mu, nu, jj, lam_c, mu_c, nu_c, rho = g.fields
    return {"rho": rho, "lam": lam, "mu": mu, "nu": nu, "j": jj, "mu_c": mu_c, "nu_c": nu_c}


def oracle_sim(prob, g, shot):
    bound = O.set_bound(g.nx, g.nz, *g.sponge)
    return O.Sim(prob.physics, g.nx, g.nz, g.dx, g.dz, prob.dt, prob.nt, oracle_model(g), bound,
                 shot.src_k, shot.rec_k, isrc=-1)


def oracle_stf(prob, shot):
    z = np.zeros(shot.src_k.size * prob.nt, np.float32)
    if prob.physics == _lib.SH:
        return [z, shot.stf[0].reshape(-1), z]
    return [shot.stf[0].reshape(-1), z, shot.stf[1].reshape(-1)]


def oracle_records(prob, g, shot, precision="f32", sae=0):
    sim = oracle_sim(prob, g, shot)
    obs, snaps = sim.forward(oracle_stf(prob, shot), sae=sae, precision=precision)
    return np.stack([obs[c] for c in COMP_NAMES[prob.physics]]), snaps, sim


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def oracle_gradient(prob, data, precision="f32"):
    """Per grid kernels and per shot misfit norms from the oracle, shots in
    order.  data: list over shots of [ncomp, nrec, nt]."""
    kern, norms, s = [], [], 0
    for g in prob.grids:
        K = None
        for shot in g.shots:
            rec, snaps, sim = oracle_records(prob, g, shot, precision, sae=prob.sae)
            ad, nn = {}, []
            for c, name in enumerate(MIS_NAMES[prob.physics]):
                m, a, _ = O.waveform(data[s][c], rec[c], prob.nt, prob.dt, precision)
                ad[name] = a
                nn.append(m)
            norms.append(nn)
            K = sim.adjoint(ad, prob.sae, snaps, K, precision)
            s += 1
        kern.append(K)
    return kern, np.array(norms)
