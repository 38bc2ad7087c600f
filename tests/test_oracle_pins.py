"""Pins of the CPU oracle (oracle/oracle.c) by closed forms, invariants and
finite differences.  CPU only: these run without a GPU."""
import os

import numpy as np
import pytest

from oracle import oracle as O

REF = "/root/reference/examples"


# ------------------------------------------------------------ operators
@pytest.mark.parametrize("which,axis,shift", [(0, 0, -0.5), (1, 0, 0.5), (2, 1, -0.5), (3, 1, 0.5)])
def test_staggered_difference_exact_for_cubics(which, axis, shift):
    """D-/D+ (solver.py:22-51) are exact at the half-grid point for cubics,
    and zero outside their valid bands."""
    nx, nz, h = 14, 12, 2.5
    i = np.arange(nx)[:, None] * np.ones((1, nz))
    j = np.ones((nx, 1)) * np.arange(nz)[None, :]
    c = i if axis == 0 else j
    p = lambda x: 0.3 + 1.7 * x - 0.4 * x ** 2 + 0.05 * x ** 3
    dp = lambda x: 1.7 - 0.8 * x + 0.15 * x ** 2
    v = p(c * h)
    out = O.diff(v, nx, nz, h, h, which)
    n = nx if axis == 0 else nz
    lo = 2 if which in (0, 2) else 1
    valid = (c >= lo) & (c < n - 2)
    np.testing.assert_allclose(out[valid], dp((c[valid] + shift) * h), rtol=1e-12, atol=1e-12)
    assert np.all(out[~valid] == 0.0)


def test_dminus_is_minus_dplus_transpose_in_the_interior():
    """Interior rows of D- equal rows of -(D+)^T; the edge bands differ
    (which is why the reference-style adjoint is exact only away from edges)."""
    nx, nz = 12, 10
    n = nx * nz
    M = {}
    for w in range(4):
        A = np.zeros((n, n))
        for k in range(n):
            e = np.zeros(n)
            e[k] = 1.0
            A[:, k] = O.diff(e, nx, nz, 1.0, 1.0, w).reshape(-1)
        M[w] = A
    i = np.arange(n) // nz
    j = np.arange(n) % nz
    inner_x = (i >= 3) & (i < nx - 3)
    inner_z = (j >= 3) & (j < nz - 3)
    np.testing.assert_allclose(M[0][inner_x], -M[1].T[inner_x], atol=1e-14)
    np.testing.assert_allclose(M[2][inner_z], -M[3].T[inner_z], atol=1e-14)
    assert not np.allclose(M[0], -M[1].T)


# ------------------------------------------------------------- host pieces
def test_ricker_closed_form():
    f0, t0, amp = 10.0, 0.12, 1e10
    x, y, z = O.ricker(t0, f0, t0, 30.0, amp)
    assert y == -amp
    assert np.isclose(x, -amp * 0.5) and np.isclose(z, amp * np.cos(np.deg2rad(30.0)))
    tz = t0 + 1.0 / (np.pi * f0 * np.sqrt(2.0))
    assert abs(O.ricker(tz, f0, t0, 0.0, amp)[1]) < 1e-6 * amp
    assert np.isclose(O.ricker(t0 - 0.03, f0, t0, 0, amp)[1], O.ricker(t0 + 0.03, f0, t0, 0, amp)[1])


def test_waveform_taper_adjoint_source_and_norm():
    """waveform.py:5-34: res = (d - u) w, adjoint source reversed 2 w^2 (d - u)."""
    nt, nrec, dt = 101, 3, 0.01
    rng = np.random.default_rng(0)
    d = rng.standard_normal((nrec, nt)).astype(np.float32)
    u = rng.standard_normal((nrec, nt)).astype(np.float32)
    norm, ad, res = O.waveform(d, u, nt, dt, "f64")
    t = np.arange(nt) * dt
    te = (nt - 1) * dt
    tw = te / 10
    w = np.ones(nt)
    w[t <= tw] = 0.5 + 0.5 * np.cos(np.pi * (tw - t[t <= tw]) / tw)
    w[t >= te - tw] = 0.5 + 0.5 * np.cos(np.pi * (te - tw - t[t >= te - tw]) / tw)
    assert w[0] == pytest.approx(0.0, abs=1e-15) and w[-1] == pytest.approx(0.0, abs=1e-15)
    np.testing.assert_allclose(res, (d.astype(np.float64) - u) * w, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(ad[:, ::-1], 2 * w ** 2 * (d.astype(np.float64) - u), rtol=1e-12, atol=1e-12)
    assert norm == pytest.approx(np.linalg.norm(res), rel=1e-12)


def test_sponge_closed_form_and_separable():
    nx, nz, width, alpha = 50, 40, 20, 0.015
    b = O.set_bound(nx, nz, width, alpha, 1, 1, 1, 0).reshape(nx, nz)
    fx = np.ones(nx)
    for i in range(nx):
        if i + 1 < width:
            fx[i] *= np.exp(-(alpha * (width - i - 1)) ** 2)
        if i > nx - width:
            fx[i] *= np.exp(-(alpha * (width + i - nx)) ** 2)
    fz = np.array([np.exp(-(alpha * (width + j - nz)) ** 2) if j > nz - width else 1.0 for j in range(nz)])
    np.testing.assert_allclose(b, np.outer(fx, fz), rtol=2e-7)
    assert b[25, 0] == 1.0 and b[nx // 2, :10].min() == 1.0  # top free


def test_vps2lm_closed_form():
    vp = np.array([3000.0, 1000.0, 5500.0], np.float32)
    vs = np.array([1730.0, 1500.0, 3500.0], np.float32)
    rho = np.array([2600.0, 2000.0, 2700.0], np.float32)
    lam, mu = O.vps2lm(vp, vs, rho)
    np.testing.assert_allclose(mu, rho.astype(np.float64) * vs ** 2.0, rtol=2e-7)
    np.testing.assert_allclose(lam[[0, 2]], (rho * (vp.astype(np.float64) ** 2 - 2 * vs.astype(np.float64) ** 2))[[0, 2]], rtol=2e-7)
    assert lam[1] == 0.0  # vp <= vs -> 0 (fd2d.py:28-31)


def test_smooth_matches_brute_force_and_keeps_constants():
    nx, nz, s = 9, 7, 2
    rng = np.random.default_rng(1)
    a = rng.standard_normal(nx * nz).astype(np.float32)
    g = lambda d: np.exp(-d * d / (2.0 * s * s)) / (np.sqrt(2 * np.pi) * s)
    A = a.reshape(nx, nz).astype(np.float64)
    ref = np.zeros((nx, nz))
    for i in range(nx):
        for j in range(nz):
            w = np.outer(g(i - np.arange(nx)), g(j - np.arange(nz)))
            ref[i, j] = (w * A).sum() / w.sum()
    out = O.smooth(a, nx, nz, s, "f64").reshape(nx, nz)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    c = np.full(nx * nz, 3.25, np.float32)
    np.testing.assert_allclose(O.smooth(c, nx, nz, s, "f32"), c, rtol=1e-6)


# ------------------------------------------------------------ simulations
def _homog(physics, nx, nz):
    n = nx * nz
    m = {"rho": np.full(n, 2600.0, np.float32), "lam": np.full(n, 7.8e9, np.float32),
         "mu": np.full(n, 7.8e9, np.float32)}
    if physics == O.SPIN:
        m.update(nu=np.full(n, 1.0e9, np.float32), j=np.full(n, 2700.0, np.float32),
                 mu_c=np.full(n, 1.5e8, np.float32), nu_c=np.full(n, 3e8, np.float32))
    return m


def _stf(nt, dt, nsrc=1, f0=40.0, amp=1e10):
    t = np.arange(nt) * dt
    s = np.concatenate([O.ricker(t, f0, 1.2 / f0, 30.0, amp * (1 + 0.25 * q))[1] for q in range(nsrc)])
    return s.astype(np.float32)


@pytest.mark.parametrize("physics", [O.SH, O.PSV, O.SPIN])
def test_linearity_superposition_and_zero_input(physics):
    nx, nz, nt, dt = 48, 44, 150, 5e-4
    bound = O.set_bound(nx, nz, 10, 0.015, 1, 1, 1, 0)
    src = np.array([20 * nz + 15, 28 * nz + 22], np.int32)
    rec = np.array([10 * nz + 5, 35 * nz + 30, 24 * nz + 2], np.int32)
    s = _stf(nt, dt, 2)
    stf = [0.5 * s, s, -0.7 * s]
    z = np.zeros_like(s)
    run = lambda st, isrc: O.Sim(physics, nx, nz, 5.0, 5.0, dt, nt, _homog(physics, nx, nz), bound, src, rec, isrc).forward(st, precision="f64")[0]
    both = run(stf, -1)
    one, two = run(stf, 0), run(stf, 1)
    dbl = run([2 * a for a in stf], -1)
    for c in both:
        np.testing.assert_allclose(both[c], one[c] + two[c], rtol=1e-10, atol=1e-12 * np.abs(both[c]).max())
        np.testing.assert_allclose(dbl[c], 2 * both[c], rtol=1e-12, atol=1e-14 * np.abs(both[c]).max())
        assert np.abs(both[c]).max() > 0
    zero = run([z, z, z], -1)
    assert all(np.all(zero[c] == 0) for c in zero)


@pytest.mark.parametrize("physics", [O.SH, O.PSV])
def test_reciprocity_in_the_interior(physics):
    """Source at A recorded at B equals source at B recorded at A (both
    components for P-SV) when no wave reaches the edge bands."""
    nx = nz = 160
    nt, dt = 200, 5e-4
    A, B = 70 * nz + 75, 86 * nz + 83
    t = np.arange(nt) * dt
    s = np.exp(-((t - 0.02) / 0.004) ** 2)
    z = np.zeros(nt)
    bound = np.ones(nx * nz, np.float32)
    m = _homog(physics, nx, nz)
    m["rho"] = (2600 * (1 + 0.1 * np.sin(np.arange(nx * nz) / 37.0))).astype(np.float32)

    def run(src, comp):
        sim = O.Sim(physics, nx, nz, 5.0, 5.0, dt, nt, m, bound, [src], [A, B], 0)
        st = [s if comp == "x" else z, s if comp == "y" else z, s if comp == "z" else z]
        return sim.forward(st, precision="f64")[0]

    comps = ["y"] if physics == O.SH else ["x", "z"]
    for c1 in comps:
        for c2 in comps:
            ab = run(A, c1)[c2][1]
            ba = run(B, c2)[c1][0]
            assert np.linalg.norm(ab - ba) <= 1e-12 * np.linalg.norm(ab)


def test_homogeneous_sh_matches_line_source_greens_function():
    """Point force in a homogeneous SH medium against the 2-D Green's function
    u(r,t) = dx dz/(2 pi rho c^2) int_0^inf s(t - (r/c) cosh th) d th (grid
    dispersion small at >= 20 cells per wavelength)."""
    nx = nz = 260
    h, nt, dt, f0 = 5.0, 700, 5e-4, 12.0
    rho, vs = 2600.0, 1730.0
    n = nx * nz
    m = {"rho": np.full(n, rho, np.float32), "lam": np.zeros(n, np.float32), "mu": np.full(n, np.float32(rho * vs * vs), np.float32)}
    src = 130 * nz + 130
    rec = np.array([(130 + d) * nz + 130 for d in (20, 40)], np.int32)
    t0 = 1.2 / f0
    t = np.arange(nt) * dt
    s = O.ricker(t, f0, t0, 0.0, 1e10)[1]
    sim = O.Sim(O.SH, nx, nz, h, h, dt, nt, m, np.ones(n, np.float32), [src], rec, 0)
    u = sim.forward([np.zeros(nt), s, np.zeros(nt)], precision="f64")[0]["y"]
    c = float(np.sqrt(np.float32(rho * vs * vs) / np.float32(rho)))
    th = np.linspace(0, 12, 24001)
    for r, d in enumerate((20, 40)):
        R = d * h
        tt = (np.arange(nt) + 1) * dt - 0.5 * dt
        arg = tt[:, None] - (R / c) * np.cosh(th)[None, :]
        g = O.ricker(arg, f0, t0, 0.0, 1e10)[1] * (arg > -1e9)
        ana = h * h / (2 * np.pi * rho * c * c) * np.trapezoid(g, th, axis=1)
        err = np.linalg.norm(u[r] - ana) / np.linalg.norm(ana)
        assert err < 0.03, err


def _discrete_energy(physics, snap, m, nx, nz, h, dt):
    """Energy of the staggered leapfrog that the reference's step sequence
    (solver.py:205-234) conserves when D- = -(D+)^T on the wavefield's
    support and nothing is injected: with v(n+1) the velocities after step n
    and the stresses written through the displacements, s(n) = C B u(n),
        E(n) = 1/2 sum M v(n+1)^2 + 1/2 sum (B u(n))^T C (B u(n+1)),
    u(n+1) = u(n) + dt v(n+1).  B: SH (D+x u, D+z u) with C = mu; P-SV
    strains (D+x ux, D+z uz, D+z ux + D+x uz) with the isotropic C; spin
    adds the rotation mismatch e = -phi - (D+x uz - D+z ux)/2 weighted 2 nu
    (syy_c = 2 nu e, entering the velocities as -2 B^T) and the curvatures
    D+x phi, D+z phi weighted mu_c + nu_c, with j as phi's inertia.
    snap: one oracle snapshot (u before step n, v after it), float64."""
    n = nx * nz
    D = lambda f, w: O.diff(f, nx, nz, h, h, w).reshape(-1)
    f = {k: np.asarray(v, np.float64) for k, v in m.items()}
    if physics == O.SH:
        u, v = snap[0], snap[1]
        u1 = u + dt * v
        pot = f["mu"] * (D(u, 1) * D(u1, 1) + D(u, 3) * D(u1, 3))
        return 0.5 * np.sum(f["rho"] * v * v) + 0.5 * np.sum(pot)
    ncomp = 2 if physics == O.PSV else 3
    u, v = snap[:ncomp], snap[ncomp:]
    u1 = u + dt * v

    def strain(ux, uz):
        return D(ux, 1), D(uz, 3), D(ux, 3) + D(uz, 1)

    exx, ezz, g = strain(u[0], u[1])
    exx1, ezz1, g1 = strain(u1[0], u1[1])
    l2m = f["lam"] + 2 * f["mu"]
    pot = l2m * (exx * exx1 + ezz * ezz1) + f["lam"] * (exx * ezz1 + ezz * exx1) + f["mu"] * g * g1
    kin = f["rho"] * (v[0] * v[0] + v[1] * v[1])
    if physics == O.SPIN:
        e = -u[2] - 0.5 * (D(u[1], 1) - D(u[0], 3))
        e1 = -u1[2] - 0.5 * (D(u1[1], 1) - D(u1[0], 3))
        c = f["mu_c"] + f["nu_c"]
        pot = pot + 4 * f["nu"] * e * e1 + c * (D(u[2], 1) * D(u1[2], 1) + D(u[2], 3) * D(u1[2], 3))
        kin = kin + f["j"] * v[2] * v[2]
    return 0.5 * np.sum(kin) + 0.5 * np.sum(pot)


@pytest.mark.parametrize("physics", [O.SH, O.PSV, O.SPIN])
def test_discrete_energy_is_conserved_without_sponge(physics):
    """Sponge off, a heterogeneous model, a source that stops at step 150:
    from then on the discrete energy stays constant to float64 rounding
    while the wavefield has not reached the edge bands (where D- differs
    from -(D+)^T).  A wrong sign, a dropped or doubled term, or a wrong
    coefficient in any stress or velocity update breaks the constancy."""
    from scipy.ndimage import gaussian_filter
    nx = nz = 300
    h, dt, nt, sae, K = 5.0, 5e-4, 360, 20, 150
    n = nx * nz
    rng = np.random.default_rng(5)
    sr = lambda: (lambda a: (a / np.abs(a).max()).reshape(-1))(gaussian_filter(rng.standard_normal((nx, nz)), 6))
    m = _homog(physics, nx, nz)
    for k in m:
        m[k] = (m[k] * (1 + 0.1 * sr())).astype(np.float32)
    t = np.arange(nt) * dt
    s = O.ricker(t, 30.0, 1.2 / 30.0, 0.0, 1e10)[1]
    s[K:] = 0.0
    src = [150 * nz + 150]
    sim = O.Sim(physics, nx, nz, h, h, dt, nt, m, np.ones(n, np.float32), src, [10 * nz + 10], 0)
    stf = [0.6 * s, s, -0.8 * s]
    _, snaps = sim.forward(stf, sae=sae, precision="f64")
    nsa = nt // sae
    E = {}
    for isa in range(nsa):
        it = (nsa - isa) * sae - 1          # the step this snapshot brackets
        if it >= K + sae:
            E[it] = _discrete_energy(physics, snaps[isa], m, nx, nz, h, dt)
    vals = np.array([E[k] for k in sorted(E)])
    assert len(vals) >= 8 and vals[0] > 0
    drift = np.abs(vals - vals[0]).max() / vals[0]
    # the last snapshot's wavefield is negligible at the edges: the check is in the interior
    last = np.abs(snaps[0]).reshape(snaps.shape[1], nx, nz)
    edge = max(last[:, :4].max(), last[:, -4:].max(), last[:, :, :4].max(), last[:, :, -4:].max())
    assert edge <= 1e-12 * last.max()
    assert drift <= 1e-11, drift
    # the same energy before the source stops is not constant (the check has teeth)
    early = _discrete_energy(physics, snaps[nsa - 3], m, nx, nz, h, dt)   # step 59, source active
    assert abs(early - vals[0]) > 1e-3 * vals[0]


def _xcorr_lag(a, b, dt):
    """Delay of b behind a (seconds) at the cross-correlation peak, with
    parabolic refinement."""
    c = np.correlate(b, a, mode="full")
    k = int(np.argmax(c))
    y0, y1, y2 = c[k - 1], c[k], c[k + 1]
    frac = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    return (k - (a.size - 1) + frac) * dt


def test_homogeneous_psv_arrival_times_follow_vp_and_vs():
    """A vertical point force in a homogeneous P-SV medium with vp/vs = 2
    (lambda = 2 mu): along the force's axis uz carries the P wave, across
    it the S wave.  Their delays between two receiver distances (30 and 80
    cells) give vp and vs within 1%; lambda and mu in swapped roles, or a
    wrong factor in the stress update, would move them by 10% or more."""
    nx = nz = 300
    h, dt, nt, f0 = 5.0, 5e-4, 720, 20.0
    vp0, vs0, rho0 = 3000.0, 1500.0, 2600.0
    n = nx * nz
    lam, mu = O.vps2lm(np.full(n, vp0, np.float32), np.full(n, vs0, np.float32), np.full(n, rho0, np.float32))
    m = {"rho": np.full(n, rho0, np.float32), "lam": lam, "mu": mu}
    c0 = 150
    d1, d2 = 30, 80
    rec = [c0 * nz + c0 + d1, c0 * nz + c0 + d2,          # along z (the force's axis): P
           (c0 + d1) * nz + c0, (c0 + d2) * nz + c0]      # along x: S
    t = np.arange(nt) * dt
    s = O.ricker(t, f0, 1.2 / f0, 0.0, 1e10)[1]
    z = np.zeros(nt)
    sim = O.Sim(O.PSV, nx, nz, h, h, dt, nt, m, np.ones(n, np.float32), [c0 * nz + c0], rec, 0)
    uz = sim.forward([z, z, s], precision="f64")[0]["z"]
    vp = (d2 - d1) * h / _xcorr_lag(uz[0], uz[1], dt)
    vs = (d2 - d1) * h / _xcorr_lag(uz[2], uz[3], dt)
    assert vp == pytest.approx(vp0, rel=0.01), vp
    assert vs == pytest.approx(vs0, rel=0.01), vs


@pytest.mark.parametrize("physics", [O.SH, O.PSV, O.SPIN])
def test_float32_storage_stays_within_tolerance_of_the_fp64_replay(physics):
    nx, nz, nt, dt = 56, 52, 400, 5e-4
    bound = O.set_bound(nx, nz, 20, 0.015, 1, 1, 1, 0)
    m = _homog(physics, nx, nz)
    rng = np.random.default_rng(3)
    m["mu"] = (m["mu"] * (1 + 0.05 * rng.random(nx * nz))).astype(np.float32)
    sim = O.Sim(physics, nx, nz, 5.0, 5.0, dt, nt, m, bound, [26 * nz + 10], [10 * nz, 30 * nz + 1, 40 * nz + 2], 0)
    s = _stf(nt, dt)
    a = sim.forward([0.5 * s, s, -s], precision="f32")[0]
    b = sim.forward([0.5 * s, s, -s], precision="f64")[0]
    for c in a:
        assert np.linalg.norm(a[c] - b[c]) <= 1e-5 * np.linalg.norm(b[c])


# ------------------------------------------------------ adjoint kernels
def _fd_problem(physics, nx=100, nz=90, nt=400, dt=5e-4, h=5.0):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(11)
    sr = lambda s=4: (lambda a: (a / np.abs(a).max()).reshape(-1))(gaussian_filter(rng.standard_normal((nx, nz)), s))
    vp = 3000 * (1 + 0.05 * sr())
    vs = 1730 * (1 + 0.05 * sr())
    rho = 2600 * (1 + 0.05 * sr())
    lam, mu = O.vps2lm(vp.astype("f4"), vs.astype("f4"), rho.astype("f4"))
    m = {"rho": rho.astype("f4"), "lam": lam, "mu": mu}
    if physics == O.SPIN:
        m.update(nu=(1.005e9 * (1 + 0.1 * sr())).astype("f4"), j=(2700 * (1 + 0.1 * sr())).astype("f4"),
                 mu_c=np.full(nx * nz, 1.5e8, "f4"), nu_c=np.full(nx * nz, 3e8, "f4"))
    bound = O.set_bound(nx, nz, 20, 0.03, 1, 1, 1, 1)
    src = np.array([[nx // 2 * h, 55 * h, 0, 40.0, 0.03, 30.0, 1e10]])
    rec = np.array([[i * h, 30 * h] for i in range(25, nx - 25, 4)])
    sid = O.grid_index(src, np.float32(h), np.float32(h), nz)
    rid = O.grid_index(rec, np.float32(h), np.float32(h), nz)
    stf = O.make_stf(src, nt, dt)
    mt = {k: v.copy() for k, v in m.items()}
    mt["mu"] = (mt["mu"] * (1 + 0.03 * sr())).astype("f4")
    mt["rho"] = (mt["rho"] * (1 - 0.03 * sr())).astype("f4")
    data, _ = O.Sim(physics, nx, nz, h, h, dt, nt, mt, bound, sid, rid, 0).forward(stf, precision="f64")
    return m, bound, sid, rid, stf, data, sr


@pytest.mark.parametrize("physics,keys", [(O.SH, ["mu", "rho"]), (O.PSV, ["lam", "mu", "rho"]),
                                          (O.SPIN, ["lam", "mu", "rho", "nu", "j", "muc"])])
def test_adjoint_kernels_match_finite_differences(physics, keys):
    """k ~= dt * dchi/dm with chi = sum_t,r w^2 (u - d)^2 (sae = 1, edges
    absorbed): the reference's SH k_mu and this project's P-SV / spin
    kernel definitions."""
    nx, nz, nt, dt, h = 100, 90, 400, 5e-4, 5.0
    m, bound, sid, rid, stf, data, sr = _fd_problem(physics, nx, nz, nt, dt, h)
    comps = ["y"] if physics == O.SH else ["x", "z"]
    sim = O.Sim(physics, nx, nz, h, h, dt, nt, m, bound, sid, rid, 0)
    _, K = O.run_kernel(sim, stf, [data], 1, adj=True, precision="f64")

    def chi(mm):
        obs, _ = O.Sim(physics, nx, nz, h, h, dt, nt, mm, bound, sid, rid, 0).forward(stf, precision="f64")
        return sum(O.waveform(data[c], obs[c], nt, dt, "f64")[0] ** 2 for c in comps)

    for key in keys:
        d = sr(3)
        pk = {"muc": "mu_c"}.get(key, key)
        eps = 1e-3
        vals = []
        for sgn in (1, -1):
            mm = dict(m)
            mm[pk] = m[pk].astype("f8") * (1 + sgn * eps * d)
            if key == "muc":
                mm["nu_c"] = m["nu_c"].astype("f8") * (1 + sgn * eps * d)
            vals.append(chi(mm))
        fd = dt * (vals[0] - vals[1]) / (2 * eps)
        w = m[pk].astype("f8") + (m["nu_c"] if key == "muc" else 0)
        adj = np.sum(K[key].astype("f8") * w * d)
        assert adj == pytest.approx(fd, rel=2e-3), (key, adj, fd)


@pytest.mark.parametrize("physics", [O.SH, O.PSV, O.SPIN])
def test_discrete_adjoint_matches_finite_differences_beside_a_free_surface(physics):
    """The discrete adjoint (oracle.orc_adjoint_exact: the transpose of the
    time stepping, edge bands and sponge included) is the gradient of a loss
    linear in the records, L = sum W . records, with the source and the
    receivers two cells below a free top and sponges on the other sides:
    float64 replay, central differences along a smooth perturbation of the
    whole model, measured along the perturbation the float32 model really
    takes.  The reference-style kernels (forward operator reused) miss it
    there, for P-SV even in sign."""
    from scipy.ndimage import gaussian_filter
    nx, nz, h, dt, nt = 60, 50, 5.0, 5e-4, 300
    rng = np.random.default_rng(1)
    sr = lambda s=3: (lambda a: (a / np.abs(a).max()).reshape(-1))(gaussian_filter(rng.standard_normal((nx, nz)), s))
    vp = (3000 * (1 + 0.05 * sr())).astype("f4")
    vs = (1730 * (1 + 0.05 * sr())).astype("f4")
    rho = (2600 * (1 + 0.05 * sr())).astype("f4")
    lam, mu = O.vps2lm(vp, vs, rho)
    m = {"rho": rho, "lam": lam, "mu": mu}
    if physics == O.SPIN:
        m.update(nu=(1.0e9 * (1 + 0.1 * sr())).astype("f4"), j=(2700 * (1 + 0.1 * sr())).astype("f4"),
                 mu_c=(1.5e8 * (1 + 0.1 * sr())).astype("f4"), nu_c=(3e8 * (1 + 0.1 * sr())).astype("f4"))
    bound = O.set_bound(nx, nz, 10, 0.03, 1, 1, 1, 0)
    src = np.array([[30 * h, 2 * h, 0, 40.0, 0.03, 30.0, 1e10]])
    rec = np.array([[i * h, 2 * h] for i in range(5, nx - 5, 3)])
    sid = O.grid_index(src, np.float32(h), np.float32(h), nz)
    rid = O.grid_index(rec, np.float32(h), np.float32(h), nz)
    stf = O.make_stf(src, nt, dt)
    comps = {O.SH: ["y"], O.PSV: ["x", "z"], O.SPIN: ["x", "z", "y"]}[physics]
    W = {c: rng.standard_normal((rid.size, nt)) for c in comps}

    def loss(mm):
        obs, _ = O.Sim(physics, nx, nz, h, h, dt, nt, mm, bound, sid, rid, 0).forward(stf, precision="f64")
        return sum(float(np.sum(W[c] * obs[c])) for c in comps)

    sim = O.Sim(physics, nx, nz, h, h, dt, nt, m, bound, sid, rid, 0)
    _, snaps = sim.forward(stf, sae=1, precision="f64")
    K = sim.adjoint_exact(W, snaps, "f64")
    # the reference-style adjoint injects x and z only (the misfit's components)
    Kr = sim.adjoint({c: -W[c][:, ::-1] for c in comps if c != "y" or physics == O.SH}, 1, snaps, None, "f64")
    worst_ref = 0.0
    keys = {O.SH: ["mu", "rho"], O.PSV: ["lam", "mu", "rho"], O.SPIN: ["lam", "mu", "rho", "nu", "j", "muc"]}[physics]
    for key in keys:
        mk = {"muc": "mu_c"}.get(key, key)   # mu_c and nu_c enter as their sum; perturb mu_c
        base = m[mk].astype(np.float64)
        d = sr(2)
        p32 = (base * (1 + 1e-3 * d)).astype(np.float32)
        q32 = (base * (1 - 1e-3 * d)).astype(np.float32)
        fd = loss(dict(m, **{mk: p32})) - loss(dict(m, **{mk: q32}))
        step = p32.astype(np.float64) - q32.astype(np.float64)
        adj = float(np.sum(K[key].astype(np.float64) / dt * step))
        assert adj == pytest.approx(fd, rel=2e-5), (key, adj, fd)
        if physics != O.SPIN:
            ref = float(np.sum(Kr[key].astype(np.float64) / dt * step))
            worst_ref = max(worst_ref, abs(ref - fd) / abs(fd))
    assert physics == O.SPIN or worst_ref > 2e-3


# ------------------------------------------------ reference-held vectors
@pytest.mark.skipif(not os.path.exists(REF), reason="reference examples not mounted")
def test_reference_trace_is_from_another_configuration():
    """examples/inverse-checker/traces/vy_000000.npy (raw float32 [132, 4800]
    SH records) does not follow from the current examples/forward inputs:
    the oracle's run of source 0 differs by ~100%, and even the best
    least-squares source wavelet leaves a ~22% residual, so the file cannot
    pin the forward pass (it was made with another model or geometry)."""
    def rd(d, n):
        with open("%s/%s/proc000000_%s.bin" % (REF, d, n), "rb") as f:
            npt = np.fromfile(f, dtype="int32", count=1)
            f.seek(4)
            return npt[0], np.fromfile(f, dtype="float32")
    npt, x = rd("forward/model_true", "x")
    _, z = rd("forward/model_true", "z")
    vp, vs, rho = (rd("forward/model_true", k)[1] for k in ("vp", "vs", "rho"))
    nx, nz, dx, dz = O.grid_shape(x, z, npt)
    assert (nx, nz, float(dx)) == (201, 201, 2400.0)
    lam, mu = O.vps2lm(vp, vs, rho)
    src = np.loadtxt(REF + "/forward/sources.dat", ndmin=2)
    rec = np.loadtxt(REF + "/forward/stations.dat", ndmin=2)
    nt, dt = 4800, 0.06
    sim = O.Sim(O.SH, nx, nz, dx, dz, dt, nt, {"rho": rho, "lam": lam, "mu": mu},
                O.set_bound(nx, nz, 20, 0.015, 1, 1, 1, 0), O.grid_index(src, dx, dz, nz),
                O.grid_index(rec, dx, dz, nz), isrc=0)
    o = sim.forward(O.make_stf(src, nt, dt))[0]["y"]
    ref = np.fromfile(REF + "/inverse-checker/traces/vy_000000.npy", dtype="float32").reshape(132, nt)
    assert np.linalg.norm(o - ref) / np.linalg.norm(ref) > 0.5
