"""Optimizers and the step-conditioner hook on synthetic problems (CPU):
convergence, the conditioned step's gain, and the fallbacks the reference
lacks or has (non-descent direction, non-finite step, line-search failure,
concave bracket)."""
import numpy as np
import pytest

from seispie_amd.optimize.cg import cg
from seispie_amd.optimize.lbfgs import lbfgs
from seispie_amd.optimize.line_search.bracket import LineSearchFitError, bracket

CONFIG = {"niter": "12", "lbfgs_mem": "5", "ls_step": "10", "ls_step_max": "0.5", "ls_step_init": "0.05",
          "ls_thresh": "1.2"}


class Quadratic:
    """f(m) = 1/2 (m - m*)' A (m - m*) with a diagonal A; the solver plugin
    surface the optimizers call."""

    def __init__(self, n=200, cond=1e3, seed=0, sign=1.0):
        rng = np.random.default_rng(seed)
        self.nx, self.nz = n, 1
        self.a = np.logspace(0, np.log10(cond), n)[rng.permutation(n)]
        self.mstar = 3.0 + rng.standard_normal(n)
        self.m = np.full(n, 3.0)
        self.sign = sign
        self.ngrad = self.nmis = 0

        class F:
            def __init__(s, o):
                s.o = o

            def copy_to_host(s, dst, stream=None):
                dst[...] = s.o.m

        class S:
            def synchronize(s):
                pass

        self.mu, self.stream = F(self), S()

    def f(self, m):
        d = m - self.mstar
        return 0.5 * float(np.dot(d, self.a * d))

    def update_model(self, m):
        self.m = np.array(m, dtype=float)

    def compute_misfit(self):
        self.nmis += 1
        return self.f(self.m)

    def compute_gradient(self):
        self.ngrad += 1
        return self.sign * self.a * (self.m - self.mstar), self.f(self.m), self.m.copy()

    def export_field(self, *a):
        pass


def _run(method, solver, precond=None, niter=12):
    opt = {"lbfgs": lbfgs, "cg": cg}[method]()
    opt.config = dict(CONFIG, niter=str(niter))
    opt.mpi = None

    class W:
        pass

    w = W()
    w.solver = solver
    opt.setup(w)
    opt.precond = precond
    opt.quiet = True
    opt.bracket.head = 0
    opt.run()
    return opt


@pytest.mark.parametrize("method", ["lbfgs", "cg"])
def test_converges_on_a_quadratic(method):
    q = Quadratic()
    f0 = q.f(q.m)
    opt = _run(method, q, niter=25)
    f = [h[1] for h in opt.history]
    assert f[-1] < 1e-2 * f0
    assert all(b <= a * (1 + 1e-12) for a, b in zip(f, f[1:])), f


@pytest.mark.parametrize("method", ["lbfgs", "cg"])
def test_diagonal_conditioner_beats_the_plain_step(method):
    """P = A^-1 (the exact diagonal inverse Hessian, what the learned
    conditioner approximates) reaches a far lower misfit in the same
    number of iterations."""
    plain = _run(method, Quadratic(), niter=8).history[-1][1]
    q = Quadratic()
    cond = _run(method, q, precond=lambda v: v / q.a, niter=8).history[-1][1]
    assert cond < 1e-3 * plain, (cond, plain)


class Refreshing:
    def __init__(self, fn):
        self.fn = fn
        self.refreshed = 0

    def refresh(self, g):
        self.refreshed += 1

    def __call__(self, v):
        return self.fn(v)


@pytest.mark.parametrize("method", ["lbfgs", "cg"])
@pytest.mark.parametrize("bad", ["ascent", "nan"])
def test_bad_conditioner_falls_back_to_the_plain_step(method, bad):
    """A conditioner whose step is not a finite descent step is ignored:
    the run matches the unconditioned one step for step."""
    fn = (lambda v: -v) if bad == "ascent" else (lambda v: v * np.nan)
    pre = Refreshing(fn)
    q = Quadratic()
    got = [h[1] for h in _run(method, q, precond=pre).history]
    want = [h[1] for h in _run(method, Quadratic()).history]
    assert pre.refreshed == 12
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_line_search_failure_stops_the_run():
    """A gradient of the wrong sign makes every trial step increase the
    misfit: the search fails and the run stops (optimize/base.py:47-50)
    with the model unchanged."""
    q = Quadratic(sign=-1.0)
    m0 = q.m.copy()
    opt = _run("lbfgs", q, niter=5)
    assert q.ngrad == 1
    np.testing.assert_array_equal(q.m, m0)
    assert opt.history[-1][1] == pytest.approx(opt.history[0][1])


def test_concave_bracket_raises():
    x = np.array([0.0, 1.0, 2.0])
    f = np.array([2.0, 1.0, 0.5])  # minimum at the end: an under-determined fit opens downwards
    with pytest.raises(LineSearchFitError):
        bracket.polyfit(x, f)


def test_not_descent_restarts_lbfgs():
    """An L-BFGS memory that yields an ascent direction restarts from the
    steepest descent direction (lbfgs.py:54-57)."""
    q = Quadratic()
    opt = lbfgs()
    opt.config = dict(CONFIG)
    opt.mpi = None

    class W:
        pass

    w = W()
    w.solver = q
    opt.setup(w)
    opt.precond = None
    g = np.ones(q.nx)
    opt.g_old, opt.m_old = g.copy(), np.zeros(q.nx)
    opt.g_new, opt.m_new = 2 * g, -np.ones(q.nx)  # y.s < 0: curvature pair of a concave function
    p = opt.compute_direction()
    np.testing.assert_array_equal(p, -opt.g_new)
    assert opt.lbfgs_s == []
