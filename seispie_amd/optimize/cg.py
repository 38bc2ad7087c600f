import numpy as np

from seispie_amd.optimize.base import base
from seispie_amd.optimize.line_search.bracket import bracket


class cg(base):
    """Polak-Ribiere conjugate gradient (seispie/optimize/cg.py:5-49).

    self.precond, when set, is a callable g -> P g applied to the gradient
    before the direction update (a step conditioner); a direction that is
    not a descent direction restarts from the plain steepest descent.
    """

    precond = None

    def setup(self, workflow):
        self.solver = workflow.solver
        self.bracket = bracket(workflow.solver, self.config)
        self.bracket.head = 0 if (getattr(self, 'mpi', None) and self.mpi.rank()) else 1

    def line_search(self, misfit):
        return self.bracket.run(self.m_new, self.g_new, self.p_new, misfit)

    def restart_search(self):
        self.bracket.restart()

    @staticmethod
    def pollak(g_new, g_old):
        return np.dot(g_new, g_new - g_old) / np.dot(g_old, g_old)

    @staticmethod
    def fletcher(g_new, g_old):
        return np.dot(g_new, g_new) / np.dot(g_old, g_old)

    def _conditioned(self, g):
        if self.precond is None:
            return g
        pg = self.precond(g)
        # the unconditioned step is the fallback when P g is not a finite descent step
        return pg if np.all(np.isfinite(pg)) and np.dot(pg, g) > 0 else g

    def compute_direction(self):
        if self.precond is not None and hasattr(self.precond, 'refresh'):
            self.precond.refresh(self.g_new)
        g_new = self._conditioned(self.g_new)
        if not hasattr(self, 'g_old'):
            return -g_new
        beta = self.pollak(g_new, self._conditioned(self.g_old))
        p_new = -g_new + beta * self.p_old
        if np.dot(p_new, self.g_new) > 0:
            print('  restarting CG: not descent')
            self.restart_search()
            return -self.g_new
        return p_new
