from time import time

import numpy as np


class base:
    """Optimizer driver (seispie/optimize/base.py:4-69).

    Each iteration: gradient, search direction, line search; the model is
    written as proc%06d_mu.bin.  run() also returns the normalised misfit
    history and keeps per-iteration wall times in self.history.
    """

    def setup(self, workflow):
        raise NotImplementedError

    def compute_direction(self):
        raise NotImplementedError

    def line_search(self, misfit):
        raise NotImplementedError

    def restart_search(self):
        raise NotImplementedError

    def run(self):
        solver = self.solver
        niter = int(self.config['niter'])
        start = time()
        misfits = []
        self.history = []
        head = 0 if (getattr(self, 'mpi', None) and self.mpi.rank()) else 1
        say = print if head and not getattr(self, 'quiet', False) else (lambda *a: None)
        report = getattr(self, 'on_iteration', None)

        for i in range(niter):
            say('Iteration %d' % (i + 1))
            self.g_new, misfit, self.m_new = solver.compute_gradient()
            if head:
                if i > 0:
                    solver.export_field(self.m_new, 'mu', i - 1)
                else:
                    ref = misfit
                misfits.append(misfit / ref)
            self.history.append((time() - start, float(misfit)))
            if report:
                report(i, self.history[-1][0], float(misfit))

            self.p_new = self.compute_direction()
            self.m_old = self.m_new
            self.p_old = self.p_new
            self.g_old = self.g_new

            if self.line_search(misfit) < 0:
                if head:
                    say('Line search failed')
                    break
            say('')

        misfit = solver.compute_misfit()
        self.history.append((time() - start, float(misfit)))
        if report:
            report(niter, self.history[-1][0], float(misfit))
        if head:
            misfits.append(misfit / ref)
            m_new = np.zeros(solver.nx * solver.nz, dtype='float32')
            solver.mu.copy_to_host(m_new, solver.stream)
            solver.stream.synchronize()
            solver.export_field(m_new, 'mu', niter)
            say('Misfits:')
            for m in misfits:
                say('  %.4f' % m)
            say('')
            say('Elapsed time: %.2fs' % (time() - start))
        return misfits
