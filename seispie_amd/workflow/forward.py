from time import time

from seispie_amd.workflow.base import base


class forward(base):
    """Forward simulation (seispie/workflow/forward.py:4-30)."""

    def setup(self):
        pass

    def run(self):
        solver = self.solver
        start = time()
        solver.import_model(1)
        solver.import_sources()
        solver.import_stations()
        solver.run_forward()
        print('elapsed time: %.2fs' % (time() - start))

    @property
    def modules(self):
        return ['solver']
