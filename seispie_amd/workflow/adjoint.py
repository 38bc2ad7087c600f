from time import time

from seispie_amd.workflow.base import base


class adjoint(base):
    """Gradient of the misfit for the initial model
    (seispie/workflow/adjoint.py:5-34); writes proc000000_kmu.bin."""

    def setup(self):
        pass

    def run(self):
        solver = self.solver
        solver.import_model(1)
        solver.import_sources()
        solver.import_stations()
        solver.import_traces()

        start = time()
        solver.import_model(0)
        gradient, _, _ = solver.compute_gradient()

        if not self.mpi or self.mpi.rank() == 0:
            print('')
            print('Elapsed time: %.2fs' % (time() - start))
            solver.export_field(gradient, 'kmu')

    @property
    def modules(self):
        return ['solver', 'postprocess']
