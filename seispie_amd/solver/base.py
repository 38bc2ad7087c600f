class base:
    """Solver plugin interface (seispie/solver/base.py:4-15)."""

    def setup(self, workflow):
        raise NotImplementedError

    def run_forward(self):
        raise NotImplementedError

    def run_adjoint(self):
        raise NotImplementedError

    def import_model(self, dir):
        raise NotImplementedError
