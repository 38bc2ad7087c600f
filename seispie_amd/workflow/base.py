class base:
    """Workflow interface (seispie/workflow/base.py:3-24)."""

    def setup(self):
        raise NotImplementedError

    def run(self):
        raise NotImplementedError

    def call(self):
        raise NotImplementedError

    @property
    def modules(self):
        """Config sections whose plugins the runner loads, in order."""
        return []
