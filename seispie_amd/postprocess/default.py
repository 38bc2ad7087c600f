class default:
    """No-op postprocess plugin (seispie/postprocess/default.py:3-5); the
    adjoint and inversion workflows list the section, so the runner loads it."""

    def setup(self, workflow):
        pass
