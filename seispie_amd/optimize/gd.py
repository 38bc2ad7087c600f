from seispie_amd.optimize.cg import cg


class gd(cg):
    """Steepest descent (seispie/optimize/gd.py:3-5)."""

    def compute_direction(self):
        return -self._conditioned(self.g_new)
