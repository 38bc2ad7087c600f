from seispie_amd.optimize.line_search.bracket import bracket


class backtrack(bracket):
    """Backtracking line search used by L-BFGS
    (seispie/optimize/line_search/backtrack.py:4-28)."""

    def calculate_step(self, step_count, step_max):
        x, f, update_count = self.get_history(step_count)
        if update_count == 0:
            return super().calculate_step(step_count, step_max)
        if step_count == 0:
            return min(1., step_max), 0
        if f.min() < f[0]:
            return x[f.argmin()], 1
        # the reference compares the trial count with the float step_max here
        if step_count <= step_max:
            slope = self.ls_gtp[-1] / self.ls_gtg[-1]
            return self.backtrack(f[0], slope, x[1], f[1], 0.1, 0.5), 0
        return 0, -1
