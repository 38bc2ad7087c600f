import numpy as np


class LineSearchFitError(Exception):
    """The parabola through the bracketing points opens downwards
    (bracket.py:150-156 raises a bare Exception there)."""


class bracket:
    """Bracketing line search (seispie/optimize/line_search/bracket.py:3-164).

    History lists hold every trial step length (0 marks the start of a
    search), its misfit, and g.g / g.p of each search.  calculate_step
    returns (alpha, status): status 1 = accept alpha, 0 = try alpha,
    -1 = failed.
    """

    max_restarts = 8  # the reference restarts without bound

    def __init__(self, solver, config):
        self.solver = solver
        self.ls_lens, self.ls_vals, self.ls_gtg, self.ls_gtp = [], [], [], []
        self.ls_step = int(config['ls_step'])
        self.ls_step_max = float(config['ls_step_max'])
        self.ls_step_init = float(config['ls_step_init'])
        self.ls_thresh = float(config['ls_thresh'])
        self.head = 1

    def restart(self):
        self.ls_lens, self.ls_vals, self.ls_gtg, self.ls_gtp = [], [], [], []

    @staticmethod
    def angle(x, y):
        c = np.dot(x, y) / (np.dot(x, x) * np.dot(y, y)) ** 0.5
        if abs(c) > 1:
            c /= abs(c)
        return np.arccos(c)

    def run(self, m_new, g_new, p_new, f, _depth=0):
        if self.head:
            print('  step 1')
        solver = self.solver
        norm_m = np.amax(abs(m_new))
        norm_p = np.amax(abs(p_new))
        if not (np.isfinite(norm_p) and norm_p > 0):
            # a vanished (exact minimum) or non-finite direction: the
            # reference's step_max is then inf and its backtracking loops
            # for ever comparing the trial count with it
            return -1
        gtg = np.dot(g_new, g_new)
        gtp = np.dot(g_new, p_new)
        step_max = self.ls_step_max * norm_m / norm_p
        step_count = 0
        self.ls_lens.append(0)
        self.ls_vals.append(f)
        self.ls_gtg.append(gtg)
        self.ls_gtp.append(gtp)

        if self.ls_step_init and len(self.ls_lens) <= 1:
            alpha = self.ls_step_init * norm_m / norm_p
        else:
            alpha, status = self.calculate_step(step_count, step_max)

        while True:
            solver.update_model(m_new + alpha * p_new)
            self.ls_lens.append(alpha)
            self.ls_vals.append(solver.compute_misfit())
            step_count += 1
            alpha, status = self.calculate_step(step_count, step_max)
            if self.head:
                print('  step %d' % (step_count + 1))
            if status > 0:
                solver.update_model(m_new + alpha * p_new)
                return status
            if status < 0:
                solver.update_model(m_new)
                if self.angle(p_new, -g_new) < 1e-3 or _depth >= self.max_restarts:
                    return status
                if self.head:
                    print('  restarting line search')
                self.restart()
                return self.run(m_new, g_new, p_new, f, _depth + 1)

    def get_history(self, i):
        """Last i+1 trials sorted by |step|, and the number of earlier searches."""
        k = len(self.ls_lens)
        x = np.array(self.ls_lens[k - i - 1:k])
        f = np.array(self.ls_vals[k - i - 1:k])
        order = abs(x).argsort()
        return x[order], f[order], sum(np.array(self.ls_lens) == 0) - 1

    def calculate_step(self, step_count, step_max):
        x, f, update_count = self.get_history(step_count)
        if step_count == 0:
            if update_count == 0:
                alpha, status = 1 / self.ls_gtg[-1], 0
            else:
                idx = np.argmin(self.ls_vals[:-1])
                alpha, status = self.ls_lens[idx] * self.ls_gtp[-2] / self.ls_gtp[-1], 0
        elif self.check_bracket(x, f):
            if self.good_enough(x, f):
                alpha, status = x[f.argmin()], 1
            else:
                alpha, status = self.polyfit(x, f), 0
        elif step_count <= self.ls_step:
            if all(f <= f[0]):
                alpha, status = 1.618034 * x[step_count], 0
            else:
                slope = self.ls_gtp[-1] / self.ls_gtg[-1]
                alpha, status = self.backtrack(f[0], slope, x[1], f[1], 0.1, 0.5), 0
        else:
            alpha, status = 0, -1

        if alpha > step_max:
            if step_count == 0:
                alpha, status = 0.618034 * step_max, 0
            else:
                alpha, status = step_max, 1
        return alpha, status

    @staticmethod
    def check_bracket(x, f):
        imin, fmin = f.argmin(), f.min()
        return 1 if (fmin < f[0]) and any(f[imin:] > fmin) else 0

    def good_enough(self, x, f):
        if not self.check_bracket(x, f):
            return 0
        x0 = self.polyfit(x, f)
        return 1 if any(np.abs(np.log10(x[1:] / x0)) < np.log10(1.2)) else 0

    @staticmethod
    def polyfit(x, f):
        i = np.argmin(f)
        p = np.polyfit(x[i - 1:i + 2], f[i - 1:i + 2], 2)
        if p[0] > 0:
            return -p[1] / (2 * p[0])
        raise LineSearchFitError('concave parabola in the bracket')

    @staticmethod
    def backtrack(f0, g0, x1, f1, b1, b2):
        """Minimiser of the quadratic through (0, f0) with slope g0 and
        (x1, f1), kept within [b1 x1, b2 x1]."""
        x2 = -g0 * x1 ** 2 / (2 * (f1 - f0 - g0 * x1))
        if x2 > b2 * x1:
            x2 = b2 * x1
        elif x2 < b1 * x1:
            x2 = b1 * x1
        return x2
