"""Optimizers of the training loop (host side, FP64), one instance per element network.

Each optimizer is a state machine fed, once per training iteration, with the loss value and the
normalised gradient at the current parameters, and answering with the next parameters -- the
protocol of ``TOptimizer_next`` (``optimizers.F90:184-221``).  The algorithms and their
parameters follow the reference's DFTB+-derived modules, because a training run must reproduce
the reference's parameter trajectory:

* ``SteepestDescent``: ``steepdesc.F90:125-204``;
* ``LineMin``: the line minimiser of ``linemin.F90`` (bracketing by golden-section growth, then
  cubic / secant / bisection steps on the directional derivative);
* ``ConjugateGradient``: ``conjgrad.F90`` (Polak-Ribiere directions, one ``LineMin`` per
  direction, first step 5 |g|);
* ``Lbfgs``: ``lbfgs.F90`` (two-loop recursion, ``LineMin`` with first step 1, or a fixed
  step capped at the maximal displacement);
* ``Fire``: ``fire.F90`` (FIRE of Bitzek et al., PRL 97, 170201 (2006)).

Tolerances and displacements are those of the Training block (``initprogram.F90:1130-1185``).
"""
import numpy as np

EPS = np.finfo(np.float64).eps
GOLDEN_MEAN_P1 = 1.6180339887498949  # (1 + sqrt(5)) / 2, constants.F90


class SteepestDescent:
    def __init__(self, n, tol, max_disp, learning_rate):
        self.tol, self.max_disp = tol, max_disp
        self.weight = np.full(n, float(learning_rate))
        self.x = None
        self.converged = False

    def reset(self, x0):
        self.x = np.array(x0, dtype=np.float64)
        self.converged = False

    def next(self, fx, g):
        if self.converged:
            return self.x.copy(), True
        if np.max(np.abs(g)) < self.tol:
            self.converged = True
            return self.x.copy(), True
        step = -self.weight * g
        m = np.max(np.abs(step))
        if m > self.max_disp:
            step = (self.max_disp / m) * step
        self.x = self.x + step
        return self.x.copy(), False


class LineMin:
    """Line minimisation along x0 + lambda d (d normalised) from directional derivatives."""

    def __init__(self, n, max_iter, tol, max_disp):
        self.max_iter, self.tol, self.max_disp = max_iter, tol, max_disp

    def reset(self, x0, d0, first_step):
        self.state = 1
        self.it = 0
        self.xx = [0.0, 0.0]  # bracket abscissae (then: minimum value, minimum lambda)
        self.dx = [0.0, 0.0]  # derivatives at the bracket
        self.x0 = np.array(x0, dtype=np.float64)
        self.d0 = np.asarray(d0, dtype=np.float64) / np.sqrt(np.sum(np.asarray(d0) ** 2))
        self.xcur = 0.0
        self.first_step = first_step
        self.max_x = self.max_disp / np.max(np.abs(self.d0))
        self.converged = False

    def next(self, fx, g):
        if self.converged:
            return self.x0.copy(), True
        self.it += 1
        dcur = float(np.dot(g, self.d0))
        if (abs(dcur) < self.tol and self.state == 3) or self.it > self.max_iter:
            self.x0 = self.x0 + self.xcur * self.d0
            self.d0 = np.array(g, dtype=np.float64)  # holds the gradient at the minimum
            self.xx = [fx, self.xcur]
            self.converged = True
            return self.x0.copy(), True
        state = self.state
        xnew = None
        nxt = state
        if state == 1:
            if dcur > 0.0:
                self.d0 = -self.d0
                dcur = -dcur
            self.xx[0] = self.xcur
            self.dx[0] = dcur
            xnew = self.first_step
            nxt = 2
        if state == 2:
            if dcur < 0.0:
                # still going down: extend the bracket by the golden ratio
                xnew = self.xcur + GOLDEN_MEAN_P1 * (self.xcur - self.xx[0])
                self.xx[0] = self.xcur
                self.dx[0] = dcur
                nxt = 2
            else:
                state = 3
        if state == 3:
            done = False
            nxt = 3
            x1, x2, d1, d2 = self.xx[0], self.xx[1], self.dx[0], self.dx[1]
            if self.it > 2 and abs(x1 - x2) > EPS:
                # cubic through the current point and both bracket ends
                qq = (self.xcur - x1) / (x1 - x2)
                tmp = dcur - d1
                cc = qq * qq * (d2 - d1)
                aa = qq * tmp + cc
                bb = 2.0 * qq * tmp + cc + tmp
                cc = (qq + 1.0) * dcur
                tmp = bb * bb - 4 * aa * cc
                if tmp >= 0.0:
                    tmp = np.sqrt(tmp)
                    aa = bb + tmp if bb > 0.0 else bb - tmp
                    if abs(aa) > EPS:
                        xn = self.xcur - (self.xcur - x1) * 2.0 * cc / aa
                        if ((x1 < xn) == (xn < x2)) and ((dcur < 0.0) == (xn > self.xcur)):
                            if xn > self.xcur:
                                self.xx[0], self.dx[0] = self.xcur, dcur
                            else:
                                self.xx[1], self.dx[1] = self.xcur, dcur
                            xnew = xn
                            done = True
            if not done:
                if dcur < 0.0:
                    self.xx[0], self.dx[0] = self.xcur, dcur
                else:
                    self.xx[1], self.dx[1] = self.xcur, dcur
                tmp = self.dx[0] - self.dx[1]
                if abs(tmp) > EPS:
                    xn = (self.xx[1] * self.dx[0] - self.xx[0] * self.dx[1]) / tmp
                    if (self.xx[0] < xn) == (xn < self.xx[1]):
                        xnew = xn
                        done = True
            if not done:
                xnew = 0.5 * (self.xx[0] + self.xx[1])
        step = xnew - self.xcur
        if abs(step) <= self.max_x:
            self.xcur = xnew
        else:
            self.xcur = self.xcur + np.copysign(self.max_x, step)
        self.state = nxt
        return self.x0 + self.xcur * self.d0, False

    # results after convergence
    def min_x(self):
        return self.x0.copy()

    def min_grad(self):
        return self.d0.copy()

    def min_lambda(self):
        return self.xx[1]


class ConjugateGradient:
    def __init__(self, n, tol, max_disp):
        self.tol = tol
        self.line = LineMin(n, 10, 10000.0, max_disp)

    def reset(self, x0):
        self.u = np.array(x0, dtype=np.float64)
        self.state = 1
        self.converged = False

    def next(self, fx, g):
        if self.converged:
            return self.u.copy(), True
        g = np.asarray(g, dtype=np.float64)
        if self.state == 1:
            if np.max(np.abs(g)) < self.tol:
                # converged at the starting point: the parameters stay where they are
                self.converged = True
                self.state = 2
                return self.u.copy(), True
            self.gg = -g
            self.hh = self.gg.copy()
            self.line.reset(self.u, self.hh, 5.0 * np.sqrt(np.sum(self.gg ** 2)))
            self.state = 2
        conv_line = True
        while not self.converged and conv_line:
            self.u, conv_line = self.line.next(fx, g)
            if conv_line:
                xi = self.line.min_grad()
                if np.max(np.abs(xi)) < self.tol:
                    self.converged = True
                else:
                    gg2 = float(np.dot(self.gg, self.gg))
                    dg2 = float(np.dot(xi + self.gg, xi))
                    if gg2 < EPS:
                        self.converged = True
                    else:
                        self.gg = -xi
                        self.hh = self.gg + (dg2 / gg2) * self.hh
                        self.u = self.line.min_x()
                        self.line.reset(self.u, self.hh, self.line.min_lambda())
        if self.converged:
            self.u = self.line.min_x()
        return self.u.copy(), self.converged


class Lbfgs:
    def __init__(self, n, tol, min_disp, max_disp, memory, line_search=True):
        self.n, self.tol, self.mem = n, tol, max(1, int(memory))
        self.line = LineMin(n, 10, tol, max_disp) if line_search else None
        self.max_disp = -1.0  # the fixed-step variant is not capped (isQNDisp is false)

    def reset(self, x0):
        self.x = np.array(x0, dtype=np.float64)
        self.x_old = self.x.copy()
        self.g_old = np.zeros(self.n)
        self.s = np.zeros((self.mem, self.n))
        self.y = np.zeros((self.mem, self.n))
        self.rho = np.zeros(self.mem)
        self.dir = np.zeros(self.n)
        self.it = 0

    def _direction(self, x, g):
        if self.it == 0:
            self.g_old = g.copy()
            self.dir = -g
            self.it = 1
            return
        new = (self.it - 1) % self.mem
        self.s[new] = x - self.x_old
        self.y[new] = g - self.g_old
        beta = float(np.dot(self.y[new], self.s[new]))
        self.rho[new] = 1.0 / max(abs(beta), EPS)
        self.x_old = x.copy()
        self.g_old = g.copy()
        if self.it == 1:
            diag = 1.0
        else:
            diag = float(np.dot(self.s[new], self.y[new])) / max(
                float(np.dot(self.y[new], self.y[new])), EPS)
        q = g.copy()
        alpha = np.zeros(self.mem)
        for m in range(self.mem):  # newest to oldest
            i = (new - m) % self.mem
            alpha[i] = self.rho[i] * float(np.dot(self.s[i], q))
            q = q - alpha[i] * self.y[i]
        d = diag * q
        oldest = (new + 1) % self.mem
        for m in range(self.mem):  # oldest to newest
            i = (oldest + m) % self.mem
            b = self.rho[i] * float(np.dot(self.y[i], d))
            d = d + self.s[i] * (alpha[i] - b)
        self.it += 1
        self.dir = -d

    def next(self, fx, g):
        g = np.asarray(g, dtype=np.float64)
        if np.max(np.abs(g)) < self.tol:
            return self.x.copy(), True
        gtmp = g
        if self.line is not None and self.it > 0:
            self.x, conv = self.line.next(fx, g)
            if not conv:
                return self.x.copy(), False
            gtmp = self.line.min_grad()
        self._direction(self.x, gtmp)
        if self.line is not None:
            self.line.reset(self.x, self.dir, 1.0)
            self.x, _ = self.line.next(fx, g)
        else:
            if self.max_disp > 0.0:
                m = np.max(np.abs(self.dir))
                if m > self.max_disp:
                    self.dir = self.dir * self.max_disp / m
            self.x = self.x + self.dir
        return self.x.copy(), False


class Fire:
    def __init__(self, n, tol, max_step):
        self.tol = tol
        self.n_min, self.a_start, self.f_inc, self.f_dec, self.f_alpha = 5, 0.1, 1.1, 0.5, 0.99
        self.dt_max = max_step
        self.dt_init = 0.1 * max_step

    def reset(self, x0):
        self.it = 0
        self.reset_step = 0
        self.a = self.a_start
        self.dt = self.dt_init
        self.v = np.zeros(len(x0))
        self.x = np.array(x0, dtype=np.float64)

    def next(self, fx, g):
        g = np.asarray(g, dtype=np.float64)
        conv = bool(np.max(np.abs(g)) < self.tol)
        if not conv:
            p = -float(np.dot(self.v, g))
            vn, gn = np.sqrt(np.sum(self.v ** 2)), np.sqrt(np.sum(g ** 2))
            self.v = (1.0 - self.a) * self.v - self.a * g * vn / gn
            if p < 0.0:
                self.v[:] = 0.0
                self.reset_step = self.it
                self.dt *= self.f_dec
                self.a = self.a_start
            elif self.it - self.reset_step > self.n_min:
                self.dt = min(self.dt * self.f_inc, self.dt_max)
                self.a *= self.f_alpha
            self.v = self.v - self.dt * g
            self.x = self.x + self.v * self.dt
        self.it += 1
        return self.x.copy(), conv


def make(kind, n, tol, min_disp, max_disp, learning_rate=0.01, memory=1000, line_search=True):
    """Optimizer of the Training block type: 'sd', 'cg', 'lbfgs' or 'fire'."""
    kind = kind.lower()
    if kind == "sd":
        return SteepestDescent(n, tol, max_disp, learning_rate)
    if kind == "cg":
        return ConjugateGradient(n, tol, max_disp)
    if kind == "lbfgs":
        return Lbfgs(n, tol, min_disp, max_disp, memory, line_search)
    if kind == "fire":
        return Fire(n, tol, max_disp)
    raise ValueError("unknown optimizer %r" % kind)
