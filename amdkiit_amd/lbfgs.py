"""Limited-memory BFGS with the More-Thuente line search, in the reverse-
communication form the reference's geometry optimiser drives
(do_opt.f90:197-266 calling LBFGS of lbfgs.f:9).

The algorithm is the published one: the two-loop recursion with M stored
corrections and the scaling gamma_k = s^T y / y^T y of the initial inverse
Hessian (Liu & Nocedal, Math. Prog. 45 (1989) 503), and the line search with
guaranteed sufficient decrease of More & Thuente (ACM TOMS 20 (1994) 286):
safeguarded cubic/quadratic steps on the modified function psi(a) = f(a) -
f(0) - ftol a f'(0) until the first point meets the sufficient-decrease test
and |f'(a)| <= gtol |f'(0)|.  The settings are the reference's: M = 5,
FTOL = 1e-4, GTOL = 0.9, MAXFEV = 20, STPMIN = 1e-20, STPMAX = 1e20,
XTOL = 1e-16 (do_opt.f90:226-231, lbfgs.f:288-289 and its LB2 defaults), the
first trial step 1/|g| and unit steps afterwards, and convergence when
|g| <= EPS (an absolute test, lbfgs.f:396).

`LBFGS.step(x, f, g)` is called after every energy/gradient evaluation and
either asks for the next evaluation (returns 1, x updated in place) or
finishes (0) or fails (< 0), as IFLAG does.  `accepted` becomes True in a
call that completed an iteration; the reference prints the geometry that is
evaluated next (i_opt, set by LB1 at lbfgs.f:473).
"""
import math

import numpy as np


class LineSearchError(RuntimeError):
    pass


def _cubic_min(a, fa, da, b, fb, db):
    """Minimiser of the cubic through (a, fa, da), (b, fb, db), and gamma's
    sign convention: returns (theta, gamma) with gamma = sqrt(theta^2 - da db)
    (scaled to avoid overflow), gamma taking the sign that picks the minimum
    when measured from a."""
    theta = 3.0 * (fa - fb) / (b - a) + da + db
    s = max(abs(theta), abs(da), abs(db))
    disc = (theta / s) ** 2 - (da / s) * (db / s)
    return theta, s, disc


def more_thuente_step(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    """One safeguarded step of the More-Thuente search (their Section 4 and
    the four cases of the trial-value selection).  (stx, fx, dx) is the best
    step so far, (sty, fy, dy) the other end of the interval, (stp, fp, dp)
    the current trial.  Returns the updated interval, the new trial step,
    brackt and a case number (0 = bad input)."""
    if (brackt and (stp <= min(stx, sty) or stp >= max(stx, sty))) or dx * (stp - stx) >= 0.0 or stpmax < stpmin:
        return stx, fx, dx, sty, fy, dy, stp, brackt, 0
    sgnd = dp * (dx / abs(dx))
    if fp > fx:
        # case 1: higher value, the minimum is bracketed; cubic step if it is
        # closer to stx than the quadratic one, else their mean
        case, bound = 1, True
        theta, s, disc = _cubic_min(stx, fx, dx, stp, fp, dp)
        gamma = s * math.sqrt(disc)
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        stpc = stx + (p / q) * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0.0:
        # case 2: derivatives of opposite sign, bracketed; the cubic or the
        # secant step, whichever is farther from stp
        case, bound = 2, False
        theta, s, disc = _cubic_min(stx, fx, dx, stp, fp, dp)
        gamma = s * math.sqrt(disc)
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        stpc = stp + (p / q) * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(dp) < abs(dx):
        # case 3: same sign, the derivative decreases in magnitude; the cubic
        # step only if it lies beyond stp (else the interval end), compared
        # with the secant step
        case, bound = 3, True
        theta, s, disc = _cubic_min(stx, fx, dx, stp, fp, dp)
        gamma = s * math.sqrt(max(0.0, disc))
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        if r < 0.0 and gamma != 0.0:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stp - stpc) < abs(stp - stpq) else stpq
        else:
            stpf = stpc if abs(stp - stpc) > abs(stp - stpq) else stpq
    else:
        # case 4: same sign, the derivative does not decrease: cubic step
        # towards sty when bracketed, else to the interval end
        case, bound = 4, False
        if brackt:
            theta, s, disc = _cubic_min(stp, fp, dp, sty, fy, dy)
            gamma = s * math.sqrt(disc)
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            stpf = stp + (p / q) * (sty - stp)
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    # new interval of uncertainty
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0.0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    stpf = max(stpmin, min(stpmax, stpf))
    stp = stpf
    if brackt and bound:  # stay within 0.66 of the interval from stx
        lim = stx + 0.66 * (sty - stx)
        stp = min(lim, stp) if sty > stx else max(lim, stp)
    return stx, fx, dx, sty, fy, dy, stp, brackt, case


class MoreThuente:
    """Reverse-communication line search along s from x0 (the state the
    reference's MCSRCH keeps between its calls).  `start` gives the first
    trial point; `check(f, g)` takes the value there and returns 1 (done: the
    strong Wolfe conditions hold), -1 (evaluate at self.x) or another code
    (2: interval below xtol, 3: maxfev, 4: stpmin, 5: stpmax, 6: rounding)."""

    P5, P66, XTRAPF = 0.5, 0.66, 4.0

    def __init__(self, ftol, gtol, xtol, stpmin, stpmax, maxfev):
        self.ftol, self.gtol, self.xtol = ftol, gtol, xtol
        self.stpmin, self.stpmax, self.maxfev = stpmin, stpmax, maxfev

    def start(self, x, f, g, s, stp):
        if stp <= 0.0:
            raise LineSearchError("non-positive initial step")
        self.dginit = float(g @ s)
        if self.dginit >= 0.0:
            raise LineSearchError("search direction is not a descent direction")
        self.s = s
        self.x0 = x.copy()
        self.brackt, self.stage1, self.infoc, self.nfev = False, True, 1, 0
        self.finit = f
        self.dgtest = self.ftol * self.dginit
        self.width = self.stpmax - self.stpmin
        self.width1 = self.width / self.P5
        self.stx, self.fx, self.dgx = 0.0, f, self.dginit
        self.sty, self.fy, self.dgy = 0.0, f, self.dginit
        self.stp = stp
        return self._next_trial()

    def _next_trial(self):
        if self.brackt:
            self.stmin, self.stmax = min(self.stx, self.sty), max(self.stx, self.sty)
        else:
            self.stmin = self.stx
            self.stmax = self.stp + self.XTRAPF * (self.stp - self.stx)
        stp = min(max(self.stp, self.stpmin), self.stpmax)
        # no progress possible: fall back to the best step so far
        if ((self.brackt and (stp <= self.stmin or stp >= self.stmax)) or self.nfev >= self.maxfev - 1
                or self.infoc == 0 or (self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax)):
            stp = self.stx
        self.stp = stp
        self.x = self.x0 + stp * self.s
        return -1

    def check(self, f, g):
        self.nfev += 1
        dg = float(g @ self.s)
        ftest1 = self.finit + self.stp * self.dgtest
        info = 0
        if (self.brackt and (self.stp <= self.stmin or self.stp >= self.stmax)) or self.infoc == 0:
            info = 6
        if self.stp == self.stpmax and f <= ftest1 and dg <= self.dgtest:
            info = 5
        if self.stp == self.stpmin and (f > ftest1 or dg >= self.dgtest):
            info = 4
        if self.nfev >= self.maxfev:
            info = 3
        if self.brackt and self.stmax - self.stmin <= self.xtol * self.stmax:
            info = 2
        if f <= ftest1 and abs(dg) <= self.gtol * (-self.dginit):
            info = 1
        if info != 0:
            return info
        # first stage: the modified function psi until it has a nonpositive
        # value and a nonnegative derivative
        if self.stage1 and f <= ftest1 and dg >= min(self.ftol, self.gtol) * self.dginit:
            self.stage1 = False
        if self.stage1 and f <= self.fx and f > ftest1:
            t = self.dgtest
            r = more_thuente_step(self.stx, self.fx - self.stx * t, self.dgx - t, self.sty, self.fy - self.sty * t,
                                  self.dgy - t, self.stp, f - self.stp * t, dg - t, self.brackt, self.stmin,
                                  self.stmax)
            stx, fxm, dgxm, sty, fym, dgym, stp, self.brackt, self.infoc = r
            self.stx, self.sty, self.stp = stx, sty, stp
            self.fx, self.fy = fxm + stx * t, fym + sty * t
            self.dgx, self.dgy = dgxm + t, dgym + t
        else:
            r = more_thuente_step(self.stx, self.fx, self.dgx, self.sty, self.fy, self.dgy, self.stp, f, dg,
                                  self.brackt, self.stmin, self.stmax)
            self.stx, self.fx, self.dgx, self.sty, self.fy, self.dgy, self.stp, self.brackt, self.infoc = r
        if self.brackt:  # force sufficient shrinking of the interval
            if abs(self.sty - self.stx) >= self.P66 * self.width1:
                self.stp = self.stx + self.P5 * (self.sty - self.stx)
            self.width1 = self.width
            self.width = abs(self.sty - self.stx)
        return self._next_trial()


class LBFGS:
    """L-BFGS minimiser driven by reverse communication (see module doc)."""

    def __init__(self, n, m=5, eps=1e-4, xtol=1e-16, gtol=0.9, ftol=1e-4, maxfev=20, stpmin=1e-20, stpmax=1e20):
        self.n, self.m, self.eps = n, m, eps
        self.ls = MoreThuente(ftol, gtol, xtol, stpmin, stpmax, maxfev)
        self.S = np.zeros((m, n))
        self.Y = np.zeros((m, n))
        self.rho = np.zeros(m)
        self.stored = 0      # corrections held (<= m)
        self.point = 0       # slot of the next correction
        self.iter = 0
        self.nfun = 1
        self.started = False
        self.accepted = False
        self.gnorm = None

    def _direction(self, g):
        """-H g by the two-loop recursion over the stored (s, y), newest first,
        with H0 = (s^T y / y^T y) of the newest pair (identity before any)."""
        q = -g.copy()
        k = self.stored
        alpha = np.zeros(self.m)
        order = [(self.point - 1 - i) % self.m for i in range(k)]  # newest first
        for j in order:
            alpha[j] = self.rho[j] * (self.S[j] @ q)
            q -= alpha[j] * self.Y[j]
        if k > 0:
            j = order[0]
            q *= (self.S[j] @ self.Y[j]) / (self.Y[j] @ self.Y[j])
        for j in reversed(order):
            beta = self.rho[j] * (self.Y[j] @ q)
            q += (alpha[j] - beta) * self.S[j]
        return q

    def step(self, x, f, g):
        """x (updated in place), f, g at x -> 1 evaluate again, 0 converged,
        < 0 line-search failure (the reference's IFLAG = -1 with INFO)."""
        x_ = np.asarray(x, np.float64)
        g = np.asarray(g, np.float64).reshape(-1)
        self.accepted = False
        if not self.started:
            self.started = True
            self.gnorm = math.sqrt(float(g @ g))
            self.iter = 1
            d = -g
            self._g_prev = g.copy()
            self.ls.start(x_.reshape(-1), f, g, d, 1.0 / self.gnorm)
            x_.reshape(-1)[:] = self.ls.x
            return 1
        info = self.ls.check(f, g)
        if info == -1:
            x_.reshape(-1)[:] = self.ls.x
            return 1
        if info != 1:
            self.info = info
            return -1
        # iteration complete: store the correction pair
        self.nfun += self.ls.nfev
        sv = self.ls.stp * self.ls.s
        yv = g - self._g_prev
        self.S[self.point] = sv
        self.Y[self.point] = yv
        self.rho[self.point] = 1.0 / float(yv @ sv)
        self.point = (self.point + 1) % self.m
        self.stored = min(self.stored + 1, self.m)
        self.gnorm = math.sqrt(float(g @ g))
        self.accepted = True
        if self.gnorm <= self.eps:
            return 0
        self.iter += 1
        self._g_prev = g.copy()
        d = self._direction(g)
        self.ls.start(x_.reshape(-1), f, g, d, 1.0)
        x_.reshape(-1)[:] = self.ls.x
        return 1
