"""An independent restatement of QUADPACK's adaptive integrator QAG with the
61-point Gauss-Kronrod rule (key 6), for pinning the oracle's C version.

Written from the published algorithm (R. Piessens, E. de Doncker-Kapenga,
C. Ueberhuber, D. Kahaner, "QUADPACK", Springer 1983: routines DQAGE, DQK61,
DQPSRT) as GSL 2.x implements it (gsl_integration_qag: the same control flow,
with GSL's names for the round-off counters and its final summation of the
interval results).  The reference calls it as
gsl_integration_qag(F, log T0, log a, 0, relerr, 200, 6, ...)
(delta_tot_table.c:597, :404).  Test infrastructure only.

The rule sums run in mpmath at 40 significant digits on an integrand given as
a Python callable of an mpf; the control flow (bisection, the error-ordered
list, the round-off and singularity tests, the stopping test) is plain Python
on those values.  The workspace keeps GSL's interval order (the bisected
interval keeps its slot for the half with the larger error, the other half
is appended) and its end points are doubles (each midpoint rounded as a
double implementation rounds 0.5 (a + b)), so its interval list can be
compared entry by entry, bit for bit, with a double-precision implementation
that makes the same decisions.
"""
import mpmath

mp = mpmath.mp

DBL_EPSILON = 2.0 ** -52
DBL_MIN = 2.0 ** -1022

SUCCESS, EROUND, EMAXITER, ESING = 0, 18, 11, 21


def _kronrod_tables():
    """Nodes and weights of the 30-point Gauss and 61-point Kronrod rules on
    [-1, 1], as QUADPACK lists them (descending positive nodes, the centre
    last), rounded to double: the committed table, generated from the rules'
    definition (oracle/gen_quad_tables.py) and checked against QUADPACK's
    published digits by the oracle's own tests.  The rule here takes these
    double values exactly, as the double implementations do."""
    import os
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    src = open(os.path.join(here, "..", "oracle", "quad_tables.h")).read()

    def arr(name):
        body = re.search(name + r"\[\d+\]\s*=\s*\{(.*?)\};", src, re.S).group(1)
        return [mpmath.mpf(float.fromhex(t)) for t in re.findall(r"-?0x[0-9a-fA-F.]+p[-+]?\d+", body)]
    return arr("orc_xgk61"), arr("orc_wgk61"), arr("orc_wg30")


class Workspace:
    def __init__(self, a, b):
        self.a, self.b, self.r, self.e, self.level = [a], [b], [0], [0], [0]
        self.order = [0]
        self.nrmax = 0
        self.imax = 0
        self.bisected = []     # (index bisected, its error, the max error of the list) per step
        self.max_level = 0

    def size(self):
        return len(self.a)


class QAG:
    """QAG-GK61 at `dps` digits.  f: callable of an mpf returning an mpf (or float)."""

    def __init__(self, dps=40):
        self.dps = dps
        with mpmath.workdps(dps):
            self.xgk, self.wgk, self.wg = _kronrod_tables()

    # ---- DQK61: the rule on [a, b] with QUADPACK's error estimate
    def rule(self, f, a, b):
        c = (a + b) / 2
        h = (b - a) / 2
        fc = mpmath.mpf(f(c))
        res_g = mpmath.mpf(0)
        res_k = self.wgk[30] * fc
        res_abs = abs(res_k)
        f1, f2 = [None] * 30, [None] * 30
        for j in range(30):
            x = h * self.xgk[j]
            f1[j], f2[j] = mpmath.mpf(f(c - x)), mpmath.mpf(f(c + x))
            res_k += self.wgk[j] * (f1[j] + f2[j])
            res_abs += self.wgk[j] * (abs(f1[j]) + abs(f2[j]))
            if j % 2 == 1:  # the Gauss nodes are the Kronrod nodes of odd index
                res_g += self.wg[j // 2] * (f1[j] + f2[j])
        mean = res_k / 2
        res_asc = self.wgk[30] * abs(fc - mean)
        for j in range(30):
            res_asc += self.wgk[j] * (abs(f1[j] - mean) + abs(f2[j] - mean))
        err = abs((res_k - res_g) * h)
        res_k, res_abs, res_asc = res_k * h, res_abs * abs(h), res_asc * abs(h)
        if res_asc != 0 and err != 0:
            err = res_asc * min(mpmath.mpf(1), (200 * err / res_asc) ** mpmath.mpf(1.5))
        if res_abs > DBL_MIN / (50 * DBL_EPSILON):
            err = max(50 * DBL_EPSILON * res_abs, err)
        return res_k, err, res_abs, res_asc

    # ---- DQPSRT: keep the error list ordered (descending) down to `top`
    @staticmethod
    def _sort(w, limit):
        last = w.size() - 1
        if last < 2:
            w.order = [0, 1]
            w.imax = w.order[w.nrmax]
            return
        e, order = w.e, w.order
        order.append(None)
        nrmax = w.nrmax
        maxerr = order[nrmax]
        errmax = e[maxerr]
        while nrmax > 0 and errmax > e[order[nrmax - 1]]:
            order[nrmax] = order[nrmax - 1]
            nrmax -= 1
        top = last if last < limit // 2 + 2 else limit - last + 1
        i = nrmax + 1
        while i < top and errmax < e[order[i]]:
            order[i - 1] = order[i]
            i += 1
        order[i - 1] = maxerr
        errmin = e[last]
        k = top - 1
        while k > i - 2 and errmin >= e[order[k]]:
            order[k + 1] = order[k]
            k -= 1
        order[k + 1] = last
        w.nrmax = nrmax
        w.imax = order[nrmax]

    def integrate(self, f, a, b, epsabs, epsrel, limit=200):
        """(status, result, abserr, workspace)."""
        with mpmath.workdps(self.dps):
            a, b = mpmath.mpf(a), mpmath.mpf(b)
            w = Workspace(a, b)
            r0, e0, ra0, rs0 = self.rule(f, a, b)
            w.r[0], w.e[0] = r0, e0
            tol = max(mpmath.mpf(epsabs), epsrel * abs(r0))
            round_off = 50 * DBL_EPSILON * ra0
            if e0 <= round_off and e0 > tol:
                return EROUND, r0, e0, w
            if (e0 <= tol and e0 != rs0) or e0 == 0:
                return SUCCESS, r0, e0, w
            if limit == 1:
                return EMAXITER, r0, e0, w
            area, errsum = r0, e0
            n_round1 = n_round2 = 0
            error_type = 0
            iteration = 1
            while True:
                i = w.imax
                ai, bi, ri, ei = w.a[i], w.b[i], w.r[i], w.e[i]
                w.bisected.append((i, ei, max(w.e)))
                # end points stay doubles: the midpoint as 0.5 (a + b) rounds in
                # double (an exact halving of the rounded sum)
                a1, b1 = ai, mpmath.mpf(float((ai + bi) / 2))
                a2, b2 = b1, bi
                r1, e1, _, s1 = self.rule(f, a1, b1)
                r2, e2, _, s2 = self.rule(f, a2, b2)
                r12, e12 = r1 + r2, e1 + e2
                errsum += e12 - ei
                area += r12 - ri
                if s1 != e1 and s2 != e2:
                    if abs(ri - r12) <= mpmath.mpf("1e-5") * abs(r12) and e12 >= mpmath.mpf("0.99") * ei:
                        n_round1 += 1
                    if iteration >= 10 and e12 > ei:
                        n_round2 += 1
                tol = max(mpmath.mpf(epsabs), epsrel * abs(area))
                if errsum > tol:
                    if n_round1 >= 6 or n_round2 >= 20:
                        error_type = 2
                    small = (1 + 100 * DBL_EPSILON) * (abs(a2) + 1000 * DBL_MIN)
                    if abs(a1) <= small and abs(b2) <= small:
                        error_type = 3
                # the half with the larger error keeps the bisected slot
                lvl = w.level[i] + 1
                if e2 > e1:
                    w.a[i], w.r[i], w.e[i] = a2, r2, e2
                    w.a.append(a1), w.b.append(b1), w.r.append(r1), w.e.append(e1)
                else:
                    w.b[i], w.r[i], w.e[i] = b1, r1, e1
                    w.a.append(a2), w.b.append(b2), w.r.append(r2), w.e.append(e2)
                w.level[i] = lvl
                w.level.append(lvl)
                w.max_level = max(w.max_level, lvl)
                self._sort(w, limit)
                iteration += 1
                if not (iteration < limit and not error_type and errsum > tol):
                    break
            result = mpmath.fsum(w.r)
            if errsum <= tol:
                return SUCCESS, result, errsum, w
            if error_type == 2:
                return EROUND, result, errsum, w
            if error_type == 3:
                return ESING, result, errsum, w
            return EMAXITER, result, errsum, w


def natural_cspline(x, y):
    """Natural cubic spline through (x, y) at the working precision: returns
    an evaluator s(t) for x[0] <= t <= x[-1] (a plain tridiagonal solve for
    the second derivatives, then the cubic on the containing interval)."""
    n = len(x)
    x = [mpmath.mpf(v) for v in x]
    y = [mpmath.mpf(v) for v in y]
    m = [mpmath.mpf(0)] * n
    if n > 2:
        hs = [x[i + 1] - x[i] for i in range(n - 1)]
        diag = [2 * (hs[i - 1] + hs[i]) for i in range(1, n - 1)]
        rhs = [6 * ((y[i + 1] - y[i]) / hs[i] - (y[i] - y[i - 1]) / hs[i - 1]) for i in range(1, n - 1)]
        for i in range(1, n - 2):  # forward elimination (sub- and super-diagonal h_i)
            w = hs[i] / diag[i - 1]
            diag[i] -= w * hs[i]
            rhs[i] -= w * rhs[i - 1]
        sol = [mpmath.mpf(0)] * (n - 2)
        sol[-1] = rhs[-1] / diag[-1]
        for i in range(n - 4, -1, -1):
            sol[i] = (rhs[i] - hs[i + 1] * sol[i + 1]) / diag[i]
        m[1:n - 1] = sol

    import bisect

    def s(t):
        t = mpmath.mpf(t)
        i = min(max(bisect.bisect_right(x, t) - 1, 0), n - 2)
        h = x[i + 1] - x[i]
        A, B = (x[i + 1] - t) / h, (t - x[i]) / h
        return A * y[i] + B * y[i + 1] + ((A ** 3 - A) * m[i] + (B ** 3 - B) * m[i + 1]) * h * h / 6
    return s
