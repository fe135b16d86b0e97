"""Pins the oracle's QUADPACK QAG-GK61 (oracle/kn_oracle.c orc_qag61, the
integrator behind get_delta_nu and fslength, delta_tot_table.c:597, :404)
against an independent restatement of the published algorithm
(tests/quadpack_restated.py: DQAGE/DQK61/DQPSRT as GSL implements them,
rule sums at 40 digits).

For each integrand both run with the same tolerance and limit; they must
return the same status, the same number of intervals and the same interval
list (QUADPACK's workspace order, end points bit for bit), i.e. the same
sequence of bisection decisions, and results within 1e-13.  The restatement
also checks the algorithm's invariants on its own run: while fewer than
limit/2 + 2 intervals exist the bisected interval carries the largest error
of the list, and a successful run stops at the first step whose error sum
is within the tolerance.

Integrands: smooth, peaked, oscillatory, a square-root kink, a natural cubic
spline through noisy knots (QAG runs out of subdivisions there: the failure
the reference turns into terminate(2001)), and the get_delta_nu integrand of
oracle histories (fslength spline, 1/(aH), the specialJ fit, the delta_tot
spline; delta_tot_table.c:492-500) with the oracle's own recorded
subdivisions.
"""
import ctypes as C
import math

import mpmath
import numpy as np
import pytest

import oracle_py as O
import quadpack_restated as Q
from batch_common import Cosmo
from dnu_common import synthetic_history

ORC_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
DP = C.POINTER(C.c_double)


def oracle_qag(f_double, a, b, epsabs, epsrel, limit=200):
    L = O.lib()
    fn = L.orc_qag61_record
    fn.restype = C.c_int
    fn.argtypes = [ORC_FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, DP, DP,
                   C.POINTER(C.c_int), DP]
    cb = ORC_FN(lambda x, p: f_double(x))
    res, err, nint = C.c_double(), C.c_double(), C.c_int()
    ivl = np.zeros(2 * limit)
    rc = fn(cb, None, a, b, epsabs, epsrel, limit, C.byref(res), C.byref(err), C.byref(nint), ivl.ctypes.data_as(DP))
    n = nint.value
    return rc, res.value, err.value, ivl[:2 * n].reshape(n, 2)


def restated_lists(w):
    return np.array([[float(x), float(y)] for x, y in zip(w.a, w.b)])


def check_invariants(w, status, errsum, epsabs, epsrel, result, limit=200):
    for t, (i, ei, emax) in enumerate(w.bisected):
        if t + 1 < limit // 2 + 2:
            assert ei == emax, "step %d: bisected an interval that is not the largest error" % t
    if status == Q.SUCCESS and w.bisected:
        assert errsum <= max(epsabs, epsrel * abs(result))


def compare(name, f_double, f_mp, a, b, epsabs, epsrel, limit=200):
    rc, res, err, ivl = oracle_qag(f_double, a, b, epsabs, epsrel, limit)
    st, r2, e2, w = Q.QAG().integrate(f_mp, a, b, epsabs, epsrel, limit)
    check_invariants(w, st, e2, epsabs, epsrel, r2, limit)
    mine = restated_lists(w)
    print("  %-24s status %d/%d, %d/%d intervals, result %.17g vs %.17g" % (name, rc, st, len(ivl), len(mine), res,
                                                                          float(r2)))
    assert (rc == 0) == (st == Q.SUCCESS)
    assert len(ivl) == len(mine)
    assert np.array_equal(ivl, mine), "%s: interval lists differ" % name
    assert abs(res - float(r2)) <= 1e-13 * max(abs(res), 1e-300)
    return rc, len(ivl)


def test_smooth_peaked_oscillatory_and_kinked_integrands():
    cases = [
        ("exp(x) cos(3x)", lambda x: math.exp(x) * math.cos(3 * x), lambda x: mpmath.exp(x) * mpmath.cos(3 * x),
         0.0, 5.0, 0.0, 1e-10),
        ("peak 1/(1e-4+(x-0.3)^2)", lambda x: 1 / (1e-4 + (x - 0.3) ** 2), lambda x: 1 / (mpmath.mpf("1e-4") + (x - mpmath.mpf(0.3)) ** 2),
         0.0, 1.0, 0.0, 1e-12),
        ("x sin(50x)", lambda x: x * math.sin(50 * x), lambda x: x * mpmath.sin(50 * x), 0.0, 3.0, 0.0, 1e-9),
        ("sqrt|x-1/3|", lambda x: math.sqrt(abs(x - 1 / 3)), lambda x: mpmath.sqrt(abs(x - mpmath.mpf(1 / 3))),
         0.0, 1.0, 0.0, 1e-13),
        ("|x-1/3|^0.2", lambda x: abs(x - 1 / 3) ** 0.2, lambda x: abs(x - mpmath.mpf(1 / 3)) ** mpmath.mpf(0.2),
         0.0, 1.0, 0.0, 1e-12),
    ]
    total = 0
    for name, fd, fm, a, b, ea, er in cases:
        _, n = compare(name, fd, fm, a, b, ea, er)
        total += n
    assert total > 60


@pytest.mark.parametrize("nknots,noise,epsrel,fails", [(300, 0.1, 1e-6, False), (600, 0.3, 1e-8, True)])
def test_noisy_spline(nknots, noise, epsrel, fails):
    """A natural cubic spline through noisy knots times a smooth kernel (the
    shape of a noisy delta_tot history): 109 intervals to converge, and a case
    that exhausts the 200-interval limit (GSL_EMAXITER, which the reference's
    error handler turns into terminate(2001)); past limit/2 + 2 intervals
    QUADPACK keeps only part of the error list ordered, and the lists still
    agree entry by entry."""
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(2)
    x = np.linspace(0.0, 4.0, nknots)
    y = np.exp(-x) * (1 + noise * rng.standard_normal(nknots))
    sd = CubicSpline(x, y, bc_type="natural")
    sm = Q.natural_cspline(x, y)
    rc, n = compare("noisy spline %d" % nknots, lambda t: float(sd(t)) * math.cos(t),
                    lambda t: sm(t) * mpmath.cos(t), 0.0, 4.0, 0.0, epsrel)
    assert (rc == Q.EMAXITER and n == 200) if fails else rc == 0


@pytest.mark.parametrize("case", [
    dict(name="C1-like Na=40", m=0.15, na=40, nk=6, amax=0.5, noise=0.0, omega0=0.3),
    dict(name="noisy Na=120", m=0.3, na=120, nk=4, amax=1.0, noise=0.1, omega0=0.3),
    dict(name="Omega0=0.25 Na=60", m=0.5, na=60, nk=4, amax=0.8, noise=0.02, omega0=0.25),
], ids=lambda c: c["name"])
def test_get_delta_nu_integrand_subdivisions(case):
    """The oracle's recorded subdivisions of get_delta_nu (per k) against the
    restated QAG on the same integrand assembled at 40 digits: natural cubic
    splines through the oracle's fslength nodes (16 Na points, each an oracle
    fslength) and through the delta_tot row, 1/(a H(a)) with the oracle's H,
    and the specialJ fit (delta_tot_table.c:417-428, :492-500)."""
    c = Cosmo([case["m"], 0.0, 0.0], omega0=case["omega0"])
    logs, rows, k, init = synthetic_history(case["na"], case["nk"], case["amax"], noise=case["noise"])
    T = c.oracle(case["nk"], logs, rows, init)
    a = case["amax"]
    rc, out, oivl, onint = T.get_delta_nu_record(a, k, case["m"])
    assert rc == 0
    L = O.lib()
    T0, light = T.d.TimeTransfer, T.d.light
    mnubykT = case["m"] / c.P.oom.kBtnu
    x0, xa = math.log(T0), math.log(a)
    nfs = 16 * case["na"]
    fx = [x0 + i * (xa - x0) / (nfs - 1.0) for i in range(nfs)]
    fy = [L.orc_fslength(v, xa, light) for v in fx]
    q = Q.QAG()
    with mpmath.workdps(q.dps):
        fsl = Q.natural_cspline(fx, fy)

        def specj(x):
            if x <= 0:
                return mpmath.mpf(1)
            x2 = x * x
            x4 = x2 * x2
            return (1 + mpmath.mpf("0.0168") * x2 + mpmath.mpf("0.0407") * x4) / (
                1 + mpmath.mpf("2.1734") * x2 + mpmath.mpf("1.6787") * mpmath.exp(mpmath.mpf("4.1811") * mpmath.log(x))
                + mpmath.mpf("0.1467") * x4 * x4)

        worst, nints = 0.0, 0
        for ik in range(case["nk"]):
            dt = Q.natural_cspline(logs, rows[ik])
            kk = k[ik]

            def f(x):
                fl = fsl(x)
                ai = mpmath.exp(x)
                return fl / (ai * L.orc_hubble(float(ai))) * specj(kk * fl / mnubykT) * dt(x)

            st, res, err, w = q.integrate(f, x0, xa, 0.0, 1e-6)
            check_invariants(w, st, err, 0.0, 1e-6, res)
            n = int(onint[ik])
            assert st == Q.SUCCESS and w.size() == n
            assert np.array_equal(oivl[ik, :n], restated_lists(w)), "k %d: interval lists differ" % ik
            nints += n
    print("  %s: %d k, %d intervals identical" % (case["name"], case["nk"], nints))
    c.free()
