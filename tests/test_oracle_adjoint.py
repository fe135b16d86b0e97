"""Pins of the oracle's adjoint of get_delta_nu (CPU).

The reference has no adjoint, so the derivative is defined on the computed
quadrature: IC + prefac * (sum of GK61 over the forward's final QAG
subdivision), subdivision held fixed.  orc_dnu_frozen evaluates exactly that
function; it is linear in delta_tot and delta_nu_init and smooth in mnu, so
central differences pin the adjoint tightly.  The oracle's adjoint is built
from a brute-force spline basis and complex-step differentiation, the GPU's
from hand-derived formulas; these pins make the oracle the trusted side.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_py as O
from conftest import UNIT_LENGTH, UNIT_TIME


def synthetic(Na, nk, amax, seed=0, noise=0.05):
    rng = np.random.default_rng(seed)
    a = np.exp(np.linspace(math.log(0.01), math.log(amax), Na))
    logs = np.log(a)
    logs[-1] = math.log(amax)
    k = np.exp(np.linspace(math.log(2e-3), math.log(5.0), nk)) * 1e-3  # h/kpc
    pk = 1.0 / (1 + (k / 2e-4) ** 2)
    rows = a[None, :] * np.sqrt(pk)[:, None] * (1 + noise * rng.standard_normal((nk, Na)))
    init = 0.01 * np.sqrt(pk) * 0.9
    return logs, rows, k, init


def make_table(masses, Na, nk, amax, hybrid=None, seed=0, noise=0.05):
    om = O.omega_nu(masses)
    if hybrid:
        vcrit, tcrit = hybrid
        O.lib().orc_init_hybrid_nu(C.byref(om.hybnu), (C.c_double * 3)(*masses), vcrit, 2.99792458e10 / 1e5, tcrit,
                                   om.kBtnu)
    O.lib().orc_init_hubble(C.byref(om), 0.3, UNIT_TIME)
    T = O.Table(om, nk, 0.01, 1.0, 0.3, UNIT_TIME, UNIT_LENGTH)
    logs, rows, k, init = synthetic(Na, nk, amax, seed=seed, noise=noise)
    T.scalefact[:Na] = logs
    T.delta_tot[:, :Na] = rows
    T.delta_nu_init[:] = init
    T.d.ia = Na
    return T, k


CASES = [
    dict(name="cspline", masses=[0.15, 0, 0], Na=40, nk=6, amax=0.5, hybrid=None),
    dict(name="linear Na=2", masses=[0.2, 0, 0], Na=2, nk=4, amax=0.03, hybrid=None),
    dict(name="hybrid", masses=[0.3, 0.06, 0.06], Na=30, nk=5, amax=0.9, hybrid=(500.0, 0.3333)),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_recorded_and_frozen_forward_equal_qag_forward(case):
    T, k = make_table(case["masses"], case["Na"], case["nk"], case["amax"], case["hybrid"])
    m, a = case["masses"][0], case["amax"]
    rc0, ref = T.get_delta_nu(a, k, m)
    rc1, rec, ivl, nint = T.get_delta_nu_record(a, k, m)
    rc2, frz = T.dnu_frozen(a, k, m, ivl, nint)
    assert rc0 == rc1 == rc2 == 0
    assert np.array_equal(ref, rec) and np.array_equal(rec, frz)
    # the recorded subintervals tile [log T0, log a]
    for i in range(case["nk"]):
        iv = ivl[i, :nint[i]]
        iv = iv[np.argsort(iv[:, 0])]
        assert iv[0, 0] == math.log(0.01) and iv[-1, 1] == math.log(a)
        assert np.array_equal(iv[1:, 0], iv[:-1, 1])


def _round(L, h):
    """Rounding error of a central difference of a value of size L with step h."""
    return 8 * np.finfo(float).eps * abs(L) / (2 * h)


def _loss(T, a, k, m, ivl, nint, g):
    rc, out = T.dnu_frozen(a, k, m, ivl, nint)
    assert rc == 0
    return float(np.dot(g, out))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_adjoint_matches_central_differences(case):
    T, k = make_table(case["masses"], case["Na"], case["nk"], case["amax"], case["hybrid"])
    m, a, Na, nk = case["masses"][0], case["amax"], case["Na"], case["nk"]
    g = np.random.default_rng(1).uniform(0.5, 1.5, nk)
    rc, out, db, ib, mb = T.get_delta_nu_adjoint(a, k, m, g)
    assert rc == 0
    _, _, ivl, nint = T.get_delta_nu_record(a, k, m)
    rng = np.random.default_rng(2)
    # delta_tot entries: the value is linear in them, differences are exact up to rounding
    picks = [(int(i), int(j)) for i, j in zip(rng.integers(0, nk, 8), rng.integers(0, Na, 8))]
    picks += [(0, 0), (nk - 1, Na - 1)]
    for i, j in picks:
        y0 = T.delta_tot[i, j]
        h = 1e-3 * np.max(np.abs(T.delta_tot[i, :Na]))
        T.delta_tot[i, j] = y0 + h
        lp = _loss(T, a, k, m, ivl, nint, g)
        T.delta_tot[i, j] = y0 - h
        lm = _loss(T, a, k, m, ivl, nint, g)
        T.delta_tot[i, j] = y0
        fd = (lp - lm) / (2 * h)
        scale = np.max(np.abs(db[i]))
        assert abs(fd - db[i, j]) <= 1e-7 * scale + _round(lp, h), (i, j, fd, db[i, j])
    # a constant added to a row: the spline reproduces constants, so the sum of the row's weights
    for i in range(nk):
        h = 1e-3 * np.max(np.abs(T.delta_tot[i, :Na]))
        row = T.delta_tot[i, :Na].copy()
        T.delta_tot[i, :Na] = row + h
        lp = _loss(T, a, k, m, ivl, nint, g)
        T.delta_tot[i, :Na] = row - h
        lm = _loss(T, a, k, m, ivl, nint, g)
        T.delta_tot[i, :Na] = row
        assert abs((lp - lm) / (2 * h) - db[i].sum()) <= 1e-7 * np.abs(db[i]).sum() + _round(lp, h)
    # delta_nu_init (linear)
    for i in range(nk):
        v0 = T.delta_nu_init[i]
        h = 0.5 * abs(v0)
        T.delta_nu_init[i] = v0 + h
        lp = _loss(T, a, k, m, ivl, nint, g)
        T.delta_nu_init[i] = v0 - h
        lm = _loss(T, a, k, m, ivl, nint, g)
        T.delta_nu_init[i] = v0
        assert abs((lp - lm) / (2 * h) - ib[i]) <= 1e-8 * abs(ib[i]) + _round(lp, h)
    # mnu: Richardson-extrapolated central differences of the frozen value
    def D(h):
        return (_loss(T, a, k, m + h, ivl, nint, g) - _loss(T, a, k, m - h, ivl, nint, g)) / (2 * h)
    h = 1e-3 * m
    rich = (4 * D(h / 2) - D(h)) / 3
    print("  %s: mnu_bar %.12e, Richardson FD %.12e" % (case["name"], mb, rich))
    assert abs(rich - mb) <= 1e-8 * abs(mb)


def test_frozen_derivative_tracks_the_qag_derivative():
    """With the QAG free to re-decide, differences of the reference computation
    itself agree with the frozen-subdivision derivative to the quadrature
    tolerance (relerr 1e-6 of the value over the step)."""
    T, k = make_table([0.15, 0, 0], 40, 6, 0.5)
    g = np.ones(6)
    rc, out, db, ib, mb = T.get_delta_nu_adjoint(0.5, k, 0.15, g)
    h = 1e-3 * 0.15
    lp = float(np.sum(T.get_delta_nu(0.5, k, 0.15 + h)[1]))
    lm = float(np.sum(T.get_delta_nu(0.5, k, 0.15 - h)[1]))
    fd = (lp - lm) / (2 * h)
    print("  d/dmu: frozen %.10e, QAG differences %.10e" % (mb, fd))
    assert abs(fd - mb) <= 1e-3 * abs(mb)


@pytest.mark.parametrize("hybrid", [0, 1])
def test_specialJ_complex_step_matches_differences(hybrid):
    """The complex restatement of specialJ differentiates the real one: compare
    its complex-step d/dmu with Richardson differences of orc_specialJ."""
    L = O.lib()
    vcrit, nfl = (500.0 / 2.99792458e5, 0.05) if hybrid else (0.0, 0.0)  # 500 km/s: qc = vcrit mu ~ 0.1..3
    for kfsl in [1e-3, 0.3, 2.0, 7.5, 40.0]:
        for mu in [50.0, 500.0, 2000.0]:
            def J(m):
                return L.orc_specialJ(kfsl / m, vcrit * m if hybrid else 0.0, nfl)
            def D(h):
                return (J(mu + h) - J(mu - h)) / (2 * h)
            h = 1e-3 * mu
            rich = (4 * D(h / 2) - D(h)) / 3
            cs = L.orc_specialJ_dmu(kfsl, mu, vcrit, hybrid, nfl)
            assert abs(cs - rich) <= 1e-7 * abs(cs) + 2 * _round(J(mu), h / 2), (kfsl, mu, cs, rich)
