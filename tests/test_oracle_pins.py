"""Pin the oracle (CPU restatement) before trusting it.

Every check here is independent of the oracle's own formulas: the reference
test suite's known answers (Mathematica values, fixtures, invariants),
closed forms, brute force and mpmath.  A mistake planted in the oracle (a
dropped term, a wrong sign or index) fails at least one of these.
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import oracle_py as O
from conftest import GOLDEN, REFDATA, ROOT, T_CMB0, UNIT_LENGTH, UNIT_TIME

mp = pytest.importorskip("mpmath")
FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


def qag(f, a, b, epsrel=1e-10, limit=200):
    L = O.lib()
    L.orc_qag61.argtypes = [FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                            C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    r, e, n = C.c_double(), C.c_double(), C.c_int()
    cb = FN(lambda x, p: f(x))
    rc = L.orc_qag61(cb, None, a, b, 0.0, epsrel, limit, C.byref(r), C.byref(e), C.byref(n))
    return rc, r.value, e.value, n.value


def table(name):
    s = open(os.path.join(ROOT, "oracle", "quad_tables.h")).read()
    m = re.search(name + r"\[\d+\] = \{(.*?)\};", s, re.S)
    return np.array([float.fromhex(t.split("/*")[0].strip()) for t in m.group(1).split(",") if t.strip()])


# ---------------------------------------------------------------- quadrature

def test_gk61_published_digits():
    xgk, wgk, wg = table("orc_xgk61"), table("orc_wgk61"), table("orc_wg30")
    # QUADPACK dqk61 (Piessens et al. 1983) values
    assert xgk[0] == pytest.approx(0.999484410050490637571325895705811, rel=0, abs=1e-16)
    assert wgk[30] == pytest.approx(0.051494729429451567558340433647099, rel=0, abs=1e-17)
    assert wg[14] == pytest.approx(0.102852652893558840341285636705415, rel=0, abs=1e-17)
    assert xgk[30] == 0.0 and np.all(np.diff(xgk) < 0)


def test_gk61_polynomial_exactness():
    xgk, wgk, wg = table("orc_xgk61"), table("orc_wgk61"), table("orc_wg30")
    for d in range(0, 92, 2):
        k = wgk[30] * (1.0 if d == 0 else 0.0) + np.sum(wgk[:30] * 2 * xgk[:30] ** d)
        assert k == pytest.approx(2.0 / (d + 1), rel=1e-13), d
    gx = xgk[1::2]
    for d in range(0, 60, 2):
        g = np.sum(wg * 2 * gx ** d)
        assert g == pytest.approx(2.0 / (d + 1), rel=1e-13), d
    # the two rules differ on a non-polynomial (the QAG error estimate depends on it)
    f = lambda x: 1.0 / (1.0 + 25 * x * x)
    kr = wgk[30] * f(0.0) + np.sum(wgk[:30] * (f(xgk[:30]) + f(-xgk[:30])))
    ga = np.sum(wg * (f(gx) + f(-gx)))
    exact = 2 * math.atan(5.0) / 5
    assert abs(kr - exact) < 1e-9 and abs(ga - exact) > 1e-8


def test_qag_closed_forms():
    rc, r, e, n = qag(math.sin, 0.0, math.pi)
    assert rc == 0 and r == pytest.approx(2.0, rel=1e-14) and n == 1
    # sharply peaked: needs bisection
    eps = 1e-3
    f = lambda x: 1.0 / ((x - 0.3) ** 2 + eps ** 2)
    exact = (math.atan(0.7 / eps) + math.atan(0.3 / eps)) / eps
    rc, r, e, n = qag(f, 0.0, 1.0, epsrel=1e-10)
    assert rc == 0 and n > 4 and r == pytest.approx(exact, rel=1e-10)
    # integrable singularity at 0: 200 subdivisions exhaust QAG (GSL_EMAXITER=11 or roundoff 18)
    rc, r, e, n = qag(lambda x: x ** -0.9 if x > 0 else 0.0, 0.0, 1.0, epsrel=1e-12)
    assert rc in (11, 18, 21)


def test_natural_spline_vs_scipy():
    interp = pytest.importorskip("scipy.interpolate")
    rng = np.random.default_rng(4)
    x = np.sort(rng.uniform(0, 3, 40))
    y = np.sin(3 * x) + 0.1 * rng.standard_normal(40)
    s = O.orc_spline = None
    L = O.lib()
    sp = O.lib  # noqa: F841
    ref = interp.CubicSpline(x, y, bc_type="natural")
    # evaluate the oracle spline through get_delta_nu's building block via a tiny ctypes shim
    xs = np.ascontiguousarray(x)
    ys = np.ascontiguousarray(y)

    class Spl(C.Structure):
        _fields_ = [("n", C.c_int), ("linear", C.c_int), ("x", O.dp), ("y", O.dp), ("c", O.dp)]
    L.orc_spline_init.argtypes = [C.POINTER(Spl), O.dp, O.dp, C.c_int, C.c_int]
    L.orc_spline_eval.argtypes = [C.POINTER(Spl), C.c_double, C.POINTER(C.c_int)]
    L.orc_spline_eval.restype = C.c_double
    L.orc_spline_free.argtypes = [C.POINTER(Spl)]
    spl = Spl()
    L.orc_spline_init(C.byref(spl), O.ptr(xs), O.ptr(ys), len(xs), 0)
    err = C.c_int(0)
    for xv in np.linspace(x[0], x[-1], 301):
        assert L.orc_spline_eval(C.byref(spl), xv, C.byref(err)) == pytest.approx(float(ref(xv)), rel=1e-11, abs=1e-12)
    assert err.value == 0
    L.orc_spline_eval(C.byref(spl), x[-1] + 1e-9, C.byref(err))
    assert err.value == 2001           # outside the table: GSL_EDOM -> terminate(2001)
    L.orc_spline_free(C.byref(spl))


# ---------------------------------------------------------- special functions

def test_specialJ_known_answers(orc):
    # delta_tot_table_test.c:159-169 (Mathematica)
    assert orc.orc_specialJ(0, -1, 0) == 1
    for x, v in [(1, 0.2117), (2, 0.0223807), (0.5, 0.614729), (0.3, 0.829763)]:
        assert abs(orc.orc_specialJ(x, -1, 0) - v) < 1e-3
    assert abs(orc.orc_specialJ(0, 1, 0) - 0.940437) < 1e-4
    assert abs(orc.orc_specialJ(0.5, 1e-2, 0.5) - 0.614729 / 0.5) < 1e-3
    assert abs(orc.orc_specialJ(0.5, 1, 0.5) - 0.556557 / 0.5) < 1e-4
    assert abs(orc.orc_specialJ(1, 0.1, 0.5) - 0.211662 / 0.5) < 1e-4


def _J_exact(x, qc=0.0):
    mp.mp.dps = 30
    f = lambda q: (mp.sin(q * x) / (q * x) if x > 0 else 1) * q * q / (mp.exp(q) + 1)
    return float(mp.quad(f, [qc, 5, 15, 40, 80]) / (1.5 * mp.zeta(3)))


def test_specialJ_against_integral(orc):
    # fit: 3% relative and 7e-4 absolute of the exact J (delta_tot_table.c:409-416)
    for x in [0.05, 0.2, 0.7, 1.5, 3.0, 6.0]:
        ex = _J_exact(x)
        fit = orc.orc_specialJ(x, -1, 0)
        assert abs(fit - ex) < max(0.03 * abs(ex), 7e-4), x
    # 19-term truncated series vs the truncated integral normalised by (1 - nufrac_low(qc))
    for x, qc in [(0.5, 1.0), (1.0, 0.1), (2.0, 0.5), (0.0, 1.0)]:
        nfl = orc.orc_nufrac_low(qc)
        ex = _J_exact(x, qc) / (1 - nfl)
        assert abs(orc.orc_specialJ(x, qc, nfl) - ex) < 1e-4, (x, qc)


def test_bessel_j0(orc):
    for x in [0.0, 1e-8, 0.1, 0.49999, 0.5, 1.0, 7.3]:
        want = 1.0 if x == 0 else math.sin(x) / x
        assert orc.orc_bessel_j0(x) == pytest.approx(want, rel=3e-16, abs=1e-17)


# -------------------------------------------------------------- background

@pytest.fixture(scope="module")
def om_test(orc):
    om = O.omega_nu([0.15, 0.15, 0.15])
    orc.orc_init_hubble(C.byref(om), 0.2793, UNIT_TIME)
    return om


def test_fslength_known_answers(orc, om_test):
    kT = 8.61734e-5 * (4 / 11.) ** (1 / 3.) * 1.00328 * T_CMB0
    # delta_tot_table_test.c:183-191
    assert abs(orc.orc_fslength(math.log(0.5), math.log(1), 299792.) / 1272.92 / (0.45 / kT) - 1) < 1e-5
    assert abs(orc.orc_fslength(math.log(0.1), math.log(0.5), 299792.) / 5427.8 / (0.6 / kT) - 1) < 1e-5
    assert orc.orc_fslength(0.0, -1.0, 1.0) == 0.0


def test_fslength_additive_and_mpmath(orc, om_test):
    x0, x1, x2 = math.log(0.01), math.log(0.2), math.log(0.9)
    f01, f12, f02 = (orc.orc_fslength(a, b, 1.0) for a, b in [(x0, x1), (x1, x2), (x0, x2)])
    # one accepted 61-point estimate per call: accurate to ~1e-10, far inside QAG's 1e-6
    assert f01 + f12 == pytest.approx(f02, rel=1e-9)
    mp.mp.dps = 25
    g = lambda x: 1 / (mp.e ** (2 * x) * orc.orc_hubble(float(mp.e ** x)))
    exact = float(mp.quad(g, [x0, x1, x2]))
    assert f02 == pytest.approx(exact, rel=1e-9)
    assert f01 == pytest.approx(float(mp.quad(g, [x0, x1])), rel=1e-9)


def test_omega_nu_reference_checks(orc):
    # omega_nu_single_test.c:37-65
    om = O.omega_nu([0.5, 0.5, 0.0])
    assert abs(om.rhocrit - 1.8784e-29 * 0.49) < 5e-5 * om.rhocrit
    z0 = orc.orc_omega_nu_single(C.byref(om), 1, 0)
    assert abs(z0 / 0.5 ** 3 - orc.orc_omega_nu_single(C.byref(om), 0.5, 0)) < 5e-5 * z0
    assert z0 / 0.01 ** 3 < orc.orc_omega_nu_single(C.byref(om), 0.01, 0)
    assert abs(z0 - 0.5 / 93.14 / 0.49) < 1e-4 * z0
    assert orc.orc_omega_nu_single(C.byref(om), 0.1, 1) == orc.orc_omega_nu_single(C.byref(om), 0.1, 0)
    m0 = orc.orc_omega_nu_single(C.byref(om), 1, 2)
    assert m0 / 0.5 ** 4 == orc.orc_omega_nu_single(C.byref(om), 0.5, 2)
    assert orc.orc_omega_nu_single(C.byref(om), 1e-4, 0) > z0 / 1e-12
    # degeneracies (omega_nu_single_test.c:112-136)
    assert list(O.omega_nu([0.2, 0.2, 0.2]).deg) == [3, 0, 0]
    assert list(O.omega_nu([0.2, 0.1, 0.3]).deg) == [1, 1, 1]
    # sum of singles (:138-149) and photons (:151-160)
    om3 = O.omega_nu([0.2, 0.1, 0.3])
    tot = sum(orc.orc_omega_nu_single(C.byref(om3), 0.5, i) for i in range(3))
    assert abs(orc.orc_get_omega_nu(C.byref(om3), 0.5) - tot) < 1e-6 * tot
    omegar = (4 * 5.670373e-5 * 8 * math.pi * 6.67408e-8 / (3 * 2.99792458e10 ** 3 * 3.24077929e-18 ** 2 * 0.49)
              * T_CMB0 ** 4)
    assert abs(orc.orc_get_omegag(C.byref(om3), 0.5) / (omegar / 0.5 ** 4) - 1) < 1e-6


def test_omega_nu_table_vs_exact(orc):
    # omega_nu_single_test.c:91-110: interpolated table within 1e-6 of direct integration
    om = O.omega_nu([0.05, 0.05, 0.05])
    for i in range(1, 123, 7):
        a = 0.01 + i / 123.
        tab = orc.orc_omega_nu_single(C.byref(om), a, 0)
        mp.mp.dps = 20
        kT = om.kBtnu
        ex = mp.quad(lambda q: q * q * mp.sqrt(q * q + (0.05 * a) ** 2) / (mp.exp(q / kT) + 1), [0, 20 * kT, 500 * kT])
        conv = 4 * math.pi * 2 * (1. / (2 * math.pi * 2.99792458e10 * 6.582119e-16)) ** 3 * (1.60217646e-12 / 2.99792458e10 ** 2)
        exact = float(ex) * conv / a ** 4 / om.rhocrit
        assert abs(1 - exact / tab) < 1e-6


def test_nufrac_and_hybrid(orc):
    # omega_nu_single_test.c:163-187
    assert orc.orc_nufrac_low(0) == 0
    assert abs(orc.orc_nufrac_low(1) / 0.0595634 - 1) < 1e-5
    assert abs(orc.orc_nufrac_low(0.5) / 0.00941738 - 1) < 1e-5
    om = O.omega_nu([0.2, 0.2, 0.2])
    orc.orc_init_hybrid_nu(C.byref(om.hybnu), (C.c_double * 3)(0.2, 0.2, 0.2), 700, 299792, 0.5, om.kBtnu)
    part = orc.orc_nufrac_low(700 / 299792. * 0.2 / om.kBtnu)
    assert abs(orc.orc_particle_nu_fraction(C.byref(om.hybnu), 0.50001, 0) / part - 1) < 1e-5
    assert orc.orc_particle_nu_fraction(C.byref(om.hybnu), 0.49999, 0) == 0
    r = orc.orc_get_omega_nu_nopart(C.byref(om), 0.499999) * (1 - part) / orc.orc_get_omega_nu_nopart(C.byref(om), 0.500001)
    assert abs(r - 1) < 1e-4


def test_transfer_reader():
    # transfer_init_test.c:18-34
    f = os.path.join(REFDATA, "ics_transfer_99.dat")
    rc, t = O.load_transfer(f, 1e9, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    assert rc == 0 and t.NPowerTable == 336
    rc, t = O.load_transfer(f, 512000, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    assert t.NPowerTable == 271 and t.logk[0] > math.log(math.pi / 512000)
    rc, t2 = O.load_transfer(f, 512, 1e3 * UNIT_LENGTH, UNIT_LENGTH * 1e3)
    assert t2.logk[0] == t.logk[0] + math.log(1e3) and t2.T_nu[0] == t.T_nu[0]
    assert abs(t.T_nu[0] - 0.508479) < 1e-6 and abs(t.T_nu[30] - 0.0122563) < 1e-6


# ------------------------------------------------------------------ binning

def _ref_4cubed_field():
    # powerspectrum_test.c:21-35: ones for x in {0,1}, field[0] = 2, FFTW forward r2c
    f = np.zeros((4, 4, 4))
    f[0:2, :, :] = 1
    f[0, 0, 0] = 2
    return np.fft.rfftn(f)


def test_powerspectrum_4cubed_known_answer():
    n, p, c, k = O.total_powerspectrum(_ref_4cubed_field(), 15)
    assert n == 9
    assert abs(k[2] - 1.73205) < 1e-5
    assert c[1] == 12 and c[0] == 6 and c[n - 1] == 1
    # powerspectrum_test.c:43-45 tolerances (the fixture came from a single-precision FFT)
    assert abs(p[0] - 0.254834) < 1e-5 * 0.04
    assert abs(p[1] - 0.00212722) < 1e-5 * 0.005
    assert abs(p[2] - 0.00323766) < 1e-5 * 0.003


def _brute_bins(F, nrbins):
    """Independent numpy statement of the binning (vectorised, exact integer k^2)."""
    N = F.shape[1]
    kv = np.where(np.arange(N) <= N // 2, np.arange(N), np.arange(N) - N)
    kx, ky, kz = np.meshgrid(kv, kv, np.arange(N // 2 + 1), indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    mult = np.where((kz == 0) | (kz == N // 2), 1, 2)
    iw = np.where(kv == 0, 1.0, np.pi * kv / (N * np.sin(np.pi * kv / N)))
    w4 = (iw[kx % N] * iw[ky % N] * iw[kz]) ** 4
    bpu = (nrbins - 1) / math.log(math.sqrt(3) * N / 2.0)
    valid = k2 > 0
    b = np.floor(bpu * np.log(np.sqrt(k2[valid].astype(float)))).astype(int)
    cnt = np.bincount(b, weights=mult[valid], minlength=nrbins)
    pw = np.bincount(b, weights=(mult * np.abs(F) ** 2 * w4)[valid], minlength=nrbins)
    ke = np.bincount(b, weights=(mult * np.sqrt(k2))[valid], minlength=nrbins)
    m2 = abs(F[0, 0, 0]) ** 2
    nz = cnt > 0
    return nz.sum(), pw[nz] / m2 / cnt[nz], cnt[nz].astype(np.int64), ke[nz] / cnt[nz]


@pytest.mark.parametrize("N,nrbins", [(8, 7), (16, 15), (16, 64), (32, 33)])
def test_powerspectrum_brute_force(N, nrbins):
    rng = np.random.default_rng(N + nrbins)
    F = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    F[0, 0, 0] = N ** 3
    n, p, c, k = O.total_powerspectrum(F, nrbins)
    bn, bp, bc, bk = _brute_bins(F, nrbins)
    assert n == bn
    np.testing.assert_array_equal(c[:n], bc)
    np.testing.assert_allclose(p[:n], bp, rtol=1e-12)
    np.testing.assert_allclose(k[:n], bk, rtol=1e-13)


def test_powerspectrum_single_mode_and_slabs():
    N = 16
    F = np.zeros((N, N, N // 2 + 1), complex)
    F[0, 0, 0] = 1.0
    F[3, 2, 5] = 2.0 + 1.0j
    n, p, c, k = O.total_powerspectrum(F, 10)
    hit = np.nonzero(p[:n])[0]
    assert len(hit) == 1
    iw = lambda q: math.pi * q / (N * math.sin(math.pi * q / N))
    w4 = (iw(3) * iw(2) * iw(5)) ** 4
    assert p[hit[0]] * c[hit[0]] == pytest.approx(2 * 5.0 * w4, rel=1e-14)
    # slab decomposition: partial sums over [0,6) + [6,16) equal the whole (before normalisation)
    _, pa, ca, _ = O.total_powerspectrum(F[:6], 10, 0, 6)
    _, pb, cb, _ = O.total_powerspectrum(F[6:], 10, 6, 10)
    assert ca[:n].sum() + cb.sum() == c[:n].sum()


def test_powerspectrum_adjoint_central_differences():
    """dL/dF of L = sum_n g_n P_n: P is quadratic in each F_m over |F_0|^2, so
    central differences in Re/Im of single entries are exact up to rounding.
    Entries: the zero mode, the kz = 0 and kz = N/2 planes, interior modes."""
    N, nrbins = 8, 12
    rng = np.random.default_rng(3)
    F = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    F[0, 0, 0] = N ** 3 * (1 + 0.3j)
    n, p, c, k = O.total_powerspectrum(F, nrbins)
    g = rng.uniform(-1, 1, n)
    rc, Fb = O.bin_power_adjoint(F, nrbins, g)
    assert rc == 0
    L = lambda G: float(np.dot(g, O.total_powerspectrum(G, nrbins)[1][:n]))
    for idx in [(0, 0, 0), (0, 1, 0), (3, 5, 0), (2, 7, N // 2), (1, 2, 3), (7, 7, 1), (4, 0, 2)]:
        for part, unit in [("re", 1.0), ("im", 1j)]:
            h = 1e-3 * max(1.0, abs(F[idx]))

            def D(h):
                Fp, Fm = F.copy(), F.copy()
                Fp[idx] += h * unit
                Fm[idx] -= h * unit
                return (L(Fp) - L(Fm)) / (2 * h)
            # exact for the quadratic modes; the zero mode enters through 1/|F_0|^2: Richardson
            fd = D(h) if idx != (0, 0, 0) else (4 * D(h / 2) - D(h)) / 3
            an = Fb[idx].real if part == "re" else Fb[idx].imag
            assert fd == pytest.approx(an, rel=1e-7, abs=1e-12 * abs(L(F)) / h), (idx, part)
    # P is homogeneous of degree 0 in F: sum_m Re(conj(F_m) dL/dF_m) = 0, the
    # zero mode balancing the others (-2 sum g P against +2 sum g P)
    dots = (F.conj() * Fb).real
    assert dots[0, 0, 0] == pytest.approx(-2 * np.dot(g, p[:n]), rel=1e-12)
    assert abs(dots.sum()) < 1e-12 * abs(dots[0, 0, 0])


# ----------------------------------------------------------------- delta_pow

def _read_powerspec_fixtures():
    with open(os.path.join(REFDATA, "powerspec_nu_004.txt")) as fh:
        toks = fh.read().split()
    scale, nb = float(toks[0]), int(toks[1])
    vals = np.array([float(t) for t in toks[2:2 + 2 * nb]]).reshape(nb, 2)
    with open(os.path.join(REFDATA, "powerspec_cdm_004.txt")) as fh:
        toks2 = fh.read().split()
    assert int(toks2[1]) == nb
    cdm = np.array([float(t) for t in toks2[2:2 + nb]])
    return scale, vals[:, 0].copy(), np.sqrt(vals[:, 1]), cdm


def test_delta_pow_reference_checks(orc):
    # delta_pow_test.c:89-112 with its fixture setup (:116-181)
    _, k, dnu, dcdm = _read_powerspec_fixtures()
    logkk, ratio = np.log(k), dnu / dcdm
    for i in range(10):
        assert O.get_dnudcdm(logkk, ratio, 1.0, logkk[5 * i]) == ratio[5 * i]
    for i in range(50):
        lk = (logkk[5 * i + 1] + logkk[5 * i + 2]) / 2
        pk = (ratio[5 * i + 1] + ratio[5 * i + 2]) / 2
        assert abs(O.get_dnudcdm(logkk, ratio, 1.0, lk) - pk) < 5e-3 * pk
    assert abs(O.get_dnudcdm(logkk, ratio, 1.0, logkk[0] - 0.01) - ratio[0]) < 5e-3 * ratio[0]
    assert abs(O.get_dnudcdm(logkk, ratio, 1.0, logkk[-1] + 0.01) - ratio[-1]) < 5e-3 * ratio[-1]
    # closed form of linear interpolation and the norm factor
    x = 0.3 * logkk[7] + 0.7 * logkk[8]
    want = ratio[7] + (x - logkk[7]) / (logkk[8] - logkk[7]) * (ratio[8] - ratio[7])
    assert O.get_dnudcdm(logkk, ratio, 2.5, x) == pytest.approx(2.5 * want, rel=1e-15)


def test_rescale_properties():
    N = 16
    rng = np.random.default_rng(3)
    F = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    logkk = np.linspace(-8, -3, 20)
    rc, G = O.rescale(F, 512000.0, logkk, np.full(20, 0.2), 0.5)
    assert rc == 0
    assert G[0, 0, 0] == F[0, 0, 0]
    mask = np.ones(F.shape, bool)
    mask[0, 0, 0] = False
    np.testing.assert_allclose(G[mask], F[mask] * 1.1, rtol=1e-15)
    rc, _ = O.rescale(F, 512000.0, logkk, np.full(20, np.nan), 0.5)
    assert rc == 5


# -------------------------------------------------------- integrator state

def read_all_nu_state(tab, path):
    """Row reader of delta_tot_table.c:254-301 ('# log a' then nk values)."""
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                parts = line[1:].split()
                rows.append([float(v) for v in parts])
    for i, r in enumerate(rows):
        tab.scalefact[i] = r[0]
        tab.delta_tot[:len(r) - 1, i] = r[1:]
    tab.d.ia = len(rows)
    return len(rows)


@pytest.fixture(scope="module")
def dtt_state(orc):
    scale, k, dnu, dcdm = _read_powerspec_fixtures()
    om = O.omega_nu([0.15, 0.15, 0.15])
    na3 = orc.orc_get_omega_nu(C.byref(om), 0.01) * 0.01 ** 3
    oma = 0.2793 - orc.orc_get_omega_nu(C.byref(om), 1) + na3
    fnu = na3 / oma
    dcdm = (dcdm - fnu * dnu) / (1. - fnu)
    rc, t = O.load_transfer(os.path.join(REFDATA, "ics_transfer_99.dat"), 512000, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    orc.orc_init_hubble(C.byref(om), 0.2793, UNIT_TIME)
    return dict(k=k, dnu=dnu, dcdm=dcdm, om=om, t=t)


def test_delta_tot_init(orc, dtt_state):
    s = dtt_state
    T = O.Table(s["om"], len(s["k"]), 0.01, 1, 0.2793, UNIT_TIME, UNIT_LENGTH)
    assert T.delta_tot_init(s["k"], s["dcdm"], s["t"], 0.01) == 0
    assert T.d.ia == 1 and T.scalefact[0] == math.log(0.01)
    na3 = orc.orc_get_omega_nu(C.byref(s["om"]), 0.01) * 0.01 ** 3
    nu1 = orc.orc_get_omega_nu(C.byref(s["om"]), 1)
    assert np.all(T.delta_nu_init > 0) and np.all(T.delta_nu_last > 0)
    assert np.max(np.abs(T.delta_nu_last / T.delta_nu_init - 1)) < 1e-4
    before = [orc.orc_get_delta_tot(T.delta_nu_init[i], s["dcdm"][i], na3, T.d.Omeganonu, nu1, 0) for i in range(len(s["k"]))]
    assert np.max(np.abs(T.delta_tot[:, 0] / before - 1)) < 1e-4


def test_get_delta_nu_update_against_saved_power(orc, dtt_state):
    # delta_tot_table_test.c:193-228
    s = dtt_state
    T = O.Table(s["om"], len(s["k"]), 0.01, 1, 0.2793, UNIT_TIME, UNIT_LENGTH)
    assert read_all_nu_state(T, os.path.join(REFDATA, "delta_tot_nu.txt")) == 25
    assert T.delta_tot_init(s["k"], s["dcdm"], s["t"], 0.33333333) == 0
    assert math.log(0.3333333) - T.scalefact[T.d.ia - 1] > 1e-5
    rc, out = T.get_delta_nu_update(0.33333333, s["k"], s["dcdm"], s["t"])
    assert rc == 0 and T.d.ia == 25
    assert np.max(np.abs(out / s["dnu"] - 1)) < 1e-2
    T.d.ia -= 1
    rc, out = T.get_delta_nu_update(0.33333333, s["k"], s["dcdm"], s["t"])
    assert rc == 0 and T.d.ia == 25
    assert np.max(np.abs(out / s["dnu"] - 1)) < 3e-2


def load_camb(i):
    z = np.load(os.path.join(GOLDEN, "camb_linear_200.npz"))
    dcdm = np.sqrt((z["T_cdm"][i] / z["T_tot"][i]) ** 2 * z["P"][i] * 1000.0 ** 3)
    dnu = np.sqrt((z["T_nu"][i] / z["T_tot"][i]) ** 2 * z["P"][i] * 1000.0 ** 3)
    return z["scalefact"][i], z["k_pk"][i] / 1000.0, dcdm, dnu


def test_reproduce_linear_theory(orc, dtt_state):
    # delta_tot_table_test.c:315-363: 99 successive updates against CAMB
    om = dtt_state["om"]
    rc, t = O.load_transfer(os.path.join(REFDATA, "ics_transfer_0.01.dat"), 512000, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    _, k, dcdm, _ = load_camb(0)
    T = O.Table(om, 200, 0.01, 1, 0.2793, UNIT_TIME, UNIT_LENGTH)
    assert T.delta_tot_init(k, dcdm, t, 0.01) == 0
    acc = 0.05
    for i in range(99):
        if i == 8:
            acc = 2e-2
        if i == 22:
            acc = 1.2e-2
        a, k, dcdm, dnu_camb = load_camb(i)
        rc, dnu = T.get_delta_nu_update(a, k, dcdm, t)
        assert rc == 0 and T.d.ia == i + 1
        assert np.all(np.abs(dnu_camb - dnu) < acc * dnu), i


# -------------------------------------------------- get_delta_nu invariants

def _history(orc, om, Na, nk, amax=0.5, seed=0, mode=0):
    T = O.Table(om, nk, 0.01, 0.01 + Na / 100., 0.3, UNIT_TIME, UNIT_LENGTH, mode=mode)
    a = np.exp(np.linspace(math.log(0.01), math.log(amax), Na))
    a[-1] = amax
    T.scalefact[:Na] = [math.log(v) for v in a]
    k = np.exp(np.linspace(math.log(2 * math.pi / 512000), math.log(0.01), nk))
    rng = np.random.default_rng(seed)
    base = np.outer(np.sqrt(k * 1e3) / (1 + (k * 2e4) ** 2), a) * (1 + 0.05 * rng.standard_normal((nk, Na)))
    T.delta_tot[:, :Na] = base
    T.delta_nu_init[:] = 0.4 * base[:, 0]
    T.d.ia = Na
    return T, k, amax


@pytest.fixture(scope="module")
def om_dnu(orc):
    om = O.omega_nu([0.15, 0.0, 0.0])
    orc.orc_init_hubble(C.byref(om), 0.3, UNIT_TIME)
    return om


def test_get_delta_nu_special_cases(orc, om_dnu):
    T, k, a = _history(orc, om_dnu, 30, 8)
    fsl = orc.orc_fslength(math.log(0.01), math.log(a), T.d.light)
    dp = 0.01 * (orc.orc_hubble(0.01) / T.d.light) * 0.01
    ic = np.array([orc.orc_specialJ(kk * fsl / (0.15 / om_dnu.kBtnu), 0, 0) for kk in k]) * T.delta_nu_init * (1 + dp * fsl)
    rc, out0 = T.get_delta_nu(a, k, 0.0)          # m = 0: initial-condition piece with m/kT -> 1
    ic0 = np.array([orc.orc_specialJ(kk * fsl, 0, 0) for kk in k]) * T.delta_nu_init * (1 + dp * fsl)
    np.testing.assert_allclose(out0, ic0, rtol=1e-15)
    T.d.ia = 1                                    # a single stored step: IC piece only
    rc, out1 = T.get_delta_nu(a, k, 0.15)
    np.testing.assert_allclose(out1, ic, rtol=1e-15)


def test_get_delta_nu_linear_in_history(orc, om_dnu):
    T, k, a = _history(orc, om_dnu, 30, 6, mode=1)
    init = T.delta_nu_init.copy()
    hist = T.delta_tot[:, :30].copy()
    _, full = T.get_delta_nu(a, k, 0.15)
    T.delta_nu_init[:] = 0
    _, part_hist = T.get_delta_nu(a, k, 0.15)
    T.delta_nu_init[:] = init
    T.delta_tot[:, :30] = 0
    _, part_ic = T.get_delta_nu(a, k, 0.15)
    np.testing.assert_allclose(part_hist + part_ic, full, rtol=1e-13)
    T.delta_tot[:, :30] = 2 * hist
    T.delta_nu_init[:] = 0
    _, twice = T.get_delta_nu(a, k, 0.15)
    np.testing.assert_allclose(twice, 2 * part_hist, rtol=1e-14)


def test_get_delta_nu_against_mpmath(orc, om_dnu):
    """Brute force: mpmath quadrature of the reference integrand built from the
    same natural splines (fslength spline over 16 Na nodes, delta_tot spline)."""
    interp = pytest.importorskip("scipy.interpolate")
    Na, nk = 4, 2
    T, k, a = _history(orc, om_dnu, Na, nk, amax=0.05)
    x0, xa = math.log(0.01), math.log(a)
    nfs = 16 * Na
    fsx = np.array([x0 + i * (xa - x0) / (nfs - 1.) for i in range(nfs)])
    fsy = np.array([orc.orc_fslength(x, xa, T.d.light) for x in fsx])
    fs = interp.CubicSpline(fsx, fsy, bc_type="natural")
    mnubykT = 0.15 / om_dnu.kBtnu
    fsl0 = orc.orc_fslength(x0, xa, T.d.light)
    dp = 0.01 * (orc.orc_hubble(0.01) / T.d.light) * 0.01
    _, out = T.get_delta_nu(a, k, 0.15)
    mp.mp.dps = 20
    for ik in range(nk):
        S = interp.CubicSpline(T.scalefact[:Na], T.delta_tot[ik, :Na], bc_type="natural")
        f = lambda x: (fs(float(x)) / (math.exp(float(x)) * orc.orc_hubble(math.exp(float(x))))
                       * orc.orc_specialJ(k[ik] * fs(float(x)) / mnubykT, 0, 0) * S(float(x)))
        val = float(mp.quad(f, list(fsx)))
        want = (orc.orc_specialJ(k[ik] * fsl0 / mnubykT, 0, 0) * T.delta_nu_init[ik] * (1 + dp * fsl0)
                + T.d.delta_nu_prefac * val)
        assert out[ik] == pytest.approx(want, rel=1e-6)   # QAG's own tolerance


def test_qag_vs_converged_gap_is_small(orc, om_dnu):
    """The reference's QAG result and the converged integral of the same
    interpolants agree to the QAG tolerance (1e-6), not better."""
    T0, k, a = _history(orc, om_dnu, 60, 6, mode=0)
    T1, _, _ = _history(orc, om_dnu, 60, 6, mode=1)
    _, q = T0.get_delta_nu(a, k, 0.15)
    _, c = T1.get_delta_nu(a, k, 0.15)
    assert np.max(np.abs(q / c - 1)) < 1e-6
