"""ctypes view of the oracle (oracle/_build/libkn_oracle.so).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# KN_ORACLE_BUILD=fast selects the build with the reference's compile flags
# (-O2 -ffast-math), used only by the CPU-baseline timings.
LIB_PATH = os.path.join(ROOT, "oracle", "_build",
                        "libkn_oracle_fast.so" if os.environ.get("KN_ORACLE_BUILD") == "fast" else "libkn_oracle.so")
FAST_LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libkn_oracle_fast.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
llp = C.POINTER(C.c_longlong)
QAG_LIMIT = 200  # GSL_VAL: QAG subinterval limit (kspace_neutrino_const.h)


class OrcRhoNu(C.Structure):
    _fields_ = [("have_table", C.c_int), ("mnu", C.c_double), ("loga", C.c_double * 200),
                ("rhonu", C.c_double * 200), ("c2", C.c_double * 200)]


class OrcHybrid(C.Structure):
    _fields_ = [("enabled", C.c_int), ("nufrac_low", C.c_double * 3), ("nu_crit_time", C.c_double),
                ("vcrit", C.c_double)]


class OrcOmegaNu(C.Structure):
    _fields_ = [("tab", OrcRhoNu * 3), ("deg", C.c_int * 3), ("rhocrit", C.c_double), ("kBtnu", C.c_double),
                ("tcmb0", C.c_double), ("hybnu", OrcHybrid)]


class OrcTransfer(C.Structure):
    _fields_ = [("NPowerTable", C.c_int), ("logk", dp), ("T_nu", dp)]


class OrcDtt(C.Structure):
    _fields_ = [("nk", C.c_int), ("nk_allocated", C.c_int), ("namax", C.c_int), ("ia", C.c_int),
                ("delta_nu_prefac", C.c_double), ("delta_tot_init_done", C.c_int),
                ("scalefact", dp), ("block_delta_tot", dp), ("delta_nu_init", dp), ("delta_nu_last", dp),
                ("wavenum", dp), ("omnu", C.POINTER(OrcOmegaNu)), ("Omeganonu", C.c_double),
                ("light", C.c_double), ("TimeTransfer", C.c_double), ("mode", C.c_int),
                ("last_error", C.c_int), ("qag_max_intervals", C.c_int)]


class OrcDeltaPow(C.Structure):
    _fields_ = [("logkk", dp), ("ratio", dp), ("nbins", C.c_int), ("norm", C.c_double)]


_lib = None
_lib_fast = None


def lib_fast():
    """The fast-math build (stateless binning/rescale functions only: the
    integrator state of this process lives in lib())."""
    global _lib_fast
    if _lib_fast is None:
        _lib_fast = _bind(FAST_LIB_PATH)
    return _lib_fast


def lib():
    global _lib
    if _lib is None:
        _lib = _bind(LIB_PATH)
    return _lib


def _bind(path):
    if True:  # one indentation level with the signature table below
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        sig = {
            "orc_init_omega_nu": (None, [C.POINTER(OrcOmegaNu), dp, C.c_double, C.c_double, C.c_double]),
            "orc_get_omega_nu": (C.c_double, [C.POINTER(OrcOmegaNu), C.c_double]),
            "orc_get_omega_nu_nopart": (C.c_double, [C.POINTER(OrcOmegaNu), C.c_double]),
            "orc_get_omegag": (C.c_double, [C.POINTER(OrcOmegaNu), C.c_double]),
            "orc_omega_nu_single": (C.c_double, [C.POINTER(OrcOmegaNu), C.c_double, C.c_int]),
            "orc_nufrac_low": (C.c_double, [C.c_double]),
            "orc_init_hybrid_nu": (None, [C.POINTER(OrcHybrid), dp, C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_particle_nu_fraction": (C.c_double, [C.POINTER(OrcHybrid), C.c_double, C.c_int]),
            "orc_rho_nu_exact": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "orc_init_hubble": (None, [C.POINTER(OrcOmegaNu), C.c_double, C.c_double]),
            "orc_hubble": (C.c_double, [C.c_double]),
            "orc_specialJ": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "orc_fslength": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "orc_get_delta_tot": (C.c_double, [C.c_double] * 6),
            "orc_bessel_j0": (C.c_double, [C.c_double]),
            "orc_transfer_load": (C.c_int, [C.POINTER(OrcTransfer), C.c_double, C.c_double, C.c_double, C.c_char_p]),
            "orc_transfer_free": (None, [C.POINTER(OrcTransfer)]),
            "orc_dtt_allocate": (None, [C.POINTER(OrcDtt), C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(OrcOmegaNu), C.c_double, C.c_double]),
            "orc_dtt_free": (None, [C.POINTER(OrcDtt)]),
            "orc_delta_tot_init": (C.c_int, [C.POINTER(OrcDtt), C.c_int, dp, dp, C.POINTER(OrcTransfer), C.c_double]),
            "orc_update_delta_tot": (None, [C.POINTER(OrcDtt), C.c_double, dp, dp, C.c_int]),
            "orc_get_delta_nu": (C.c_int, [C.POINTER(OrcDtt), C.c_double, dp, dp, C.c_double]),
            "orc_get_delta_nu_combined": (C.c_int, [C.POINTER(OrcDtt), C.c_double, dp, dp]),
            "orc_get_delta_nu_update": (C.c_int, [C.POINTER(OrcDtt), C.c_double, C.c_int, dp, dp, dp,
                                                  C.POINTER(OrcTransfer)]),
            "orc_total_powerspectrum": (C.c_int, [C.c_int, dp, C.c_int, C.c_int, C.c_int, dp, llp, dp]),
            "orc_bin_sums": (None, [C.c_int, dp, C.c_int, C.c_int, C.c_int, dp, llp, dp]),
            "orc_green": (None, [C.c_int, dp, C.c_int, C.c_int, C.c_double]),
            "orc_psindex": (C.c_int, [C.c_longlong, C.c_int, C.c_int]),
            "orc_get_dnudcdm": (C.c_double, [C.POINTER(OrcDeltaPow), C.c_double]),
            "orc_rescale": (C.c_int, [C.c_int, dp, C.c_int, C.c_int, C.c_double, C.POINTER(OrcDeltaPow)]),
            "orc_kspace_prefac": (C.c_double, [C.POINTER(OrcDtt), C.c_double]),
            "orc_get_delta_nu_record": (C.c_int, [C.POINTER(OrcDtt), C.c_double, dp, dp, C.c_double, dp, ip]),
            "orc_dnu_frozen": (C.c_int, [C.POINTER(OrcDtt), C.c_double, dp, dp, C.c_double, dp, ip]),
            "orc_get_delta_nu_adjoint": (C.c_int, [C.POINTER(OrcDtt), C.c_double, dp, C.c_double, dp, dp, dp, dp,
                                                   dp]),
            "orc_specialJ_dmu": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_double]),
            "orc_bin_power_adjoint": (C.c_int, [C.c_int, dp, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        return L


def ptr(a):
    """Pointer to a contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


def omega_nu(masses, a0=0.01, h=0.7, tcmb=2.7255):
    om = OrcOmegaNu()
    m = (C.c_double * 3)(*masses)
    lib().orc_init_omega_nu(C.byref(om), m, a0, h, tcmb)
    return om


def total_powerspectrum(field_c, nrbins, startslab=0, nslab=None):
    """field_c: complex128 array [nslab][N][N/2+1]."""
    dims = field_c.shape[1]
    if nslab is None:
        nslab = field_c.shape[0]
    f = np.ascontiguousarray(field_c).view(np.float64).ravel()
    power = np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    keffs = np.zeros(nrbins)
    n = lib().orc_total_powerspectrum(dims, ptr(f), nrbins, startslab, nslab, ptr(power),
                                      count.ctypes.data_as(llp), ptr(keffs))
    return n, power, count, keffs


def bin_power_adjoint(field_c, nrbins, gbar, startslab=0):
    """dL/dF (complex array like field_c) for L = sum_n gbar[n] power[n]."""
    dims = field_c.shape[1]
    f = np.ascontiguousarray(field_c).view(np.float64).ravel()
    g = np.ascontiguousarray(gbar, dtype=np.float64)
    out = np.zeros_like(f)
    rc = lib().orc_bin_power_adjoint(dims, ptr(f), nrbins, startslab, field_c.shape[0], ptr(g), len(g), ptr(out))
    return rc, out.view(np.complex128).reshape(field_c.shape)


def bin_sums(field_c, nrbins, startslab=0, L=None):
    """Raw per-bin sums (power, count, keff) of the slabs in field_c [nslab][N][N/2+1]."""
    dims = field_c.shape[1]
    f = np.ascontiguousarray(field_c).view(np.float64).ravel()
    power, keffs = np.zeros(nrbins), np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    (L or lib()).orc_bin_sums(dims, ptr(f), nrbins, startslab, field_c.shape[0], ptr(power),
                              count.ctypes.data_as(llp), ptr(keffs))
    return power, count, keffs


def rescale(field_c, boxsize, logkk, ratio, norm, startslab=0, L=None, in_place=False):
    dims = field_c.shape[1]
    f = field_c if in_place else np.ascontiguousarray(field_c.copy())
    assert f.flags.c_contiguous
    fv = f.view(np.float64).ravel()
    lk = np.ascontiguousarray(logkk, dtype=np.float64)
    rt = np.ascontiguousarray(ratio, dtype=np.float64)
    dpow = OrcDeltaPow(ptr(lk), ptr(rt), len(lk), norm)
    rc = (L or lib()).orc_rescale(dims, ptr(fv), startslab, f.shape[0], boxsize, C.byref(dpow))
    return rc, f


def green(field_c, boxsize, asmth, startslab=0):
    """The PM Green's-function multiply (a copy of field_c)."""
    dims = field_c.shape[1]
    f = np.ascontiguousarray(field_c.copy())
    s = 2 * np.pi * asmth / boxsize
    lib().orc_green(dims, ptr(f.view(np.float64).ravel()), startslab, f.shape[0], s * s)
    return f


def get_dnudcdm(logkk, ratio, norm, x):
    lk = np.ascontiguousarray(logkk, dtype=np.float64)
    rt = np.ascontiguousarray(ratio, dtype=np.float64)
    dpow = OrcDeltaPow(ptr(lk), ptr(rt), len(lk), norm)
    return lib().orc_get_dnudcdm(C.byref(dpow), x)


class Table:
    """Oracle integrator state with numpy views of its history block."""

    def __init__(self, om, nk, time_transfer, time_max, omega0, unit_time, unit_length, mode=0):
        self.om = om
        self.d = OrcDtt()
        lib().orc_dtt_allocate(C.byref(self.d), nk, time_transfer, time_max, omega0, C.byref(om),
                               unit_time, unit_length)
        self.d.mode = mode
        namax = self.d.namax
        self.scalefact = np.ctypeslib.as_array(self.d.scalefact, shape=(namax,))
        self.delta_tot = np.ctypeslib.as_array(self.d.block_delta_tot, shape=(nk, namax))
        self.delta_nu_init = np.ctypeslib.as_array(self.d.delta_nu_init, shape=(nk,))
        self.delta_nu_last = np.ctypeslib.as_array(self.d.delta_nu_last, shape=(nk,))
        self.wavenum = np.ctypeslib.as_array(self.d.wavenum, shape=(nk,))

    def __del__(self):
        try:
            lib().orc_dtt_free(C.byref(self.d))
        except Exception:
            pass

    def get_delta_nu(self, a, wavenum, mnu):
        out = np.zeros(self.d.nk)
        w = np.ascontiguousarray(wavenum, dtype=np.float64)
        rc = lib().orc_get_delta_nu(C.byref(self.d), a, ptr(w), ptr(out), mnu)
        return rc, out

    def get_delta_nu_combined(self, a, wavenum):
        out = np.zeros(self.d.nk)
        w = np.ascontiguousarray(wavenum, dtype=np.float64)
        rc = lib().orc_get_delta_nu_combined(C.byref(self.d), a, ptr(w), ptr(out))
        return rc, out

    def get_delta_nu_record(self, a, wavenum, mnu):
        """Forward plus each k's final QAG subdivision (ivl[k, t] = (lo, hi), nint[k])."""
        nk = self.d.nk
        out = np.zeros(nk)
        ivl = np.zeros((nk, QAG_LIMIT, 2))
        nint = np.zeros(nk, dtype=np.int32)
        w = np.ascontiguousarray(wavenum, dtype=np.float64)
        rc = lib().orc_get_delta_nu_record(C.byref(self.d), a, ptr(w), ptr(out), mnu, ptr(ivl),
                                           nint.ctypes.data_as(ip))
        return rc, out, ivl, nint

    def dnu_frozen(self, a, wavenum, mnu, ivl, nint):
        out = np.zeros(self.d.nk)
        w = np.ascontiguousarray(wavenum, dtype=np.float64)
        rc = lib().orc_dnu_frozen(C.byref(self.d), a, ptr(w), ptr(out), mnu, ptr(ivl), nint.ctypes.data_as(ip))
        return rc, out

    def get_delta_nu_adjoint(self, a, wavenum, mnu, gbar):
        """(rc, delta_nu, dtot_bar[nk, Na], init_bar[nk], mnu_bar)."""
        nk, na = self.d.nk, self.d.ia
        out, ib = np.zeros(nk), np.zeros(nk)
        db = np.zeros((nk, na))
        mb = C.c_double()
        w = np.ascontiguousarray(wavenum, dtype=np.float64)
        g = np.ascontiguousarray(gbar, dtype=np.float64)
        rc = lib().orc_get_delta_nu_adjoint(C.byref(self.d), a, ptr(w), mnu, ptr(g), ptr(out), ptr(db), ptr(ib),
                                            C.byref(mb))
        return rc, out, db, ib, mb.value

    def delta_tot_init(self, wavenum, delta_cdm, transfer, time):
        w = np.ascontiguousarray(wavenum, dtype=np.float64)
        c = np.ascontiguousarray(delta_cdm, dtype=np.float64)
        return lib().orc_delta_tot_init(C.byref(self.d), len(w), ptr(w), ptr(c), C.byref(transfer), time)

    def get_delta_nu_update(self, a, keff, delta_cdm, transfer):
        k = np.ascontiguousarray(keff, dtype=np.float64)
        c = np.ascontiguousarray(delta_cdm, dtype=np.float64)
        out = np.zeros(len(k))
        rc = lib().orc_get_delta_nu_update(C.byref(self.d), a, len(k), ptr(k), ptr(c), ptr(out),
                                           C.byref(transfer))
        return rc, out


def load_transfer(path, boxsize, unit_length, input_unit_length):
    t = OrcTransfer()
    rc = lib().orc_transfer_load(C.byref(t), boxsize, unit_length, input_unit_length, path.encode())
    return rc, t
