"""Matched integrator states for the product (C ABI) and the oracle.

Both sides get the same cosmology (their own, bit-identical Omega_nu tables),
the same Hubble function form (delta_tot_table_test.c:25-45) and the same
history bytes, so their get_delta_nu results can be compared directly.
"""
import ctypes as C
import math

import numpy as np

import knhip
import oracle_py as O
from conftest import UNIT_LENGTH, UNIT_TIME, load_testhost

DP = C.POINTER(C.c_double)


class Pair:
    def __init__(self, masses, omega0, nk, time_max, hybrid=None, time_transfer=0.01, mode=0):
        kn = knhip.lib()
        host = load_testhost()
        self.kn = kn
        self.om = knhip.OmegaNu()
        kn.init_omega_nu(C.byref(self.om), (C.c_double * 3)(*masses), time_transfer, 0.7, 2.7255)
        self.oom = O.omega_nu(masses, a0=time_transfer)
        if hybrid:
            vcrit, tcrit = hybrid
            kn.init_hybrid_nu(C.byref(self.om.hybnu), (C.c_double * 3)(*masses), vcrit, 2.99792458e10 / 1e5, tcrit,
                              self.om.kBtnu)
            O.lib().orc_init_hybrid_nu(C.byref(self.oom.hybnu), (C.c_double * 3)(*masses), vcrit,
                                       2.99792458e10 / 1e5, tcrit, self.oom.kBtnu)
        host.kn_test_init_hubble(C.byref(self.om), omega0, UNIT_TIME)
        O.lib().orc_init_hubble(C.byref(self.oom), omega0, UNIT_TIME)
        self.d = knhip.DeltaTotTable()
        kn.allocate_delta_tot_table(C.byref(self.d), nk, time_transfer, time_max, omega0, C.byref(self.om),
                                    UNIT_TIME, UNIT_LENGTH, 0)
        self.T = O.Table(self.oom, nk, time_transfer, time_max, omega0, UNIT_TIME, UNIT_LENGTH, mode=mode)
        namax = self.d.namax
        assert namax == self.T.d.namax
        self.scalefact = np.ctypeslib.as_array(self.d.scalefact, shape=(namax * (nk + 1),))
        self.nk = nk

    def set_history(self, logs, rows, dnu_init):
        """logs[Na], rows[nk][Na], dnu_init[nk] into both tables."""
        Na = len(logs)
        namax = self.d.namax
        blk = np.ctypeslib.as_array(self.d.scalefact, shape=(namax * (self.nk + 1),))
        blk[:Na] = logs
        for k in range(self.nk):
            blk[namax * (k + 1): namax * (k + 1) + Na] = rows[k]
        init = np.ctypeslib.as_array(self.d.delta_nu_init, shape=(self.nk,))
        init[:] = dnu_init
        self.d.ia = Na
        self.T.scalefact[:Na] = logs
        self.T.delta_tot[:, :Na] = rows
        self.T.delta_nu_init[:] = dnu_init
        self.T.d.ia = Na

    def gpu_get_delta_nu(self, a, k, mnu):
        out = np.zeros(self.nk)
        kk = np.ascontiguousarray(k, dtype=np.float64)
        self.kn.get_delta_nu(C.byref(self.d), a, kk.ctypes.data_as(DP), out.ctypes.data_as(DP), mnu)
        return out

    def gpu_get_delta_nu_combined(self, a, k):
        out = np.zeros(self.nk)
        kk = np.ascontiguousarray(k, dtype=np.float64)
        self.kn.get_delta_nu_combined(C.byref(self.d), a, kk.ctypes.data_as(DP), out.ctypes.data_as(DP))
        return out

    def free(self):
        self.kn.free_delta_tot_table(C.byref(self.d))


def synthetic_history(Na, nk, amax, seed=0, noise=0.05, kmax_box=1024):
    """log-spaced a in [0.01, amax], delta_tot = a sqrt(P(k)) (1 + noise N(0,1))."""
    a = np.exp(np.linspace(math.log(0.01), math.log(amax), Na))
    logs = np.array([math.log(v) for v in a])
    logs[-1] = math.log(amax)
    k = np.exp(np.linspace(math.log(2 * math.pi / 512000), math.log(math.sqrt(3) * math.pi * kmax_box / 512000), nk))
    q = k * 1000 / 0.2
    t = np.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25
    pk = 1e9 * (k * 1000) ** 0.96 * t * t
    rng = np.random.default_rng(seed)
    rows = np.outer(np.sqrt(pk), np.exp(logs)) * (1 + noise * rng.standard_normal((nk, Na)))
    return logs, rows, k, 0.4 * rows[:, 0]
