"""QUADPACK's error rescaling as the device evaluates it (dnu_device.h
gk_rescale, run through the CPU emulation build of the same source):
resasc * min(1, (200 err / resasc)^1.5) with the power rounded once.
Pinned against mpmath at 50 digits (the correctly rounded t^1.5) and against
the oracle's libm pow(t, 1.5), which GSL's qk uses (gsl/integration/qk.c
rescale_error)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT

EMU = os.path.join(ROOT, "tests", "emu", "_build", "libkn_dnu_emu.so")


@pytest.fixture(scope="module")
def emu(kn):
    L = C.CDLL(EMU)
    L.emu_gk_rescale.restype = C.c_double
    L.emu_gk_rescale.argtypes = [C.c_double] * 3
    return L


def test_rescale_power_is_correctly_rounded(emu):
    import mpmath
    mpmath.mp.dps = 50
    rng = np.random.default_rng(12)
    t = np.concatenate([np.exp(-27.0 * rng.uniform(0, 1, 20000)), 1 - 1e-3 * rng.uniform(0, 1, 5000)])
    wrong_dev = wrong_libm = 0
    for ti in t:
        err = float(ti) / 200.0
        tt = 200 * err                     # the t the rule forms (resasc = 1)
        exact = float(mpmath.mpf(tt) ** mpmath.mpf(1.5))
        dev = emu.emu_gk_rescale(err, 1e-300, 1.0)
        if exact >= 1:
            assert dev == 1.0
            continue
        wrong_dev += dev != exact
        wrong_libm += tt ** 1.5 != exact
    print("  device form wrong in %d of %d; libm pow wrong in %d" % (wrong_dev, len(t), wrong_libm))
    assert wrong_dev == 0


def test_rescale_floor_and_zero_cases(emu):
    # err 0 stays 0 (no rescale), resasc 0 leaves |err|, the 50 eps resabs floor applies
    assert emu.emu_gk_rescale(0.0, 0.0, 1.0) == 0.0
    assert emu.emu_gk_rescale(-3e-5, 0.0, 0.0) == 3e-5
    eps = np.finfo(float).eps
    assert emu.emu_gk_rescale(1e-30, 1.0, 1.0) == 50 * eps * 1.0
