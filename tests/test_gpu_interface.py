"""The drop-in entry of the path, end to end on the GPU: InitOmegaNu,
allocate_kspace_memory, then add_nu_power_to_rhogrid over several PM steps
(interface_common.c:77-188, interface_gadget.c:75-194), against the oracle
composed the same way (binning -> get_delta_nu_update -> ratio table ->
rescale) on copies of the same grids.

The library keeps this state in globals, as the reference does, so the whole
flow is one test.  Bars: the rescaled grid to 1e-8 of the correction it
received (delta_nu agrees to ~1e-10; the correction is norm * ratio), and to
1e-12 overall.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import knhip
import oracle_py as O
from conftest import REFDATA, UNIT_LENGTH, UNIT_TIME, load_testhost

pytestmark = pytest.mark.gpu


def test_add_nu_power_to_rhogrid_over_pm_steps(kn, torch_cuda):
    torch = torch_cuda
    N, box, omega0, masses = 64, 512000.0, 0.2793, [0.15, 0.15, 0.15]
    nk_in = N // 2  # Gadget passes PMGRID/2 (patch 0002 line 59)
    path = os.path.join(REFDATA, "ics_transfer_99.dat")
    # parameters as Gadget's parameter file would set them (interface_gadget.c:25-58)
    kp = knhip.KspaceParams.in_dll(kn, "kspace_params")
    kp.KspaceTransferFunction = path.encode()
    kp.TimeTransfer = 0.01
    kp.InputSpectrum_UnitLength_in_cm = UNIT_LENGTH * 1e3
    for i in range(3):
        kp.MNu[i] = masses[i]
    kp.hybrid_neutrinos_on = 0
    kn.InitOmegaNu(0.7, 2.7255, 0)
    # the host's hubble_function on an identical Omega_nu table
    om = knhip.OmegaNu()
    kn.init_omega_nu(C.byref(om), (C.c_double * 3)(*masses), 0.01, 0.7, 2.7255)
    load_testhost().kn_test_init_hubble(C.byref(om), omega0, UNIT_TIME)
    kn.allocate_kspace_memory(nk_in, 0, box, UNIT_TIME, UNIT_LENGTH, omega0, None, 1.0, 0)
    # the oracle's mirror of the same state
    oom = O.omega_nu(masses, a0=0.01)
    O.lib().orc_init_hubble(C.byref(oom), omega0, UNIT_TIME)
    rc, ot = O.load_transfer(path, box, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    assert rc == 0
    T = O.Table(oom, nk_in, 0.01, 1.0, omega0, UNIT_TIME, UNIT_LENGTH)

    grid = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
    worst_corr, worst_all = 0.0, 0.0
    for step, a in enumerate([0.01, 0.02, 0.03, 0.05, 0.08, 0.1]):
        # a growing Gaussian field: amplitude ~ a^2 in power
        assert kn.kn_synth_grid(grid.data_ptr(), N, 0, N, box, 1e4 * (a / 0.01) ** 2, 0.96, 0.2 / 1000.0,
                                100 + step) == 0
        F = grid.cpu().numpy()
        on_host = a == 0.1  # the last step hands over a host grid (staged through HBM)
        print("add_nu_power_to_rhogrid a=%g (%s grid)" % (a, "host" if on_host else "device"), flush=True)
        if on_host:
            G = np.ascontiguousarray(F.copy())
            kn.add_nu_power_to_rhogrid(a, box, G.ctypes.data, N, 0, N, 0)
        else:
            kn.add_nu_power_to_rhogrid(a, box, grid.data_ptr(), N, 0, N, 0)
            G = grid.cpu().numpy()
        # oracle composition (interface_gadget.c:75-102, interface_common.c:125-148, :158-194)
        n, p, c, k = O.total_powerspectrum(F, nk_in)
        dcdm = np.sqrt(p[:n] / box ** -3)
        keff = k[:n] * 2 * math.pi / box
        rc, dnu = T.get_delta_nu_update(a, keff, dcdm, ot)
        assert rc == 0
        nopart = O.lib().orc_get_omega_nu_nopart(C.byref(oom), a)
        hybrid = O.lib().orc_get_omega_nu(C.byref(oom), a) - nopart
        prefac = nopart / (T.d.Omeganonu / a ** 3 + hybrid)
        rc2, W = O.rescale(F, box, np.log(keff), dnu / dcdm, prefac)
        assert rc2 == 0
        corr = np.abs(W - F)
        e_corr = float(np.max(np.abs(G - W)) / np.max(corr))
        e_all = float(np.max(np.abs(G - W) / np.maximum(np.abs(W), 1e-300)))
        print("  a=%g: ia=%d, error %.2e of the correction, %.2e overall" % (a, T.d.ia, e_corr, e_all))
        worst_corr, worst_all = max(worst_corr, e_corr), max(worst_all, e_all)
    assert T.d.ia >= 5
    assert worst_corr < 1e-8 and worst_all < 1e-12
