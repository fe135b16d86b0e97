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


def test_compute_neutrino_power_from_cdm_over_steps(kn, torch_cuda):
    """The code-agnostic entry (interface_common.c:106-123): a caller that
    bins its own CDM power hands (keff, P_cdm, Nmodes) over; bins with no
    modes are skipped, delta_cdm = sqrt(P), and the returned delta_pow holds
    log k, delta_nu/delta_cdm and kspace_prefac (:125-148).  Five steps with
    empty bins at the start, inside and at the end, against the oracle
    composed the same way; get_dnudcdm_powerspec read back at bin centres,
    between bins and beyond both ends (the clamps of delta_pow.c:25-34; the argument is log k)."""
    N_K, box, omega0, masses = 40, 512000.0, 0.2793, [0.1, 0.1, 0.1]
    path = os.path.join(REFDATA, "ics_transfer_99.dat")
    kp = knhip.KspaceParams.in_dll(kn, "kspace_params")
    kp.KspaceTransferFunction = path.encode()
    kp.TimeTransfer = 0.01
    kp.InputSpectrum_UnitLength_in_cm = UNIT_LENGTH * 1e3
    for i in range(3):
        kp.MNu[i] = masses[i]
    kp.hybrid_neutrinos_on = 0
    kn.InitOmegaNu(0.7, 2.7255, 0)
    om = knhip.OmegaNu()
    kn.init_omega_nu(C.byref(om), (C.c_double * 3)(*masses), 0.01, 0.7, 2.7255)
    load_testhost().kn_test_init_hubble(C.byref(om), omega0, UNIT_TIME)
    table = knhip.DeltaTotTable.in_dll(kn, "delta_tot_table")
    if table.nk_allocated > 0:  # an earlier test of this process set the module table up
        kn.free_delta_tot_table(C.byref(table))
    kn.allocate_kspace_memory(N_K, 0, box, UNIT_TIME, UNIT_LENGTH, omega0, None, 1.0, 0)
    oom = O.omega_nu(masses, a0=0.01)
    O.lib().orc_init_hubble(C.byref(oom), omega0, UNIT_TIME)
    rc, ot = O.load_transfer(path, box, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    assert rc == 0

    k_in = np.exp(np.linspace(math.log(2 * math.pi / box), math.log(2 * math.pi / box * 40), N_K))
    nmodes = np.arange(N_K, dtype=np.int64) * 6 + 1
    nmodes[[0, 1, 17, N_K - 1]] = 0  # empty bins are skipped
    keep = nmodes > 0
    T = None
    worst_r, worst_i = 0.0, 0.0
    for step, a in enumerate([0.01, 0.02, 0.04, 0.07, 0.1]):
        q = k_in * 1000 / 0.2
        t = np.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25
        p_cdm = 1e9 * (k_in * 1000) ** 0.96 * t * t * (a / 0.01) ** 2 * (1 + 0.05 * np.sin(step + np.arange(N_K)))
        print("compute_neutrino_power_from_cdm a=%g" % a, flush=True)
        kin = np.ascontiguousarray(k_in)
        pin = np.ascontiguousarray(p_cdm)
        nm = np.ascontiguousarray(nmodes.astype(np.int64))
        dp = kn.compute_neutrino_power_from_cdm(a, kin.ctypes.data_as(C.POINTER(C.c_double)),
                                                pin.ctypes.data_as(C.POINTER(C.c_double)),
                                                nm.ctypes.data_as(C.POINTER(C.c_long)), N_K, 0)
        # the oracle composition on the same inputs
        dcdm = np.sqrt(p_cdm[keep])
        keff = k_in[keep]
        if T is None:
            T = O.Table(oom, len(keff), 0.01, 1.0, omega0, UNIT_TIME, UNIT_LENGTH)
        rc, dnu = T.get_delta_nu_update(a, keff, dcdm, ot)
        assert rc == 0
        nopart = O.lib().orc_get_omega_nu_nopart(C.byref(oom), a)
        hybrid = O.lib().orc_get_omega_nu(C.byref(oom), a) - nopart
        prefac = nopart / (T.d.Omeganonu / a ** 3 + hybrid)
        n = int(np.sum(keep))
        assert dp.nbins == n
        logk = np.array(dp.logkk[:n])
        ratio = np.array(dp.delta_ratio[:n])
        assert np.max(np.abs(logk - np.log(keff))) <= 1e-15 * np.max(np.abs(np.log(keff)))
        assert abs(dp.norm / prefac - 1) < 1e-14
        want = dnu / dcdm
        e_r = float(np.max(np.abs(ratio - want)) / np.max(np.abs(want)))
        # the interpolant read back: centres, midpoints, and beyond both ends
        xs = np.concatenate([np.log(keff), 0.5 * (np.log(keff[1:]) + np.log(keff[:-1])),
                             [np.log(keff[0]) - 0.5, np.log(keff[-1]) + 0.5]])
        got = np.array([kn.get_dnudcdm_powerspec(C.byref(dp), float(x)) for x in xs])
        ref = np.array([O.get_dnudcdm(np.log(keff), want, prefac, float(x)) for x in xs])
        e_i = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        print("  a=%g: %d of %d bins kept, ratio %.2e, interpolant %.2e" % (a, n, N_K, e_r, e_i))
        worst_r, worst_i = max(worst_r, e_r), max(worst_i, e_i)
        kn.free_d_pow(C.byref(dp))
    assert T.d.ia >= 4
    assert worst_r < 1e-9 and worst_i < 1e-9
