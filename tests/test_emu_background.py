"""(f)4: the background service (kn_background: Omega_nu(a) and H(a)
tabulated on the device) in the CPU SIMT emulation build of
csrc/delta_nu.hip, against the oracle.

The density tables are rho_nu_init's (omega_nu_single.c:117-153): 200 log-a
points, each a QAG-GK61 integral at 1e-9, here one emulated wavefront each;
Omega_nu(a) follows get_omega_nu (:55-67) and H(a) the N-body code's
hubble_function (gadget-2 patch 0003:38-55), which for a flat background
equals the reference test's form (delta_tot_table_test.c:25-45) that the
oracle's H uses.  Groups whose histories take their H from a background
build its Chebyshev panels (with their split panels) on the device.
"""
import ctypes as C

import numpy as np
import pytest

import knhip
import oracle_py as O
from batch_common import RTOL, append_flow, mixed_batch, rel
from conftest import UNIT_TIME
from test_emu_group import _forward_errors, emu, host_memcpy  # noqa: F401

H0 = 3.24077929e-18 * UNIT_TIME


def _grid():
    """log-spaced a in [0.005, 1] plus both sides of every species' switch
    to the non-relativistic series (a = 100 kT/m) and of the table's lower end."""
    a = list(np.exp(np.linspace(np.log(0.005), 0.0, 1501)))
    kT = 8.61734e-5 * (4 / 11.) ** (1 / 3.) * 1.00328 * 2.7255
    for m in (0.06, 0.3):
        for x in (100 * kT / m, 0.01 / 1.2):
            a += [x * (1 - 1e-9), x, x * (1 + 1e-9)]
    return np.array(sorted(a))


@pytest.mark.parametrize("masses", [[0.06, 0.06, 0.3], [0.15, 0, 0], [0.1, 0.1, 0.1], [0, 0, 0]],
                         ids=["0.06-0.06-0.3", "0.15", "3x0.1", "massless"])
def test_emu_background_matches_the_oracle(emu, masses):
    """Density tables, Omega_nu(a) and H(a) (device kernel and host form)
    against the oracle's init_omega_nu / get_omega_nu / H."""
    bg = knhip.Background(0.3, 0.7, H0, masses, a0=0.01, L=emu)
    om = O.omega_nu(masses, a0=0.01)
    O.lib().orc_init_hubble(C.byref(om), 0.3, UNIT_TIME)
    for i in range(3):
        t = bg.rho_table(i)
        assert (t is not None) == bool(om.tab[i].have_table), i
        if t is None:
            continue
        assert np.array_equal(t[0], np.array(om.tab[i].loga))
        e = rel(t[1], np.array(om.tab[i].rhonu))
        print("  species %d (m %g): density table max rel err %.2e" % (i, masses[i], e))
        assert e < 1e-13
    a = _grid()
    out = np.zeros_like(a)
    assert emu.kn_background_omega_nu(bg.ctx, len(a), a.ctypes.data, out.ctypes.data, None) == 0
    o = np.array([O.lib().orc_get_omega_nu(C.byref(om), x) for x in a])
    e_o = rel(out, o)
    assert emu.kn_background_hubble(bg.ctx, len(a), a.ctypes.data, out.ctypes.data, None) == 0
    h = np.array([O.lib().orc_hubble(x) for x in a])
    e_h = rel(out, h)
    e_hh = rel(np.array([bg.hubble_host(x) for x in a]), h)
    print("  Omega_nu %.2e, H device %.2e, H host %.2e" % (e_o, e_h, e_hh))
    assert e_o < 1e-13 and e_h < 1e-13 and e_hh < 1e-13
    bg.close()


def test_emu_background_checks(emu):
    p = knhip.BackgroundParams()
    p.Omega0, p.OmegaLambda, p.Hubble, p.HubbleParam, p.tcmb0, p.a0 = 0.3, 0.7, H0, 0.7, 2.7255, 0.01
    h = C.c_void_p()
    for field, bad in (("Hubble", 0.0), ("tcmb0", -1.0), ("a0", 0.0), ("a0", 2.0), ("HubbleParam", float("nan"))):
        q = knhip.BackgroundParams.from_buffer_copy(p)
        setattr(q, field, bad)
        assert emu.kn_background_create(C.byref(h), C.byref(q)) == 3002, field
    q = knhip.BackgroundParams.from_buffer_copy(p)
    q.MNu[1] = -0.1
    assert emu.kn_background_create(C.byref(h), C.byref(q)) == 3002
    assert emu.kn_background_rho_table(None, 0, None, None) == -3002
    assert emu.kn_background_hubble(None, 1, None, None, None) == 3002


def test_emu_group_on_a_background(emu):
    """Twelve histories over three cosmologies (one hybrid, its H with two
    split panels in range), their H panels built from device backgrounds,
    forward and adjoint against the oracle, and against the same group on
    the host's H."""
    cosmos, specs, B = mixed_batch(4, seed=3, extra=2, L=emu, memcpy=host_memcpy, background=True)
    B.solve()
    worst = _forward_errors(B, cosmos, specs)
    cosmos_h, specs_h, Bh = mixed_batch(4, seed=3, extra=2, L=emu, memcpy=host_memcpy)
    Bh.solve()
    diff = max(rel(B.delta_nu(i), Bh.delta_nu(i)) for i in range(B.n))
    print("  %d histories: worst forward rel err %.3e vs the oracle, %.3e vs host-H group" % (B.n, worst, diff))
    assert worst < RTOL and diff < RTOL
    for X, cs in ((B, cosmos), (Bh, cosmos_h)):
        X.close()
        for c in cs:
            c.free()


def test_emu_append_flow_on_a_background(emu):
    """The append flow (IC-only start, hybrid switches, set_mass) with the
    histories' H from device backgrounds: every step against the oracle."""
    worst = append_flow(L=emu, memcpy=host_memcpy, mode="eager", background=True)
    print("  worst rel err over the steps: %.2e" % worst)
    assert worst < RTOL


def test_emu_process_table_from_a_background(emu):
    """The per-call H table (ln T0 - 20 to ln T0 + 8, 7168 panels) built from
    a background on the device: the panels read back on the host within
    1e-12 of the oracle's H over the integrals' range, across both masses'
    switches to the non-relativistic series (split panels)."""
    masses = [0.06, 0.06, 0.3]
    bg = knhip.Background(0.3, 0.7, H0, masses, a0=0.01, L=emu)
    om = O.omega_nu(masses, a0=0.01)
    O.lib().orc_init_hubble(C.byref(om), 0.3, UNIT_TIME)
    emu.kn_htable_eval_host.restype = C.c_double
    emu.kn_htable_eval_host.argtypes = [C.c_double]
    emu.kn_htable_ensure.argtypes = [C.c_double, C.c_double, C.c_double]
    x0 = np.log(0.01)
    bg.set_process(True)
    try:
        emu.kn_htable_ensure(x0, x0, 0.0)
        xs = np.concatenate([np.random.default_rng(2).uniform(x0, 0.0, 3000), np.log(_grid()[_grid() >= 0.01])])
        e = max(abs(emu.kn_htable_eval_host(x) / O.lib().orc_hubble(np.exp(x)) - 1) for x in xs)
    finally:
        bg.set_process(False)
    print("  H panels from the background: max rel err %.2e" % e)
    assert e < 1e-12
    bg.close()
