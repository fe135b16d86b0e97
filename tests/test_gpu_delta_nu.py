"""(b) get_delta_nu and its update/init flows on the GPU, through the C ABI,
against the oracle's restatement of the reference QAG procedure.

Bar: 1e-9 relative.  The GPU replays the reference's
own interpolants and QAG decisions, so agreement is typically ~1e-13; the
max error against the oracle is printed per case."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import knhip
import oracle_py as O
from conftest import REFDATA, UNIT_LENGTH
from batch_common import adjoint_case, compare_adjoint
from dnu_common import DP, Pair, synthetic_history
from test_oracle_pins import _read_powerspec_fixtures, load_camb, read_all_nu_state

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def rel(a, b):
    return float(np.max(np.abs(a / b - 1)))


def test_hubble_interpolant_matches_host(kn, host, torch_cuda):
    """H factor panels anchored at ln T0 (H read back from 1/(a H)).  Away from the host's own branch switch
    (rho_nu's table -> series at a = 100 kT/m, a ~1e-8 jump of H) the
    interpolant is exact to ~1e-14; the panel holding the jump smears it."""
    P = Pair([0.1, 0.1, 0.1], 0.3, 8, 1.0)
    kn.kn_htable_eval_host.restype = C.c_double
    kn.kn_htable_eval_host.argtypes = [C.c_double]
    kn.kn_htable_ensure.argtypes = [C.c_double, C.c_double, C.c_double]
    x0 = math.log(0.01)
    kn.kn_htable_ensure(x0, x0, 0.5)
    xs = np.random.default_rng(0).uniform(x0, 0.4, 4000)
    h = host.hubble_function
    h.restype = C.c_double
    h.argtypes = [C.c_double]
    err = np.array([abs(kn.kn_htable_eval_host(x) / h(math.exp(x)) - 1) for x in xs])
    a_sw = 100 * P.om.kBtnu / 0.1
    far = np.abs(xs - math.log(a_sw)) > 2.0 / 256
    print("H interpolant: max rel err %.3e away from the switch, %.3e overall" % (err[far].max(), err.max()))
    assert err[far].max() < 1e-13 and err.max() < 1e-7
    P.free()


@pytest.mark.parametrize("case", [
    dict(name="C1-like", masses=[0.15, 0, 0], Na=100, nk=64, amax=0.5, noise=0.0),
    dict(name="3x0.1eV", masses=[0.1, 0.1, 0.1], Na=300, nk=48, amax=1.0, noise=0.0),
    dict(name="noisy", masses=[0.3, 0, 0], Na=200, nk=32, amax=1.0, noise=0.1),
    dict(name="Na=3", masses=[0.2, 0, 0], Na=3, nk=16, amax=0.05, noise=0.0),
    dict(name="Na=2-linear", masses=[0.2, 0, 0], Na=2, nk=16, amax=0.03, noise=0.0),
])
def test_get_delta_nu_matches_oracle(kn, torch_cuda, case):
    P = Pair(case["masses"], 0.3, case["nk"], 0.01 + case["Na"] / 100.0 + 0.02)
    logs, rows, k, init = synthetic_history(case["Na"], case["nk"], case["amax"], noise=case["noise"])
    P.set_history(logs, rows, init)
    a = case["amax"]
    m = case["masses"][0]
    print("get_delta_nu", case["name"], flush=True)
    g = P.gpu_get_delta_nu(a, k, m)
    rc, o = P.T.get_delta_nu(a, k, m)
    assert rc == 0
    e = rel(g, o)
    print("  %s: max rel err vs oracle %.3e" % (case["name"], e))
    assert e < RTOL
    P.free()


def test_get_delta_nu_ic_only_cases(kn, torch_cuda):
    P = Pair([0.15, 0, 0], 0.3, 12, 0.3)
    logs, rows, k, init = synthetic_history(20, 12, 0.2)
    P.set_history(logs, rows, init)
    for m in [0.0]:                      # massless: IC piece only
        g = P.gpu_get_delta_nu(0.2, k, m)
        rc, o = P.T.get_delta_nu(0.2, k, m)
        assert rel(g, o) < 1e-14
    P.d.ia = 1                           # one stored step: IC piece only
    P.T.d.ia = 1
    g = P.gpu_get_delta_nu(0.2, k, 0.15)
    rc, o = P.T.get_delta_nu(0.2, k, 0.15)
    assert rel(g, o) < 1e-14
    P.free()


def test_hybrid_species_matches_oracle(kn, torch_cuda):
    """Hybrid neutrinos on (vcrit = 500 km/s, particles from a = 0.3333): the
    truncated Fermi-Dirac series and the loosened tolerance."""
    P = Pair([0.3, 0.06, 0.06], 0.3, 24, 0.01 + 150 / 100.0 + 0.02, hybrid=(500.0, 0.3333))
    logs, rows, k, init = synthetic_history(150, 24, 0.9)
    P.set_history(logs, rows, init)
    for m in [0.3, 0.06]:
        g = P.gpu_get_delta_nu(0.9, k, m)
        rc, o = P.T.get_delta_nu(0.9, k, m)
        e = rel(g, o)
        print("  hybrid m=%g: max rel err %.3e" % (m, e))
        assert rc == 0 and e < RTOL
    g = P.gpu_get_delta_nu_combined(0.9, k)
    rc, o = P.T.get_delta_nu_combined(0.9, k)
    assert rel(g, o) < RTOL
    P.free()


def test_combined_nondegenerate_species(kn, torch_cuda):
    P = Pair([0.05, 0.1, 0.2], 0.3, 20, 0.8)
    logs, rows, k, init = synthetic_history(60, 20, 0.6)
    P.set_history(logs, rows, init)
    g = P.gpu_get_delta_nu_combined(0.6, k)
    rc, o = P.T.get_delta_nu_combined(0.6, k)
    e = rel(g, o)
    print("  combined 3 species: %.3e" % e)
    assert rc == 0 and e < RTOL
    P.free()


def _transfer(kn, path, box=512000):
    t = knhip.TransferInitTable()
    kn.allocate_transfer_init_table(C.byref(t), box, UNIT_LENGTH, UNIT_LENGTH * 1e3, path.encode())
    rc, ot = O.load_transfer(path, box, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    return t, ot


def test_reference_init_and_update_flow(kn, torch_cuda):
    """delta_tot_table_test.c:127-152 and :193-228 on the GPU path, against the oracle."""
    scale, k, dnu, dcdm = _read_powerspec_fixtures()
    P = Pair([0.15, 0.15, 0.15], 0.2793, len(k), 1.0)
    na3 = kn.get_omega_nu(C.byref(P.om), 0.01) * 0.01 ** 3
    oma = 0.2793 - kn.get_omega_nu(C.byref(P.om), 1) + na3
    fnu = na3 / oma
    dcdm = (dcdm - fnu * dnu) / (1. - fnu)
    t, ot = _transfer(kn, os.path.join(REFDATA, "ics_transfer_99.dat"))
    kk = np.ascontiguousarray(k)
    cd = np.ascontiguousarray(dcdm)
    # init from scratch at a = 0.01
    kn.delta_tot_init(C.byref(P.d), len(k), kk.ctypes.data_as(DP), cd.ctypes.data_as(DP), C.byref(t), 0.01)
    assert P.T.delta_tot_init(k, dcdm, ot, 0.01) == 0
    last = np.ctypeslib.as_array(P.d.delta_nu_last, shape=(len(k),))
    assert P.d.ia == 1 and rel(last, P.T.delta_nu_last) < 1e-14
    P.free()
    # resume 25 rows, then update at a = 1/3 (ia stays 25), then drop a row and update again
    P = Pair([0.15, 0.15, 0.15], 0.2793, len(k), 1.0)
    read_all_nu_state(P.T, os.path.join(REFDATA, "delta_tot_nu.txt"))
    P.set_history(P.T.scalefact[:25].copy(), P.T.delta_tot[:, :25].copy(), np.zeros(len(k)))
    kn.delta_tot_init(C.byref(P.d), len(k), kk.ctypes.data_as(DP), cd.ctypes.data_as(DP), C.byref(t), 0.33333333)
    assert P.T.delta_tot_init(k, dcdm, ot, 0.33333333) == 0
    out = np.zeros(len(k))
    for step in range(2):
        kn.get_delta_nu_update(C.byref(P.d), 0.33333333, len(k), kk.ctypes.data_as(DP), cd.ctypes.data_as(DP),
                               out.ctypes.data_as(DP), C.byref(t))
        rc, o = P.T.get_delta_nu_update(0.33333333, k, dcdm, ot)
        e = rel(out, o)
        print("  update step %d: ia=%d rel err %.3e, vs saved power %.3e" % (step, P.d.ia, e, rel(out, dnu)))
        assert rc == 0 and P.d.ia == 25 == P.T.d.ia and e < RTOL
        assert rel(out, dnu) < (1e-2 if step == 0 else 3e-2)
        P.d.ia -= 1
        P.T.d.ia -= 1
    P.free()


def test_reproduce_linear_theory_on_gpu(kn, torch_cuda):
    """delta_tot_table_test.c:315-363: 99 updates against CAMB, GPU vs oracle at each step."""
    P = Pair([0.15, 0.15, 0.15], 0.2793, 200, 1.0)
    t, ot = _transfer(kn, os.path.join(REFDATA, "ics_transfer_0.01.dat"))
    _, k, dcdm, _ = load_camb(0)
    kk = np.ascontiguousarray(k)
    cd = np.ascontiguousarray(dcdm)
    kn.delta_tot_init(C.byref(P.d), 200, kk.ctypes.data_as(DP), cd.ctypes.data_as(DP), C.byref(t), 0.01)
    assert P.T.delta_tot_init(k, dcdm, ot, 0.01) == 0
    acc, worst = 0.05, 0.0
    out = np.zeros(200)
    for i in range(99):
        if i == 8:
            acc = 2e-2
        if i == 22:
            acc = 1.2e-2
        a, k, dcdm, dnu_camb = load_camb(i)
        kk = np.ascontiguousarray(k)
        cd = np.ascontiguousarray(dcdm)
        kn.get_delta_nu_update(C.byref(P.d), a, 200, kk.ctypes.data_as(DP), cd.ctypes.data_as(DP),
                               out.ctypes.data_as(DP), C.byref(t))
        rc, o = P.T.get_delta_nu_update(a, k, dcdm, ot)
        assert P.d.ia == i + 1 == P.T.d.ia
        assert np.all(np.abs(dnu_camb - out) < acc * out), i
        worst = max(worst, rel(out, o))
    print("  99-step linear theory: worst rel err vs oracle %.3e" % worst)
    assert worst < RTOL
    P.free()


# ------------------------------------------------------------- adjoint (per-call path)

@pytest.mark.parametrize("case", [
    dict(name="cspline", masses=[0.15, 0, 0], Na=40, nk=10, amax=0.4, hybrid=None),
    dict(name="linear Na=2", masses=[0.2, 0, 0], Na=2, nk=8, amax=0.03, hybrid=None),
    dict(name="hybrid", masses=[0.3, 0.06, 0.06], Na=80, nk=8, amax=0.9, hybrid=(500.0, 0.3333)),
    dict(name="noisy", masses=[0.3, 0, 0], Na=150, nk=6, amax=1.0, hybrid=None, noise=0.1),
    dict(name="C1-like", masses=[0.15, 0, 0], Na=100, nk=64, amax=0.5, hybrid=None),
    # 76800 row entries: chunked upload and download, two knot parts per item
    dict(name="long Na=1200 nk=64", masses=[0.2, 0, 0], Na=1200, nk=64, amax=0.9, hybrid=None),
], ids=lambda c: c["name"])
def test_get_delta_nu_adjoint_per_call_matches_oracle(kn, torch_cuda, case):
    """kn_get_delta_nu_adjoint (csrc/delta_nu.hip) against the oracle's adjoint:
    d/d delta_tot, d/d delta_nu_init and d/d mnu to 1e-9, and the forward it
    returns against the oracle's get_delta_nu."""
    P = Pair(case["masses"], 0.3, case["nk"], 0.01 + case["Na"] / 100.0 + 0.02, hybrid=case["hybrid"])
    logs, rows, k, init = synthetic_history(case["Na"], case["nk"], case["amax"], noise=case.get("noise", 0.0))
    P.set_history(logs, rows, init)
    print("adjoint (per call)", case["name"], flush=True)
    gpu, orc = adjoint_case(P, case["amax"], k, case["masses"][0])
    compare_adjoint(case["name"], gpu, orc)
    P.free()


def test_get_delta_nu_adjoint_per_call_massless(kn, torch_cuda):
    P = Pair([0.15, 0, 0], 0.3, 8, 0.5)
    logs, rows, k, init = synthetic_history(20, 8, 0.2)
    P.set_history(logs, rows, init)
    (dnu, db, ib, mb), orc = adjoint_case(P, 0.2, k, 0.0)
    assert np.all(db == 0) and mb == 0.0
    assert rel(dnu, orc[1]) < 1e-14 and rel(ib, orc[3]) < 1e-14
    P.free()


def test_get_delta_nu_adjoint_per_call_gradcheck(kn, torch_cuda):
    """torch.autograd.gradcheck of the per-call binding (central differences,
    relative step 1e-6 on O(1) inputs), then against the oracle's gradient."""
    import torch
    P = Pair([0.15, 0, 0], 0.3, 4, 0.5)
    logs, rows, k, init = synthetic_history(10, 4, 0.3, noise=0.05)
    scale = np.max(np.abs(rows))
    rows, init = rows / scale, init / scale
    P.set_history(logs, rows, init)
    F = knhip.delta_nu_plain_function()
    dt = torch.tensor(rows, dtype=torch.float64, requires_grad=True)
    di = torch.tensor(init, dtype=torch.float64, requires_grad=True)
    m = torch.tensor(0.15, dtype=torch.float64, requires_grad=True)
    print("per-call gradcheck", flush=True)
    # linear in delta_tot and delta_nu_init: differences exact up to rounding (~1e-9)
    assert torch.autograd.gradcheck(lambda a, b, c: F.apply(a, b, c, P.d, 0.3, k), (dt, di, m), eps=1e-6,
                                    atol=1e-8, rtol=1e-6, nondet_tol=1e-12)
    g = torch.linspace(-1, 1, 4, dtype=torch.float64)
    (F.apply(dt, di, m, P.d, 0.3, k) * g).sum().backward()
    P.set_history(logs, rows, init)
    orc = P.T.get_delta_nu_adjoint(0.3, k, 0.15, g.numpy())
    e = max(float(np.max(np.abs(dt.grad.numpy() - orc[2]) / np.max(np.abs(orc[2])))),
            rel(di.grad.numpy(), orc[3]), abs(float(m.grad) / orc[4] - 1))
    print("  autograd vs oracle gradient: %.3e" % e)
    assert e < RTOL
    P.free()


@pytest.mark.parametrize("cfg", ["medium", "full"])
def test_get_delta_nu_at_configuration_shapes(kn, torch_cuda, cfg):
    """The free-streaming integral at the medium and full configurations'
    shapes, delta_tot = a sqrt(P(k)) over a in [0.01, 1]: 300 k x 1000 steps
    with three degenerate 0.1 eV species (all k against the oracle), and
    757 k x 2000 steps hybrid 0.06/0.06/0.3 eV (vcrit 500 km/s, particles
    from a = 0.3333; 24 k against the oracle)."""
    if cfg == "medium":
        masses, hybrid, nk, na, sel = [0.1, 0.1, 0.1], None, 300, 1000, None
    else:
        masses, hybrid, nk, na = [0.06, 0.06, 0.3], (500.0, 0.3333), 757, 2000
        sel = np.linspace(0, nk - 1, 24).astype(int)
    P = Pair(masses, 0.3, nk, 0.01 + na / 100.0 + 0.02, hybrid=hybrid)
    logs, rows, k, init = synthetic_history(na, nk, 1.0, noise=0.0)
    P.set_history(logs, rows, init)
    print("get_delta_nu_combined", cfg, nk, "x", na, flush=True)
    g = P.gpu_get_delta_nu_combined(1.0, k)
    if sel is None:
        rc, o = P.T.get_delta_nu_combined(1.0, k)
        e = rel(g, o)
    else:
        Q = Pair(masses, 0.3, len(sel), 0.01 + na / 100.0 + 0.02, hybrid=hybrid)
        Q.set_history(logs, rows[sel], init[sel])
        rc, o = Q.T.get_delta_nu_combined(1.0, k[sel])
        e = rel(g[sel], o)
        Q.free()
    print("  %s: max rel err vs oracle %.3e" % (cfg, e))
    assert rc == 0 and e < RTOL
    P.free()


def test_history_stays_on_the_device_and_sees_host_edits(kn, torch_cuda):
    """get_delta_nu keeps the delta_tot history on the device and uploads only
    the columns that differ from what it holds: a first call uploads every
    stored step, an appended step one column, an overwritten last step one,
    an interior edit written straight into the host table (no library call)
    the columns from the edit on, a dropped step none; every call against the
    oracle on the same table."""
    masses = [0.15, 0.0, 0.0]
    P = Pair(masses, 0.3, 12, 0.01 + 80 / 100.0 + 0.02)
    logs, rows, k, init = synthetic_history(80, 12, 0.8, noise=0.02)
    n0 = 60

    def check(na, expect_cols, what):
        P.set_history(logs[:na], rows[:, :na], init)
        a = math.exp(logs[na - 1])
        g = P.gpu_get_delta_nu(a, k, 0.15)
        cols = kn.kn_dnu_last_upload_columns()
        rc, o = P.T.get_delta_nu(a, k, 0.15)
        e = rel(g, o)
        print("  %-34s Na=%d: %d columns uploaded, rel err %.2e" % (what, na, cols, e))
        assert rc == 0 and e < RTOL
        assert cols == expect_cols, (what, cols, expect_cols)

    print("device-resident history", flush=True)
    check(n0, n0, "first call")
    check(n0 + 1, 1, "one step appended")
    rows[:, n0] *= 1.01
    check(n0 + 1, 1, "last step overwritten")
    rows[3, 10] *= 1.5
    check(n0 + 1, n0 + 1 - 10, "interior entry edited")
    check(n0, 0, "last step dropped")
    check(n0 + 2, 2, "two steps appended")
    P.free()


@pytest.mark.parametrize("noise", [0.1, 1.0], ids=["108 intervals", "171 intervals"])
def test_long_interval_lists_match_oracle(kn, torch_cuda, noise):
    """Histories whose QAGs need 108 and 171 subintervals (per-entry noise on
    300 stored steps): past 64 intervals the error order leaves registers for
    the LDS list, and past limit / 2 + 2 = 102 QUADPACK keeps it only
    partially sorted.  Two calls, the first with one wavefront per item, the
    second with the long items on speculative blocks, each against the oracle
    (a different subdivision would move the result by far more than 1e-9)."""
    P = Pair([0.3, 0, 0], 0.3, 8, 0.01 + 300 / 100.0 + 0.02)
    logs, rows, k, init = synthetic_history(300, 8, 1.0, noise=noise)
    P.set_history(logs, rows, init)
    rc, o = P.T.get_delta_nu(1.0, k, 0.3)
    assert rc == 0 and P.T.d.qag_max_intervals > 102
    for call in range(2):
        g = P.gpu_get_delta_nu(1.0, k, 0.3)
        e = rel(g, o)
        print("  noise %.1f call %d: %d intervals (oracle max %d), rel err %.2e"
              % (noise, call, kn.kn_dnu_last_intervals(C.byref(C.c_int())), P.T.d.qag_max_intervals, e))
        assert e < RTOL
    P.free()
