"""(b) batched free-streaming integrals and the adjoints, in the CPU SIMT
emulation build of the library's HIP kernels (tests/emu), against the oracle.

The kernels are the product's own source (csrc/batch.hip) compiled with g++
against an emulation of the HIP subset they use: every lane is a host thread,
wavefront shuffles go through a 64-lane barrier, fresh "device" memory is
filled with garbage.  The device build keeps these kernels switched off (a
device run did not complete); this is the evidence that their arithmetic and
control flow are right.  Bars as on the device: 1e-9 relative.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import knhip
from batch_common import RTOL, Cosmo, adjoint_case, compare_adjoint, mixed_batch, rel
from conftest import ROOT
from dnu_common import Pair, synthetic_history

EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libkn_batch_emu.so")
DNU_EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libkn_dnu_emu.so")


def host_memcpy(dst, src, n, kind):
    C.memmove(dst, src, n)


@pytest.fixture(scope="module")
def emu(kn):
    if not os.path.exists(EMU_LIB):
        raise RuntimeError("emulation build missing: make -C tests/emu")
    return knhip.bind(C.CDLL(EMU_LIB))


@pytest.fixture(scope="module")
def dnu_emu(kn):
    if not os.path.exists(DNU_EMU_LIB):
        raise RuntimeError("emulation build missing: make -C tests/emu")
    return knhip.bind(C.CDLL(DNU_EMU_LIB))


@pytest.mark.parametrize("case", [
    dict(name="cspline", masses=[0.15, 0, 0], Na=40, nk=10, amax=0.4, hybrid=None),
    dict(name="linear Na=2", masses=[0.2, 0, 0], Na=2, nk=8, amax=0.03, hybrid=None),
    dict(name="hybrid", masses=[0.3, 0.06, 0.06], Na=80, nk=8, amax=0.9, hybrid=(500.0, 0.3333)),
    dict(name="noisy", masses=[0.3, 0, 0], Na=150, nk=6, amax=1.0, hybrid=None, noise=0.1),
    dict(name="massless", masses=[0.15, 0, 0], Na=20, nk=8, amax=0.2, hybrid=None, m=0.0),
    dict(name="long Na=2300 (3 adjoint parts)", masses=[0.2, 0, 0], Na=2300, nk=3, amax=0.9, hybrid=None),
    dict(name="Na=1100 nk=64 (chunked transfers)", masses=[0.2, 0, 0], Na=1100, nk=64, amax=0.9, hybrid=None),
], ids=lambda c: c["name"])
def test_emu_plain_adjoint_matches_oracle(dnu_emu, case):
    """The per-call path (csrc/delta_nu.hip: k_dnu_qag with recorded
    subdivisions, k_dnu_adj, k_dnu_adj_rows) in the emulation."""
    P = Pair(case["masses"], 0.3, case["nk"], 0.01 + case["Na"] / 100.0 + 0.02, hybrid=case["hybrid"])
    logs, rows, k, init = synthetic_history(case["Na"], case["nk"], case["amax"], noise=case.get("noise", 0.0))
    P.set_history(logs, rows, init)
    m = case.get("m", case["masses"][0])
    gpu, orc = adjoint_case(P, case["amax"], k, m, L=dnu_emu)
    if m > 0:
        compare_adjoint(case["name"], gpu, orc)
    else:
        assert np.all(gpu[1] == 0) and gpu[3] == 0.0 and rel(gpu[2], orc[3]) < 1e-14
    P.free()


def test_product_build_keeps_the_batch_switched_off(kn):
    """The device library refuses to create a batch (3005) instead of running it."""
    hh = knhip.BatchHistory()
    h = C.c_void_p()
    assert kn.kn_batch_create(C.byref(h), 1, C.byref(hh)) == 3005
    assert b"switched off" in kn.kn_last_error()


@pytest.mark.parametrize("case", [
    dict(name="cspline", masses=[0.15, 0, 0], Na=60, nk=16, amax=0.5, hybrid=None),
    dict(name="linear Na=2", masses=[0.2, 0, 0], Na=2, nk=8, amax=0.03, hybrid=None),
    dict(name="hybrid", masses=[0.3, 0.06, 0.06], Na=50, nk=12, amax=0.9, hybrid=(500.0, 0.3333)),
    dict(name="noisy", masses=[0.3, 0, 0], Na=120, nk=6, amax=1.0, hybrid=None, noise=0.1),
], ids=lambda c: c["name"])
def test_emu_get_delta_nu_adjoint_matches_oracle(emu, case):
    P = Pair(case["masses"], 0.3, case["nk"], 0.01 + case["Na"] / 100.0 + 0.02, hybrid=case["hybrid"])
    logs, rows, k, init = synthetic_history(case["Na"], case["nk"], case["amax"], noise=case.get("noise", 0.0))
    P.set_history(logs, rows, init)
    gpu, orc = adjoint_case(P, case["amax"], k, case["masses"][0], L=emu, fn="kn_batch_get_delta_nu_adjoint")
    compare_adjoint(case["name"], gpu, orc)
    P.free()


def test_emu_adjoint_massless_is_ic_only(emu):
    P = Pair([0.15, 0, 0], 0.3, 8, 0.5)
    logs, rows, k, init = synthetic_history(20, 8, 0.2)
    P.set_history(logs, rows, init)
    (dnu, db, ib, mb), orc = adjoint_case(P, 0.2, k, 0.0, L=emu, fn="kn_batch_get_delta_nu_adjoint")
    assert np.all(db == 0) and mb == 0.0
    # delta_nu here is J(k fsl) delta_nu_init down to 1e-26, on J's x^-4 tail:
    # fslength's few-ulp difference from the oracle (the device's 1/(a^2 H)
    # interpolant against the host's H, ~2e-15) comes out about four times larger.
    assert rel(dnu, orc[1]) < 5e-14 and rel(ib, orc[3]) < 5e-14
    P.free()


def test_emu_batch_mixed_forward_and_adjoint(emu):
    cosmos, specs, B = mixed_batch(4, seed=2, L=emu, memcpy=host_memcpy)
    B.solve()
    worst = 0.0
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        rc, o = T.get_delta_nu(s["amax"], s["k"], s["m"])
        assert rc == 0
        worst = max(worst, rel(B.delta_nu(i), o))
    print("  %d histories: worst forward rel err %.3e" % (B.n, worst))
    assert worst < RTOL
    g = np.random.default_rng(7).uniform(-1, 1, B.off_k[B.n])
    B.adjoint(g)
    mb = B.mnu_bar()
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        gi = g[B.off_k[i]:B.off_k[i] + s["nk"]]
        orc = T.get_delta_nu_adjoint(s["amax"], s["k"], s["m"], gi)
        compare_adjoint("hist %d (Na=%d nk=%d)" % (i, s["na"], s["nk"]),
                        (B.delta_nu(i), B.dtot_bar(i, s["na"]), B.init_bar(i), mb[i]), orc)
    B.close()
    for c in cosmos:
        c.free()


def test_emu_append_steps_and_capacity(emu):
    """Append two stored steps on the device side, solve, compare with the
    oracle on the extended histories; a third append overflows na_max and is
    reported as 3004 without writing."""
    cosmos, specs, B = mixed_batch(2, seed=4, extra=2, L=emu, memcpy=host_memcpy)
    rng = np.random.default_rng(5)
    for step in range(2):
        loga, rows = np.zeros(B.n), np.zeros(B.off_k[B.n])
        for i, s in enumerate(specs):
            new_a = math.exp(s["logs"][-1]) * 1.02
            s["logs"] = np.append(s["logs"], math.log(new_a))
            r = s["rows"][:, -1] * 1.02 * (1 + 0.01 * rng.standard_normal(s["nk"]))
            s["rows"] = np.concatenate([s["rows"], r[:, None]], axis=1)
            s["amax"] = new_a
            loga[i] = math.log(new_a)
            rows[B.off_k[i]:B.off_k[i] + s["nk"]] = r
        B.append(loga, rows)
        B.solve()
        worst = 0.0
        for i, s in enumerate(specs):
            T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
            rc, o = T.get_delta_nu(s["amax"], s["k"], s["m"])
            worst = max(worst, rel(B.delta_nu(i), o))
        print("  append step %d: worst rel err %.3e" % (step, worst))
        assert worst < RTOL
    B.append(np.zeros(B.n) + math.log(0.99), np.zeros(B.off_k[B.n]))
    assert B.status() == 3004
    B.close()
    for c in cosmos:
        c.free()


def test_emu_out_of_range_a_reports_2001(emu):
    c = Cosmo([0.15, 0, 0])
    logs, rows, k, init = synthetic_history(10, 4, 0.2)
    hh = c.history(4, 10, 10, 0.25, 0.25, 0.15)
    B = knhip.Batch([hh], keep=(c,), L=emu, memcpy=host_memcpy)
    B.load(0, logs, rows, k, init)
    with pytest.raises(RuntimeError, match="code 2001"):
        B.solve()
    B.close()
    c.free()


def test_emu_autograd_gradcheck_and_oracle_gradient(emu):
    """torch.autograd.gradcheck through the C-ABI adjoint (central differences,
    relative step 1e-6), then the gradient against the oracle's adjoint."""
    import torch
    P = Pair([0.15, 0, 0], 0.3, 4, 0.5)
    logs, rows, k, init = synthetic_history(8, 4, 0.3, noise=0.05)
    scale = np.max(np.abs(rows))  # O(1) inputs: eps = 1e-6 is then a relative step
    rows, init = rows / scale, init / scale
    P.set_history(logs, rows, init)
    state = knhip.DeltaNuState(P.d, 0.3, k, 0.15, L=emu, memcpy=host_memcpy)
    F = knhip.delta_nu_function()
    dt = torch.tensor(rows, dtype=torch.float64, requires_grad=True)
    di = torch.tensor(init, dtype=torch.float64, requires_grad=True)
    m = torch.tensor(0.15, dtype=torch.float64, requires_grad=True)
    # delta_nu is linear in delta_tot and delta_nu_init, so the differences are
    # exact up to rounding (2e-16 |delta_nu| / eps ~ 1e-9): atol 1e-8
    assert torch.autograd.gradcheck(lambda a, b, c: F.apply(a, b, c, state), (dt, di, m), eps=1e-6, atol=1e-8,
                                    rtol=1e-6, nondet_tol=1e-12)
    g = torch.linspace(-1, 1, 4, dtype=torch.float64)
    (F.apply(dt, di, m, state) * g).sum().backward()
    orc = P.T.get_delta_nu_adjoint(0.3, k, 0.15, g.numpy())
    e = max(float(np.max(np.abs(dt.grad.numpy() - orc[2]) / np.max(np.abs(orc[2])))),
            rel(di.grad.numpy(), orc[3]), abs(float(m.grad) / orc[4] - 1))
    print("  autograd vs oracle gradient: %.3e" % e)
    assert e < RTOL
    state.batch.close()
    P.free()
