"""(f)4 on the GPU: the background service (kn_background), Omega_nu(a) and
H(a) tabulated on the device, against the oracle, and the integrals on H
panels built from it.

* The density tables of rho_nu_init (omega_nu_single.c:117-153: 200 log-a
  points, QAG-GK61 at 1e-9, one wavefront each), Omega_nu(a) (get_omega_nu,
  :55-67) and H(a) (gadget-2 patch 0003:38-55; for a flat background the
  reference test's form, delta_tot_table_test.c:25-45, which the oracle's H
  uses): within 1e-12 relative.
* Groups whose histories take H from a background (panels and split panels
  built on the device) and the per-call get_delta_nu with the process's H set
  to a background: within 1e-9 of the oracle (the path's bar), and the
  per-call path no longer calls the host's hubble_function.
"""
import ctypes as C
import math
import time

import numpy as np
import pytest

import knhip
import oracle_py as O
from batch_common import RTOL, append_flow, mixed_batch, rel
from conftest import UNIT_TIME
from dnu_common import Pair, synthetic_history

pytestmark = pytest.mark.gpu

H0 = 3.24077929e-18 * UNIT_TIME


def _grid():
    a = list(np.exp(np.linspace(np.log(0.005), 0.0, 4001)))
    kT = 8.61734e-5 * (4 / 11.) ** (1 / 3.) * 1.00328 * 2.7255
    for m in (0.06, 0.3):
        for x in (100 * kT / m, 0.01 / 1.2):
            a += [x * (1 - 1e-9), x, x * (1 + 1e-9)]
    return np.array(sorted(a))


@pytest.mark.parametrize("masses", [[0.06, 0.06, 0.3], [0.15, 0, 0], [0.1, 0.1, 0.1], [0, 0, 0]],
                         ids=["0.06-0.06-0.3", "0.15", "3x0.1", "massless"])
def test_background_matches_the_oracle(kn, torch_cuda, masses):
    torch = torch_cuda
    print("background %s" % masses, flush=True)
    bg = knhip.Background(0.3, 0.7, H0, masses, a0=0.01)
    om = O.omega_nu(masses, a0=0.01)
    O.lib().orc_init_hubble(C.byref(om), 0.3, UNIT_TIME)
    for i in range(3):
        t = bg.rho_table(i)
        assert (t is not None) == bool(om.tab[i].have_table), i
        if t is None:
            continue
        assert np.allclose(t[0], np.array(om.tab[i].loga), rtol=1e-15, atol=0)
        e = rel(t[1], np.array(om.tab[i].rhonu))
        print("  species %d (m %g): density table max rel err %.2e" % (i, masses[i], e))
        assert e < 1e-12
    a = _grid()
    da = torch.from_numpy(a).cuda()
    o = np.array([O.lib().orc_get_omega_nu(C.byref(om), x) for x in a])
    h = np.array([O.lib().orc_hubble(x) for x in a])
    e_o = rel(bg.omega_nu(da).cpu().numpy(), o)
    e_h = rel(bg.hubble(da).cpu().numpy(), h)
    e_hh = rel(np.array([bg.hubble_host(x) for x in a]), h)
    print("  Omega_nu %.2e, H device %.2e, H host %.2e" % (e_o, e_h, e_hh))
    assert e_o < 1e-12 and e_h < 1e-12 and e_hh < 1e-12
    bg.close()


@pytest.mark.parametrize("extra", [0, 3])
def test_group_on_a_background(kn, torch_cuda, extra):
    """48 histories over three cosmologies (one hybrid), their H panels built
    on the device from backgrounds: forward and adjoint against the oracle,
    forward against the same group on the host's H."""
    print("group on backgrounds, extra capacity %d" % extra, flush=True)
    t0 = time.perf_counter()
    cosmos, specs, B = mixed_batch(16, seed=2, extra=extra, background=True)
    t1 = time.perf_counter()
    cosmos_h, specs_h, Bh = mixed_batch(16, seed=2, extra=extra)
    t2 = time.perf_counter()
    B.solve()
    Bh.solve()
    worst = 0.0
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        rc, o = T.get_delta_nu(s["amax"], s["k"], s["m"])
        assert rc == 0
        worst = max(worst, rel(B.delta_nu(i), o))
    diff = max(rel(B.delta_nu(i), Bh.delta_nu(i)) for i in range(B.n))
    print("  %d histories: worst rel err %.3e vs the oracle, %.3e vs host-H group; set-up %.3fs (host H %.3fs)"
          % (B.n, worst, diff, t1 - t0, t2 - t1))
    assert worst < RTOL and diff < RTOL
    g = np.random.default_rng(7).uniform(-1, 1, B.off_k[B.n])
    B.adjoint(g)
    mb = B.mnu_bar()
    for i, s in enumerate(specs):
        T = cosmos[s["c"]].oracle(s["nk"], s["logs"], s["rows"], s["init"])
        gi = g[B.off_k[i]:B.off_k[i] + s["nk"]]
        _, odnu, odb, oib, omb = T.get_delta_nu_adjoint(s["amax"], s["k"], s["m"], gi)
        db = B.dtot_bar(i, s["na"])
        scale = np.maximum(np.max(np.abs(odb), axis=1, keepdims=True), 1e-300)
        assert float(np.max(np.abs(db - odb) / scale)) < RTOL
        assert rel(B.init_bar(i), oib) < RTOL if np.all(oib != 0) else True
        assert (abs(mb[i] / omb - 1) if omb != 0 else abs(mb[i])) < RTOL
    for X, cs in ((B, cosmos), (Bh, cosmos_h)):
        X.close()
        for c in cs:
            c.free()


@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_append_flow_on_a_background(kn, torch_cuda, mode):
    print("append flow on backgrounds, %s" % mode, flush=True)
    s = torch_cuda.cuda.Stream() if mode == "graph" else None  # a graph is captured on a stream of its own
    worst = append_flow(mode=mode, stream=C.c_void_p(s.cuda_stream) if s is not None else None, background=True)
    print("  worst rel err over the steps: %.2e" % worst)
    assert worst < RTOL


def test_get_delta_nu_on_the_process_background(kn, host, torch_cuda):
    """get_delta_nu_combined (hybrid 0.06/0.06/0.3 eV, 600 steps) with the
    process's H set to a background: the H panels come from the device, the
    host's hubble_function is not called, and the result matches the oracle;
    with the background unset the host's H is used again."""
    print("per-call get_delta_nu on the process background", flush=True)
    masses = [0.06, 0.06, 0.3]
    P = Pair(masses, 0.3, 16, 6.5, hybrid=(500.0, 0.3333))
    logs, rows, k, init = synthetic_history(600, 16, 1.0, noise=0.0)
    P.set_history(logs, rows, init)
    rc, o = P.T.get_delta_nu_combined(1.0, k)
    assert rc == 0
    bg = knhip.Background(0.3, 0.7, H0, masses, a0=0.01)
    bg.set_process(True)
    kn.kn_htable_eval_host.restype = C.c_double
    kn.kn_htable_eval_host.argtypes = [C.c_double]
    try:
        n0 = host.kn_test_hubble_calls()
        g = P.gpu_get_delta_nu_combined(1.0, k)
        g2 = P.gpu_get_delta_nu_combined(1.0, k)
        calls = host.kn_test_hubble_calls() - n0
        e = rel(g, o)
        # the device-built panels read back on the host, against the oracle's H
        xs = np.random.default_rng(1).uniform(math.log(0.01), 0.0, 3000)
        e_h = max(abs(kn.kn_htable_eval_host(x) / O.lib().orc_hubble(math.exp(x)) - 1) for x in xs)
        print("  rel err %.2e vs the oracle, repeat bit-identical %s, host H calls %d, H panels %.2e"
              % (e, np.array_equal(g, g2), calls, e_h))
        assert calls == 0 and e < RTOL and np.array_equal(g, g2) and e_h < 1e-12
    finally:
        bg.set_process(False)
    n0 = host.kn_test_hubble_calls()
    g3 = P.gpu_get_delta_nu_combined(1.0, k)
    assert host.kn_test_hubble_calls() > n0 and rel(g3, o) < RTOL
    bg.close()
    P.free()
