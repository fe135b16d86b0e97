"""Parity at the configurations' full shapes, on the GPU through the C ABI.

* (a) and (c) at 512^3 (300 bins) and 1024^3 (1024 bins): the device bins and
  rescales the whole grid; three slabs (the first, the middle and the last)
  are copied before and after and compared with the oracle: per-slab raw bin
  sums (powerspectrum.c:56-89) with counts bit-exact and sums to 1e-12, and
  the rescaled modes (interface_gadget.c:163-188) to 1e-14.
* The adjoint of get_delta_nu at the differentiation shape (1024 k x 2000
  steps), per call and for a 64-history group, with dL/d delta_nu nonzero
  only on sampled k: then d/d mnu is the sum over those k alone, and the
  oracle's adjoint of a table holding just those rows gives every gradient.
* QAG failure: a history whose per-entry noise makes the oracle's QAG fail
  (GSL error -> terminate 2001 in the reference) fails on the device too.
"""
import ctypes as C
import math

import numpy as np
import pytest

import knhip
import oracle_py as O
from batch_common import RTOL, Cosmo, rel
from conftest import dptr
from dnu_common import Pair, synthetic_history

pytestmark = pytest.mark.gpu

LL = C.POINTER(C.c_longlong)
DP = C.POINTER(C.c_double)


@pytest.mark.parametrize("N,nrbins", [(512, 300), (1024, 1024)])
def test_binning_and_rescale_full_size_sampled_slabs(kn, torch_cuda, N, nrbins):
    torch = torch_cuda
    box = 512000.0
    rowlen = N // 2 + 1
    buf = torch.empty((N, N, rowlen), dtype=torch.complex128, device="cuda")
    assert kn.kn_synth_grid(buf.data_ptr(), N, 0, N, box, 1e-6, 0.96, 0.2 / 1000.0, 3) == 0
    slabs = [0, N // 2, N - 1]
    before = {s: buf[s:s + 1].cpu().numpy() for s in slabs}
    # per-slab raw sums: the device sweep on one slab against the oracle's loop
    d_sum = torch.zeros(nrbins + 1, dtype=torch.float64, device="cuda")
    for s in slabs:
        print("full size %d^3: bin sums of slab %d" % (N, s), flush=True)
        ptr = buf.data_ptr() + 16 * s * N * rowlen
        assert kn.kn_bin_power_device(ptr, N, nrbins, s, 1, d_sum.data_ptr(), d_sum.data_ptr() + 8 * nrbins,
                                      None) == 0
        torch.cuda.synchronize()
        dev = d_sum[:nrbins].cpu().numpy()
        op, oc, ok = O.bin_sums(before[s], nrbins, startslab=s)
        nz = oc > 0
        assert np.all(dev[~nz] == 0)
        np.testing.assert_allclose(dev[nz], op[nz], rtol=1e-12)
        # counts and keffs of the slab through the reference entry point
        power, keffs = np.zeros(nrbins), np.zeros(nrbins)
        count = np.zeros(nrbins, dtype=np.int64)
        n = kn.total_powerspectrum(N, ptr, nrbins, s, 1, dptr(power), count.ctypes.data_as(LL), dptr(keffs), 0)
        assert n == int(nz.sum())
        np.testing.assert_array_equal(count[:n], oc[nz])
        np.testing.assert_allclose(keffs[:n], ok[nz] / oc[nz], rtol=1e-12)
    # the whole grid binned, a ratio table on its k, the whole grid rescaled
    power, keffs = np.zeros(nrbins), np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    n = kn.total_powerspectrum(N, buf.data_ptr(), nrbins, 0, N, dptr(power), count.ctypes.data_as(LL), dptr(keffs), 0)
    assert count[:n].sum() == N ** 3 - 1
    logkk = np.log(keffs[:n] * 2 * math.pi / box)
    kfs = 2 * math.pi / box * 10.0
    ratio = 0.3 / (1 + (np.exp(logkk) / kfs) ** 2) + 0.02 * np.sin(3 * logkk)
    norm = 0.07
    tab = torch.from_numpy(np.concatenate([logkk, ratio])).cuda()
    print("full size %d^3: rescale" % N, flush=True)
    assert kn.kn_rescale_device(buf.data_ptr(), N, 0, N, box, tab.data_ptr(), tab.data_ptr() + 8 * n, n, norm,
                                None) == 0
    torch.cuda.synchronize()
    for s in slabs:
        after = buf[s:s + 1].cpu().numpy()
        rc, want = O.rescale(before[s], box, logkk, ratio, norm, startslab=s)
        assert rc == 0
        np.testing.assert_allclose(after.view(np.float64), want.view(np.float64), rtol=1e-14, atol=0)
    assert buf[0, 0, 0].item() == before[0][0, 0, 0]


def _sampled_adjoint_check(dnu, dbar_rows, ibar, mbar, T, a, k_s, m, g_s):
    """Device results on the sampled k against the oracle table T of those rows."""
    orc = T.get_delta_nu_adjoint(a, k_s, m, g_s)
    assert orc[0] == 0
    e_f = rel(dnu, orc[1])
    scale = np.maximum(np.max(np.abs(orc[2]), axis=1, keepdims=True), 1e-300)
    e_y = float(np.max(np.abs(dbar_rows - orc[2]) / scale))
    e_i = float(np.max(np.abs(ibar - orc[3]) / np.max(np.abs(orc[3]))))
    e_m = abs(mbar / orc[4] - 1)
    return e_f, e_y, e_i, e_m


def test_get_delta_nu_adjoint_at_differentiation_shape(kn, torch_cuda):
    """kn_get_delta_nu_adjoint on one 1024 k x 2000 step history (m = 0.3 eV,
    delta_tot = a sqrt(P(k)) with 5% noise per k), dL/d delta_nu on 8 k."""
    nk, na, m = 1024, 2000, 0.3
    P = Pair([m, 0.0, 0.0], 0.3, nk, 0.01 + na / 100.0 + 0.02)
    logs, rows, k, init = synthetic_history(na, nk, 1.0, noise=0.0)
    # noise drawn per k (smooth along a): drawn per (k, a) entry at 2000 steps
    # it makes the QAG fail, as in the failure test below
    rows = rows * (1 + 0.05 * np.random.default_rng(8).standard_normal(nk))[:, None]
    init = 0.4 * rows[:, 0]
    P.set_history(logs, rows, init)
    sel = np.linspace(0, nk - 1, 8).astype(int)
    g = np.zeros(nk)
    g[sel] = np.random.default_rng(7).uniform(-1, 1, len(sel))
    dnu, db, ib, mb = np.zeros(nk), np.zeros(nk * na), np.zeros(nk), C.c_double()
    kk = np.ascontiguousarray(k)
    print("adjoint at 1024 k x 2000 steps", flush=True)
    assert kn.kn_get_delta_nu_adjoint(C.byref(P.d), 1.0, kk.ctypes.data_as(DP), m, g.ctypes.data_as(DP),
                                      dnu.ctypes.data_as(DP), db.ctypes.data_as(DP), ib.ctypes.data_as(DP),
                                      C.byref(mb)) == 0
    db = db.reshape(nk, na)
    unsel = np.setdiff1d(np.arange(nk), sel)
    assert np.all(db[unsel] == 0) and np.all(ib[unsel] == 0)
    Q = Pair([m, 0.0, 0.0], 0.3, len(sel), 0.01 + na / 100.0 + 0.02)
    Q.set_history(logs, rows[sel], init[sel])
    errs = _sampled_adjoint_check(dnu[sel], db[sel], ib[sel], mb.value, Q.T, 1.0, k[sel], m, g[sel])
    print("  forward %.2e  d/d delta_tot %.2e  d/d init %.2e  d/d mnu %.2e" % errs)
    assert max(errs) < RTOL
    Q.free()
    P.free()


def test_group_differentiation_workload(kn, torch_cuda):
    """The differentiation batch: 64 histories of 1024 k x 2000 steps, m ~
    U[0.06, 0.6] eV, solved and differentiated in one group; 8 of the
    histories checked on 2 sampled k each (forward, all gradients)."""
    rng = np.random.default_rng(5000)
    nk, na, nh = 1024, 2000, 64
    c = Cosmo([0.3, 0.0, 0.0])
    logs = np.linspace(math.log(0.01), 0.0, na)
    k = np.exp(np.linspace(math.log(2 * math.pi / 512), math.log(math.sqrt(3) * math.pi * 1024 / 512), nk)) / 1e3
    masses = rng.uniform(0.06, 0.6, nh)
    amp = np.sqrt(1e9 * (k * 1000) ** 0.96)
    hists, data = [], []
    for h in range(nh):
        rows = np.outer(amp * (1 + 0.05 * rng.standard_normal(nk)), np.exp(logs))
        hists.append(c.history(nk, na, na, 1.0, 1.0, float(masses[h])))
        data.append(rows)
    B = knhip.Batch(hists, keep=(c,))
    for h in range(nh):
        B.load(h, logs, data[h], k, 0.4 * data[h][:, 0])
    checked = np.linspace(0, nh - 1, 8).astype(int)
    sel = {h: np.array(sorted(rng.choice(nk, 2, replace=False))) for h in checked}
    g = np.zeros(B.off_k[nh])
    for h in checked:
        g[B.off_k[h] + sel[h]] = rng.uniform(-1, 1, 2)
    print("group 64 x 1024 k x 2000 steps", flush=True)
    B.solve()
    B.adjoint(g)
    mb = B.mnu_bar()
    assert np.all(mb[np.setdiff1d(np.arange(nh), checked)] == 0)
    worst = [0.0] * 4
    for h in checked:
        s = sel[h]
        T = c.oracle(len(s), logs, data[h][s], 0.4 * data[h][s, 0])
        errs = _sampled_adjoint_check(B.delta_nu(h)[s], B.dtot_bar(h, na)[s], B.init_bar(h)[s], mb[h], T, 1.0,
                                      k[s], float(masses[h]), g[B.off_k[h] + s])
        worst = [max(w, e) for w, e in zip(worst, errs)]
        del T
    print("  forward %.2e  d/d delta_tot %.2e  d/d init %.2e  d/d mnu %.2e" % tuple(worst))
    assert max(worst) < RTOL
    B.close()
    c.free()


def test_qag_failure_reported_like_the_oracle(kn, torch_cuda):
    """10% noise drawn per (k, a) entry over 1000 steps: the oracle's QAG
    fails on the low k (GSL error -> terminate 2001) and succeeds on the high
    k.  In a group the failing history is reported (2001, named), the others
    still match; the per-call adjoint returns 2001 on the failing table."""
    c = Cosmo([0.3, 0.0, 0.0])
    rng = np.random.default_rng(5)
    nk, na = 6, 1000
    logs = np.linspace(math.log(0.01), 0.0, na)
    k = np.geomspace(2e-5, 1e-2, nk)
    noisy = np.outer(np.ones(nk), np.exp(logs)) * (1 + 0.1 * rng.standard_normal((nk, na)))
    smooth = np.outer(np.linspace(1, 2, nk), np.exp(logs))
    fails = []
    for ik in range(nk):
        T = c.oracle(1, logs, noisy[ik:ik + 1], 0.4 * noisy[ik:ik + 1, 0])
        fails.append(T.get_delta_nu(1.0, k[ik:ik + 1], 0.3)[0] == 2001)
    assert any(fails) and not all(fails)
    good = [i for i in range(nk) if not fails[i]]
    # group: [smooth, noisy low k (fails), noisy high k only (passes)]
    cases = [(smooth, k), (noisy, k), (noisy[good], k[good])]
    hists = [c.history(len(kk), na, na, 1.0, 1.0, 0.3) for _, kk in cases]
    B = knhip.Batch(hists, keep=(c,))
    for h, (rows, kk) in enumerate(cases):
        B.load(h, logs, rows, kk, 0.4 * rows[:, 0])
    print("QAG failure parity", flush=True)
    with pytest.raises(RuntimeError, match="code 2001.*history 1"):
        B.solve()
    for h in (0, 2):
        rows, kk = cases[h]
        T = c.oracle(len(kk), logs, rows, 0.4 * rows[:, 0])
        rc, o = T.get_delta_nu(1.0, kk, 0.3)
        assert rc == 0 and rel(B.delta_nu(h), o) < RTOL
    B.close()
    P = Pair([0.3, 0.0, 0.0], 0.3, nk, 0.01 + na / 100.0 + 0.02)
    P.set_history(logs, noisy, 0.4 * noisy[:, 0])
    gb = np.ones(nk)
    dnu, db, ib, mb = np.zeros(nk), np.zeros(nk * na), np.zeros(nk), C.c_double()
    kk = np.ascontiguousarray(k)
    rc = kn.kn_get_delta_nu_adjoint(C.byref(P.d), 1.0, kk.ctypes.data_as(DP), 0.3, gb.ctypes.data_as(DP),
                                    dnu.ctypes.data_as(DP), db.ctypes.data_as(DP), ib.ctypes.data_as(DP),
                                    C.byref(mb))
    assert rc == 2001
    P.free()
    c.free()
