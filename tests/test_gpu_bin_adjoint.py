"""Adjoint of (a) on the GPU, through the C ABI: kn_bin_power_adjoint against
the oracle's adjoint (which central differences of the reference forward pin,
tests/test_oracle_pins.py), size-independent identities at full size, and
torch.autograd.gradcheck of the binding.

Bars: 1e-12 relative to the largest |dL/dF| entry (the kernel forms the
window weight in another order than the oracle); the homogeneity identity
sum Re(conj(F) dL/dF) = 0 to 1e-10 of its zero-mode term.
"""
import ctypes as C

import numpy as np
import pytest

import knhip
import oracle_py as O
from test_gpu_binning import _rand_grid, gpu_bin

pytestmark = pytest.mark.gpu
dp = C.POINTER(C.c_double)


def gpu_adjoint(kn, torch, F, nrbins, g):
    N = F.shape[1]
    d = torch.from_numpy(np.ascontiguousarray(F)).cuda()
    fb = torch.empty_like(d)
    gg = np.ascontiguousarray(g, dtype=np.float64)
    print("bin adjoint", N, nrbins, flush=True)
    rc = kn.kn_bin_power_adjoint(d.data_ptr(), N, nrbins, 0, N, gg.ctypes.data_as(dp), len(gg), fb.data_ptr(), None)
    assert rc == 0, kn.kn_last_error()
    return fb.cpu().numpy()


# (32, 20000): the adjoint sweep reads its tables from global memory
@pytest.mark.parametrize("N,nrbins", [(8, 12), (16, 20), (32, 40), (48, 64), (9, 5), (15, 16), (32, 20000)])
def test_bin_adjoint_matches_oracle(kn, torch_cuda, N, nrbins):
    F = _rand_grid(N, seed=N)
    F[0, 0, 0] = N ** 3 * (1 - 0.2j)
    n = O.total_powerspectrum(F, nrbins)[0]
    g = np.random.default_rng(N).uniform(-1, 1, n)
    fb = gpu_adjoint(kn, torch_cuda, F, nrbins, g)
    rc, ofb = O.bin_power_adjoint(F, nrbins, g)
    assert rc == 0
    e = float(np.max(np.abs(fb - ofb)) / np.max(np.abs(ofb)))
    print("  N=%d: max err %.2e of max |dL/dF|" % (N, e))
    assert e < 1e-12


def test_bin_adjoint_rejects_wrong_nout(kn, torch_cuda):
    F = _rand_grid(8, seed=1)
    n = O.total_powerspectrum(F, 12)[0]
    d = torch_cuda.from_numpy(F).cuda()
    fb = torch_cuda.empty_like(d)
    g = np.zeros(n + 1)
    assert kn.kn_bin_power_adjoint(d.data_ptr(), 8, 12, 0, 8, g.ctypes.data_as(dp), n + 1, fb.data_ptr(), None) == 3002
    assert kn.kn_bin_power_adjoint(d.data_ptr(), 8, 12, 2, 6, g.ctypes.data_as(dp), n, fb.data_ptr(), None) == 3002


@pytest.mark.parametrize("N,nrbins", [(512, 512)])
def test_bin_adjoint_full_size_identities(kn, torch_cuda, N, nrbins):
    """At 512^3 (size-independent checks): P is homogeneous of degree 0 in F,
    so sum_m Re(conj F_m dL/dF_m) = 0 with the zero mode holding -2 sum g P;
    and the adjoint is linear in g."""
    torch = torch_cuda
    d = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
    assert kn.kn_synth_grid(d.data_ptr(), N, 0, N, 512000.0, 1e3, 0.96, 0.2 / 1000.0, 11) == 0
    power, keffs = np.zeros(nrbins), np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    n = kn.total_powerspectrum(N, d.data_ptr(), nrbins, 0, N, power.ctypes.data_as(dp),
                               count.ctypes.data_as(C.POINTER(C.c_longlong)), keffs.ctypes.data_as(dp), 0)
    rng = np.random.default_rng(2)
    g1, g2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    out = []
    for g in (g1, g2, g1 + g2):
        fb = torch.empty_like(d)
        gg = np.ascontiguousarray(g)
        print("bin adjoint full size", N, flush=True)
        assert kn.kn_bin_power_adjoint(d.data_ptr(), N, nrbins, 0, N, gg.ctypes.data_as(dp), n, fb.data_ptr(),
                                       None) == 0
        out.append(fb)
    dots = (d.conj() * out[0]).real
    zero = float(dots[0, 0, 0])
    total = float(dots.sum())
    print("  zero-mode term %.6e (expected %.6e), total %.3e" % (zero, -2 * np.dot(g1, power[:n]), total))
    assert zero == pytest.approx(-2 * np.dot(g1, power[:n]), rel=1e-12)
    assert abs(total) < 1e-10 * abs(zero)
    lin = float((out[2] - out[0] - out[1]).abs().max()) / float(out[2].abs().max())
    print("  linearity in g: %.2e" % lin)
    assert lin < 1e-14


def test_power_spectrum_autograd_gradcheck(kn, torch_cuda):
    torch = torch_cuda
    N, nrbins = 8, 10
    F = _rand_grid(N, seed=5)
    F[0, 0, 0] = 2.0 * (1 + 0.5j)  # an O(1) zero mode keeps every input O(1): eps 1e-6 is relative
    field = torch.from_numpy(np.ascontiguousarray(F).view(np.float64).reshape(N, N, N // 2 + 1, 2)).cuda()
    field.requires_grad_(True)
    Fn = knhip.power_spectrum_function()
    print("bin gradcheck", flush=True)
    assert torch.autograd.gradcheck(lambda f: Fn.apply(f, nrbins), (field,), eps=1e-6, atol=1e-8, rtol=1e-6)
