"""(a) k-shell binning and (c) rescale on the GPU, through the C ABI,
against the oracle: mode counts bit-exact, P(k) and <|k|> to 1e-12 relative
(summation order differs), rescaled modes to 1e-14 relative."""
import ctypes as C
import math

import numpy as np
import pytest

import knhip
import oracle_py as O
from conftest import dptr

pytestmark = pytest.mark.gpu

LL = C.POINTER(C.c_longlong)


def _rand_grid(N, seed, nslab=None):
    rng = np.random.default_rng(seed)
    ns = N if nslab is None else nslab
    F = rng.standard_normal((ns, N, N // 2 + 1)) + 1j * rng.standard_normal((ns, N, N // 2 + 1))
    F *= np.exp(-0.05 * rng.uniform(0, 1, F.shape))
    F[0, 0, 0] = N ** 3
    return F


def gpu_bin(kn, torch, F, nrbins, startslab=0, on_device=True):
    N = F.shape[1]
    power = np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    keffs = np.zeros(nrbins)
    if on_device:
        buf = torch.from_numpy(np.ascontiguousarray(F)).cuda()
        ptr = buf.data_ptr()
    else:
        buf = np.ascontiguousarray(F)
        ptr = buf.ctypes.data
    print("bin", N, nrbins, startslab, F.shape[0], "device" if on_device else "host", flush=True)
    n = kn.total_powerspectrum(N, ptr, nrbins, startslab, F.shape[0], dptr(power), count.ctypes.data_as(LL),
                               dptr(keffs), 0)
    return n, power, count, keffs


def check_against_oracle(kn, torch, F, nrbins, startslab=0, on_device=True):
    n, p, c, k = gpu_bin(kn, torch, F, nrbins, startslab, on_device)
    on, op, oc, ok = O.total_powerspectrum(F, nrbins, startslab, F.shape[0])
    assert n == on
    np.testing.assert_array_equal(c[:n], oc[:n])
    np.testing.assert_allclose(p[:n], op[:n], rtol=1e-12)
    np.testing.assert_allclose(k[:n], ok[:n], rtol=1e-12)   # 1e5-term sums in another order
    return n, p, c, k


def test_reference_4cubed_known_answer(kn, torch_cuda):
    f = np.zeros((4, 4, 4))
    f[0:2] = 1
    f[0, 0, 0] = 2
    n, p, c, k = check_against_oracle(kn, torch_cuda, np.fft.rfftn(f), 15)
    assert n == 9 and c[0] == 6 and c[1] == 12 and c[8] == 1 and abs(k[2] - 1.73205) < 1e-5
    assert abs(p[0] - 0.254834) < 4e-7 and abs(p[1] - 0.00212722) < 5e-8 and abs(p[2] - 0.00323766) < 3e-8


# odd N: the z = N/2 plane keeps multiplicity 1 as in the reference (counts
# then total N^2 (N - 1) - 1); one bin: binsperunit = 0, every mode in bin 0
@pytest.mark.parametrize("N,nrbins", [(8, 7), (16, 15), (16, 64), (32, 33), (64, 100), (128, 64), (96, 300),
                                      (9, 5), (15, 16), (7, 3), (12, 1)])
def test_binning_matches_oracle(kn, torch_cuda, N, nrbins):
    check_against_oracle(kn, torch_cuda, _rand_grid(N, N * 7 + nrbins), nrbins)


def test_binning_host_pointer_path(kn, torch_cuda):
    check_against_oracle(kn, torch_cuda, _rand_grid(32, 5), 40, on_device=False)


def test_binning_slab_subset(kn, torch_cuda):
    full = _rand_grid(32, 11)
    check_against_oracle(kn, torch_cuda, np.ascontiguousarray(full[5:12]), 30, startslab=5)
    # slabs without the zero mode normalise by |F0|^2 = 0 like the reference (single rank)
    n, p, c, k = gpu_bin(kn, torch_cuda, np.ascontiguousarray(full[5:12]), 30, startslab=5)
    assert np.all(np.isinf(p[:n]) | np.isnan(p[:n]))


def test_binning_run_to_run(kn, torch_cuda):
    """Counts and <|k|> come from the geometry pass (bit-identical run to run);
    the wavefronts of a block add into shared bin sums in no fixed order, so
    P(k) may move in its last bits between runs (1e-14 relative at most)."""
    F = _rand_grid(96, 3)
    a = gpu_bin(kn, torch_cuda, F, 77)
    b = gpu_bin(kn, torch_cuda, F, 77)
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(a[3], b[3])
    np.testing.assert_allclose(a[1], b[1], rtol=1e-14, atol=0)


# Many bins: the tables leave LDS.  (64, 20000) and (16, 5000): thresholds,
# window and bin sums in global memory (and the geometry pass's sums too).
@pytest.mark.parametrize("N,nrbins", [(64, 20000), (16, 5000)])
def test_binning_tables_in_global_memory(kn, torch_cuda, N, nrbins):
    check_against_oracle(kn, torch_cuda, _rand_grid(N, N + nrbins), nrbins)


@pytest.mark.parametrize("N,nrbins,nslab", [(4096, 2048, 16), (8192, 8000, 1)])
def test_binning_large_grid_slab_subset(kn, torch_cuda, N, nrbins, nslab):
    """Gadget passes nk = PMGRID/2 bins (0002-Add-hooks-to-kspace-neutrino-code.patch:59):
    PMGRID 4096 with 2048 bins on a 16-slab share (2.1 GB; the window table
    read from global memory), and 8000 bins on one slab of an 8192 grid
    (537 MB; the geometry pass's sums in global memory), from slab 0, against
    the oracle on the same bytes."""
    torch = torch_cuda
    buf = torch.empty((nslab, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
    assert kn.kn_synth_grid(buf.data_ptr(), N, 0, nslab, 512000.0, 1e-6, 0.96, 0.2 / 1000.0, 11) == 0
    power, keffs = np.zeros(nrbins), np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    print("large grid %d^3, %d bins, %d slabs" % (N, nrbins, nslab), flush=True)
    n = kn.total_powerspectrum(N, buf.data_ptr(), nrbins, 0, nslab, dptr(power), count.ctypes.data_as(LL),
                               dptr(keffs), 0)
    F = buf.cpu().numpy()
    del buf
    on, op, oc, ok = O.total_powerspectrum(F, nrbins, 0, nslab)
    assert n == on
    np.testing.assert_array_equal(count[:n], oc[:n])
    np.testing.assert_allclose(power[:n], op[:n], rtol=1e-12)
    np.testing.assert_allclose(keffs[:n], ok[:n], rtol=1e-12)


def test_synth_grid_binning_256(kn, torch_cuda):
    """C5 grid size: synthetic BBKS field generated on the device, parity on the same bytes."""
    torch = torch_cuda
    N = 256
    buf = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
    assert kn.kn_synth_grid(buf.data_ptr(), N, 0, N, 512000.0, 1e-6, 0.96, 0.2 / 1000.0, 5) == 0
    F = buf.cpu().numpy()
    assert F[0, 0, 0] == N ** 3
    n, p, c, k = check_against_oracle(kn, torch, F, 128)
    assert c[:n].sum() == N ** 3 - 1


@pytest.mark.parametrize("N,nrbins", [(512, 300), (1024, 1024)])
def test_binning_full_size_invariants(kn, torch_cuda, N, nrbins):
    """Full sizes (512^3, 1024^3): counts sum to N^3 - 1, and sum_b P_b N_b |F0|^2 equals
    sum over modes of multiplicity W^-4 |F|^2 computed by an independent torch reduction."""
    torch = torch_cuda
    buf = torch.empty((N, N, N // 2 + 1), dtype=torch.complex128, device="cuda")
    assert kn.kn_synth_grid(buf.data_ptr(), N, 0, N, 512000.0, 1e-6, 0.96, 0.2 / 1000.0, 3) == 0
    power = np.zeros(nrbins)
    count = np.zeros(nrbins, dtype=np.int64)
    keffs = np.zeros(nrbins)
    print("full-size bin", N, nrbins, flush=True)
    n = kn.total_powerspectrum(N, buf.data_ptr(), nrbins, 0, N, dptr(power), count.ctypes.data_as(LL), dptr(keffs), 0)
    assert count[:n].sum() == N ** 3 - 1
    kv = torch.arange(N, device="cuda", dtype=torch.float64)
    kv = torch.where(kv <= N // 2, kv, kv - N)
    iw = torch.where(kv == 0, torch.ones_like(kv), math.pi * kv / (N * torch.sin(math.pi * kv / N)))
    kz = torch.arange(N // 2 + 1, device="cuda")
    mult = torch.where((kz == 0) | (kz == N // 2), 1.0, 2.0).double()
    total = 0.0
    for x in range(0, N, 64):
        sl = buf[x:x + 64]
        w = (iw[x:x + 64, None, None] * iw[None, :, None] * iw[None, None, :N // 2 + 1]) ** 4 * mult
        if x == 0:
            w[0, 0, 0] = 0.0                                   # the zero mode is not binned
        total += float((sl.abs() ** 2 * w).sum())
    m2 = abs(complex(buf[0, 0, 0].item())) ** 2
    assert np.sum(power[:n] * count[:n]) * m2 == pytest.approx(total, rel=1e-10)
    assert np.all(np.diff(keffs[:n]) > 0)


# ------------------------------------------------------------------ rescale

@pytest.mark.parametrize("N,nb", [(16, 12), (64, 64), (128, 300), (9, 4), (15, 7)])
def test_rescale_matches_oracle(kn, torch_cuda, N, nb):
    torch = torch_cuda
    F = _rand_grid(N, N + nb)
    box = 512000.0
    kmin, kmax = 2 * math.pi / box, math.sqrt(3) * math.pi * N / box
    logkk = np.sort(np.log(np.exp(np.random.default_rng(nb).uniform(math.log(kmin) + 0.2, math.log(kmax), nb))))
    ratio = 0.3 + 0.1 * np.sin(np.arange(nb))
    d = torch.from_numpy(F.copy()).cuda()
    tab = torch.from_numpy(np.concatenate([logkk, ratio])).cuda()
    print("rescale", N, nb, flush=True)
    rc = kn.kn_rescale_device(d.data_ptr(), N, 0, N, box, tab.data_ptr(), tab.data_ptr() + 8 * nb, nb, 0.07, None)
    torch.cuda.synchronize()
    assert rc == 0
    G = d.cpu().numpy()
    rc, W = O.rescale(F, box, logkk, ratio, 0.07)
    assert rc == 0
    assert G[0, 0, 0] == F[0, 0, 0]
    np.testing.assert_allclose(G.view(np.float64), W.view(np.float64), rtol=1e-14, atol=0)


def test_rescale_constant_ratio_is_uniform(kn, torch_cuda):
    torch = torch_cuda
    N = 64
    F = _rand_grid(N, 2)
    d = torch.from_numpy(F.copy()).cuda()
    tab = torch.tensor([-12.0, -2.0, 0.5, 0.5], dtype=torch.float64, device="cuda")
    assert kn.kn_rescale_device(d.data_ptr(), N, 0, N, 512000.0, tab.data_ptr(), tab.data_ptr() + 16, 2, 0.2, None) == 0
    G = d.cpu().numpy()
    want = F * 1.1
    want[0, 0, 0] = F[0, 0, 0]
    np.testing.assert_array_equal(G, want)


# ------------------------------------------------- rescale + Green's function

@pytest.mark.parametrize("N,nb,slab0,nslab", [(64, 64, 0, 64), (33, 20, 0, 33), (48, 30, 7, 12), (128, 300, 0, 128)])
def test_rescale_green_matches_the_two_passes(kn, torch_cuda, N, nb, slab0, nslab):
    """One sweep of kn_rescale_green_device against the oracle's two passes:
    the neutrino rescale (interface_gadget.c:163-188) then the PM Green's
    function with CiC deconvolution (Gadget-2's pmforce_periodic, after the
    hook of 0002-Add-hooks-to-kspace-neutrino-code.patch:116-125), to 1e-14;
    and the Green's function alone (nbins 0)."""
    torch = torch_cuda
    F = _rand_grid(N, N + nb, nslab=nslab)
    if slab0:
        F[0, 0, 0] = 1.5 + 0.5j
    box, asmth = 512000.0, 1.25 * 512000.0 / N
    kmin, kmax = 2 * math.pi / box, math.sqrt(3) * math.pi * N / box
    logkk = np.sort(np.log(np.exp(np.random.default_rng(nb).uniform(math.log(kmin) + 0.2, math.log(kmax), nb))))
    ratio = 0.3 + 0.1 * np.sin(np.arange(nb))
    tab = torch.from_numpy(np.concatenate([logkk, ratio])).cuda()
    for with_nu in (True, False):
        d = torch.from_numpy(F.copy()).cuda()
        print("rescale+green N=%d nb=%d slabs [%d,%d) nu=%d" % (N, nb, slab0, slab0 + nslab, with_nu), flush=True)
        if with_nu:
            rc = kn.kn_rescale_green_device(d.data_ptr(), N, slab0, nslab, box, tab.data_ptr(), tab.data_ptr() + 8 * nb,
                                            nb, 0.07, asmth, None)
        else:
            rc = kn.kn_rescale_green_device(d.data_ptr(), N, slab0, nslab, box, None, None, 0, 0.0, asmth, None)
        torch.cuda.synchronize()
        assert rc == 0
        G = d.cpu().numpy()
        W = F
        if with_nu:
            rc2, W = O.rescale(F, box, logkk, ratio, 0.07, startslab=slab0)
            assert rc2 == 0
        W = O.green(W, box, asmth, startslab=slab0)
        np.testing.assert_allclose(G.view(np.float64), W.view(np.float64), rtol=1e-14, atol=0)
        if slab0 == 0:
            assert G[0, 0, 0] == 0
