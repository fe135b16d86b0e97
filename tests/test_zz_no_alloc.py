"""kn_bin_power_device_noalloc (the bin sweep with its tables and partial
sums in the library's fixed device array, no allocation) against
kn_bin_power_device on the same grid, through the C ABI: two calls in a row,
the second on a smaller grid with fewer bins, each matching the allocating
entry point (raw bin sums to 1e-14: the per-block sums may differ in their
last bit between sweeps) and the oracle's raw sums (1e-12); then a third
call back at the first geometry (its tables copied again), and sizes beyond
the array's limits rejected with 3004."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def _grid(N, seed):
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    F[0, 0, 0] = N ** 3
    return F


def _sums(kn, torch, fn, buf, N, nrbins):
    out = torch.zeros(nrbins + 1, dtype=torch.float64, device="cuda")
    rc = getattr(kn, fn)(buf.data_ptr(), N, nrbins, 0, N, out.data_ptr(), out.data_ptr() + 8 * nrbins, None)
    torch.cuda.synchronize()
    assert rc == 0, kn.kn_last_error()
    return out.cpu().numpy()


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_binning_matches_allocating_entry(kn, torch_cuda, order):
    torch = torch_cuda
    for N, nrbins, seed in ((96, 200, 1), (40, 33, 2), (96, 200, 3)):
        F = _grid(N, seed)
        buf = torch.from_numpy(F).cuda()
        print("no-alloc binning N=%d nrbins=%d" % (N, nrbins), flush=True)
        a = _sums(kn, torch, "kn_bin_power_device_noalloc", buf, N, nrbins)
        b = _sums(kn, torch, "kn_bin_power_device", buf, N, nrbins)
        op, oc, _ = O.bin_sums(F, nrbins)
        np.testing.assert_allclose(a[:nrbins], b[:nrbins], rtol=1e-14, atol=0)
        assert a[nrbins] == b[nrbins] == abs(F[0, 0, 0]) ** 2
        nz = oc > 0
        assert np.all(a[:nrbins][~nz] == 0)
        np.testing.assert_allclose(a[:nrbins][nz], op[nz], rtol=1e-12)


def test_noalloc_binning_rejects_sizes_beyond_its_array(kn, torch_cuda):
    torch = torch_cuda
    buf = torch.zeros((4, 8, 5), dtype=torch.complex128, device="cuda")
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    assert kn.kn_bin_power_device_noalloc(buf.data_ptr(), 8, (1 << 18) + 1, 0, 4, out.data_ptr(),
                                          out.data_ptr() + 8, None) == 3004
    assert kn.kn_bin_power_device_noalloc(buf.data_ptr(), 8, 3, 2, 7, out.data_ptr(), out.data_ptr() + 8, None) == 3002
