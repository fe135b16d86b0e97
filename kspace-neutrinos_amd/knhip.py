"""ctypes mirror of include/kspace_neutrinos_hip.h (libkspace_neutrinos_hip.so).

The product is the C ABI; this module only binds it for Python callers
(tests, bench.py) and carries the thin torch.autograd.Function used for
gradient checks.  Loading fails loudly when the shared library is missing:
there is no Python or CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkspace_neutrinos_hip.so")

dp = C.POINTER(C.c_double)
llp = C.POINTER(C.c_longlong)
NUSPECIES = 3


class RhoNuSingle(C.Structure):
    _fields_ = [("loga", dp), ("rhonu", dp), ("interp", C.c_void_p), ("acc", C.c_void_p), ("mnu", C.c_double)]


class HybridNu(C.Structure):
    _fields_ = [("enabled", C.c_int), ("nufrac_low", C.c_double * NUSPECIES), ("nu_crit_time", C.c_double),
                ("vcrit", C.c_double)]


class OmegaNu(C.Structure):
    _fields_ = [("RhoNuTab", C.POINTER(RhoNuSingle) * NUSPECIES), ("nu_degeneracies", C.c_int * NUSPECIES),
                ("rhocrit", C.c_double), ("kBtnu", C.c_double), ("tcmb0", C.c_double), ("hybnu", HybridNu)]


class TransferInitTable(C.Structure):
    _fields_ = [("NPowerTable", C.c_int), ("logk", dp), ("T_nu", dp)]


class DeltaTotTable(C.Structure):
    _fields_ = [("nk", C.c_int), ("nk_allocated", C.c_int), ("namax", C.c_int), ("ia", C.c_int),
                ("ThisTask", C.c_int), ("delta_nu_prefac", C.c_double), ("delta_tot_init_done", C.c_int),
                ("debug", C.c_int), ("delta_tot", C.POINTER(dp)), ("scalefact", dp), ("delta_nu_init", dp),
                ("delta_nu_last", dp), ("wavenum", dp), ("omnu", C.POINTER(OmegaNu)), ("Omeganonu", C.c_double),
                ("light", C.c_double), ("TimeTransfer", C.c_double)]


class DeltaPow(C.Structure):
    _fields_ = [("logkk", dp), ("delta_ratio", dp), ("spline", C.c_void_p), ("acc", C.c_void_p),
                ("nbins", C.c_int), ("norm", C.c_double)]


class KspaceParams(C.Structure):
    _fields_ = [("KspaceTransferFunction", C.c_char * 500), ("TimeTransfer", C.c_double),
                ("InputSpectrum_UnitLength_in_cm", C.c_double), ("MNu", C.c_double * NUSPECIES),
                ("hybrid_neutrinos_on", C.c_int), ("vcrit", C.c_double), ("nu_crit_time", C.c_double)]


class BackgroundParams(C.Structure):
    """kn_background_params: the background whose Omega_nu(a) and H(a) the device tabulates."""
    _fields_ = [("Omega0", C.c_double), ("OmegaLambda", C.c_double), ("Hubble", C.c_double),
                ("HubbleParam", C.c_double), ("tcmb0", C.c_double), ("a0", C.c_double),
                ("MNu", C.c_double * NUSPECIES), ("radiation", C.c_int)]


class BatchHistory(C.Structure):
    """kn_batch_history: one get_delta_nu integral of a group (hubble: C function address)."""
    _fields_ = [("nk", C.c_int), ("na", C.c_int), ("na_max", C.c_int), ("a", C.c_double), ("a_max", C.c_double),
                ("time_transfer", C.c_double), ("mnu", C.c_double), ("kBtnu", C.c_double), ("light", C.c_double),
                ("delta_nu_prefac", C.c_double), ("qc", C.c_double), ("relerr", C.c_double),
                ("nufrac_low", C.c_double), ("vcrit", C.c_double), ("hybrid", C.c_int), ("nu_crit_time", C.c_double),
                ("ic_only", C.c_int), ("hubble", C.c_void_p), ("hubble_ctx", C.c_void_p)]


vpp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); the reference entry points first, then kn_ additions
SIGNATURES = {
    "total_powerspectrum": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, dp, llp, dp, C.c_int]),
    "allocate_delta_tot_table": (None, [C.POINTER(DeltaTotTable), C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(OmegaNu), C.c_double, C.c_double, C.c_int]),
    "free_delta_tot_table": (None, [C.POINTER(DeltaTotTable)]),
    "delta_tot_init": (None, [C.POINTER(DeltaTotTable), C.c_int, dp, dp, C.POINTER(TransferInitTable), C.c_double]),
    "update_delta_tot": (None, [C.POINTER(DeltaTotTable), C.c_double, dp, dp, C.c_int]),
    "get_delta_nu_update": (None, [C.POINTER(DeltaTotTable), C.c_double, C.c_int, dp, dp, dp,
                                   C.POINTER(TransferInitTable)]),
    "get_delta_nu": (None, [C.POINTER(DeltaTotTable), C.c_double, dp, dp, C.c_double]),
    "get_delta_nu_combined": (None, [C.POINTER(DeltaTotTable), C.c_double, dp, dp]),
    "specialJ": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "fslength": (C.c_double, [C.c_double, C.c_double, C.c_double]),
    "get_delta_tot": (C.c_double, [C.c_double] * 6),
    "init_omega_nu": (None, [C.POINTER(OmegaNu), dp, C.c_double, C.c_double, C.c_double]),
    "get_omega_nu": (C.c_double, [C.POINTER(OmegaNu), C.c_double]),
    "get_omega_nu_nopart": (C.c_double, [C.POINTER(OmegaNu), C.c_double]),
    "get_omegag": (C.c_double, [C.POINTER(OmegaNu), C.c_double]),
    "omega_nu_single": (C.c_double, [C.POINTER(OmegaNu), C.c_double, C.c_int]),
    "rho_nu_init": (None, [C.POINTER(RhoNuSingle), C.c_double, C.c_double, C.c_double, C.c_double]),
    "rho_nu": (C.c_double, [C.POINTER(RhoNuSingle), C.c_double, C.c_double]),
    "init_hybrid_nu": (None, [C.POINTER(HybridNu), dp, C.c_double, C.c_double, C.c_double, C.c_double]),
    "particle_nu_fraction": (C.c_double, [C.POINTER(HybridNu), C.c_double, C.c_int]),
    "nufrac_low": (C.c_double, [C.c_double]),
    "allocate_transfer_init_table": (None, [C.POINTER(TransferInitTable), C.c_double, C.c_double, C.c_double,
                                            C.c_char_p]),
    "free_transfer_init_table": (None, [C.POINTER(TransferInitTable)]),
    "init_delta_pow": (None, [C.POINTER(DeltaPow), dp, dp, C.c_int, C.c_double]),
    "get_dnudcdm_powerspec": (C.c_double, [C.POINTER(DeltaPow), C.c_double]),
    "free_d_pow": (None, [C.POINTER(DeltaPow)]),
    "OmegaNu": (C.c_double, [C.c_double]),
    "OmegaNu_nopart": (C.c_double, [C.c_double]),
    "InitOmegaNu": (None, [C.c_double, C.c_double, C.c_int]),
    "allocate_kspace_memory": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_char_p,
                                      C.c_double, C.c_int]),
    "compute_neutrino_power_from_cdm": (DeltaPow, [C.c_double, dp, dp, C.POINTER(C.c_long), C.c_int, C.c_int]),
    "get_nu_state": (None, [C.POINTER(dp), C.POINTER(dp), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "set_nu_state": (None, [dp, dp, C.c_size_t, C.c_size_t, C.c_int]),
    "save_delta_tot": (None, [C.POINTER(DeltaTotTable), C.c_int, C.c_char_p]),
    "save_all_nu_state": (None, [C.POINTER(DeltaTotTable), C.c_char_p]),
    "save_nu_power": (C.c_int, [C.POINTER(DeltaTotTable), C.c_double, C.c_int, C.c_char_p]),
    "read_all_nu_state": (None, [C.POINTER(DeltaTotTable), C.c_char_p]),
    "save_nu_state": (None, [C.c_char_p]),
    "save_neutrino_power": (C.c_int, [C.c_double, C.c_int, C.c_char_p]),
    "particle_nu_active": (C.c_int, [C.c_double]),
    "add_nu_power_to_rhogrid": (None, [C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "kn_set_hubble_function": (None, [C.c_void_p]),
    "kn_synth_grid": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_ulonglong]),
    "kn_bin_power_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "kn_rescale_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                    C.c_int, C.c_double, C.c_void_p]),
    "kn_bin_power_device_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                              C.c_void_p, C.c_void_p]),
    "kn_dnu_last_upload_columns": (C.c_int, []),
    "kn_dnu_last_max_intervals": (C.c_int, []),
    "kn_dnu_last_intervals": (C.c_longlong, [C.POINTER(C.c_int)]),
    "kn_dnu_qag_timing": (C.c_int, [C.c_int]),
    "kn_dnu_last_qag_ms": (C.c_double, []),
    "kn_rescale_green_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                          C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p]),
    "kn_last_error": (C.c_char_p, []),
    "kn_bin_power_adjoint": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_void_p,
                                       C.c_void_p]),
    "kn_batch_history_from_table": (C.c_int, [C.POINTER(DeltaTotTable), C.c_double, C.c_double,
                                              C.POINTER(BatchHistory)]),
    "kn_hubble_ctx": (C.c_double, [C.c_double, C.c_void_p]),
    "kn_group_export": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(DeltaTotTable)]),
    "kn_group_import": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(DeltaTotTable), C.c_double]),
    "kn_background_create": (C.c_int, [vpp, C.POINTER(BackgroundParams)]),
    "kn_background_destroy": (None, [C.c_void_p]),
    "kn_background_omega_nu": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kn_background_hubble": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kn_background_hubble_ctx": (C.c_double, [C.c_double, C.c_void_p]),
    "kn_background_rho_table": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
    "kn_set_background": (C.c_int, [C.c_void_p]),
    "kn_group_create": (C.c_int, [vpp, C.c_int, C.POINTER(BatchHistory)]),
    "kn_group_destroy": (None, [C.c_void_p]),
    "kn_group_views": (C.c_int, [C.c_void_p] + [vpp] * 9),
    "kn_group_layout": (C.c_int, [C.c_void_p, llp, llp, llp]),
    "kn_group_solve": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kn_group_adjoint": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kn_group_append": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kn_group_append_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kn_group_set_capacity_plan": (C.c_int, [C.c_void_p, C.c_int]),
    "kn_group_graph_capture": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "kn_group_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kn_group_status": (C.c_int, [C.c_void_p]),
    "kn_group_subdivision": (C.c_int, [C.c_void_p, C.c_int, dp, C.POINTER(C.c_int)]),
    "kn_group_autotune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kn_group_set_mass": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "kn_group_streams": (C.c_int, [C.c_void_p]),
    "kn_group_set_streams": (C.c_int, [C.c_void_p, C.c_int]),
    "kn_get_delta_nu_adjoint": (C.c_int, [C.POINTER(DeltaTotTable), C.c_double, dp, C.c_double, dp, dp, dp, dp,
                                          dp]),
}

_lib = None


def lib():
    """Load libkspace_neutrinos_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libkspace_neutrinos_hip.so missing: run __graft_entry__.build() "
                               "(make -C kspace-neutrinos_amd)")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names this binding expects the library to export."""
    return sorted(SIGNATURES)


# ---------------------------------------------------------------- batches

H2D, D2H, D2D = 1, 2, 3
_hip = None


def _memcpy(dst, src, nbytes, kind):
    """hipMemcpy through the HIP runtime the library itself links."""
    global _hip
    if _hip is None:
        lib()
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    if nbytes:
        rc = _hip.hipMemcpy(dst, src, nbytes, kind)
        if rc:
            raise RuntimeError("hipMemcpy failed: %d" % rc)


def _check(rc, what, L=None):
    if rc:
        raise RuntimeError("%s failed with code %d: %s" % (what, rc, (L or lib()).kn_last_error().decode()))


def bind(L):
    """Set this module's signatures on another build of the same ABI (returns L)."""
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name, None)
        if f is not None:
            f.restype = res
            f.argtypes = args
    return L


class Background:
    """kn_background: Omega_nu(a) and H(a) tabulated on the device from the
    background's parameters (omega_nu_single.c:16-205; gadget-2 patch
    0003:38-55).  hubble_fn / ctx go into a BatchHistory so that its H panels
    are built on the device."""

    def __init__(self, Omega0, HubbleParam, Hubble, masses, a0=0.01, tcmb0=2.7255, OmegaLambda=None,
                 radiation=1, L=None):
        self.L = L or lib()
        p = BackgroundParams()
        p.Omega0 = Omega0
        p.OmegaLambda = 1.0 - Omega0 if OmegaLambda is None else OmegaLambda
        p.Hubble, p.HubbleParam, p.tcmb0, p.a0, p.radiation = Hubble, HubbleParam, tcmb0, a0, radiation
        for i in range(NUSPECIES):
            p.MNu[i] = masses[i]
        self.params = p
        h = C.c_void_p()
        _check(self.L.kn_background_create(C.byref(h), C.byref(p)), "kn_background_create", self.L)
        self.ctx = h.value
        self.hubble_fn = C.cast(self.L.kn_background_hubble_ctx, C.c_void_p).value

    def hubble_host(self, a):
        return self.L.kn_background_hubble_ctx(a, self.ctx)

    def _eval(self, fn, a):
        """a: float64 tensor on the device; returns the values in a new tensor."""
        import torch
        a = a.contiguous()
        out = torch.empty_like(a)
        _check(fn(self.ctx, a.numel(), a.data_ptr(), out.data_ptr(), None), "kn_background eval", self.L)
        torch.cuda.synchronize()
        return out

    def omega_nu(self, a):
        return self._eval(self.L.kn_background_omega_nu, a)

    def hubble(self, a):
        return self._eval(self.L.kn_background_hubble, a)

    def rho_table(self, i):
        """(loga, rhonu) of species i, or None when it has no table."""
        import numpy as np
        loga, rho = np.zeros(200), np.zeros(200)
        n = self.L.kn_background_rho_table(self.ctx, i, loga.ctypes.data_as(dp), rho.ctypes.data_as(dp))
        if n < 0:
            raise RuntimeError("kn_background_rho_table: bad arguments")
        return (loga, rho) if n else None

    def set_process(self, on=True):
        """Make this background the process's H (kn_set_background)."""
        self.L.kn_set_background(self.ctx if on else None)

    def close(self):
        if self.ctx:
            self.L.kn_background_destroy(self.ctx)
            self.ctx = None


class Batch:
    """Owner of a kn_group: loads histories into its device arrays and reads
    results back.

    hists: list of BatchHistory (their hubble_ctx objects must outlive the batch).
    L, memcpy: the library build and its device copy (defaults: this library, hipMemcpy).
    """

    VIEWS = ("scalefact", "delta_tot", "wavenum", "delta_nu_init", "delta_nu", "gbar", "dtot_bar", "init_bar",
             "mnu_bar")

    def __init__(self, hists, keep=(), L=None, memcpy=None, api="kn_group"):
        L = L or lib()
        self.L = L
        self.api = api
        self._mc = memcpy or _memcpy
        self.h = None
        self.n = len(hists)
        self.hists = list(hists)
        self._keep = keep
        arr = (BatchHistory * self.n)(*hists)
        h = C.c_void_p()
        _check(self._f("create")(C.byref(h), self.n, arr), self.api + "_create", L)
        self.h = h
        ptrs = [C.c_void_p() for _ in self.VIEWS]
        _check(self._f("views")(h, *[C.byref(p) for p in ptrs]), self.api + "_views", L)
        self.ptr = {name: p.value for name, p in zip(self.VIEWS, ptrs)}
        self.off_k = (C.c_longlong * (self.n + 1))()
        self.off_y = (C.c_longlong * (self.n + 1))()
        self.off_s = (C.c_longlong * (self.n + 1))()
        _check(self._f("layout")(h, self.off_k, self.off_y, self.off_s), self.api + "_layout", L)

    def _f(self, name):
        return getattr(self.L, self.api + "_" + name)

    def _put(self, name, off, arr):
        import numpy as np
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._mc(self.ptr[name] + 8 * off, a.ctypes.data, a.nbytes, H2D)

    def _get(self, name, off, n):
        import numpy as np
        a = np.empty(n, dtype=np.float64)
        self._mc(a.ctypes.data, self.ptr[name] + 8 * off, a.nbytes, D2H)
        return a

    def load(self, i, logs, rows, wavenum, init):
        """History i: stored log a [na], delta_tot rows [nk][na], wavenumbers [nk], delta_nu_init [nk]."""
        import numpy as np
        hs = self.hists[i]
        rows = np.asarray(rows, dtype=np.float64)
        blk = np.zeros((hs.nk, hs.na_max))
        blk[:, :rows.shape[1]] = rows
        self._put("scalefact", self.off_s[i], logs)
        self._put("delta_tot", self.off_y[i], blk)
        self._put("wavenum", self.off_k[i], wavenum)
        self._put("delta_nu_init", self.off_k[i], init)

    def solve(self):
        _check(self._f("solve")(self.h, None), self.api + "_solve", self.L)

    def adjoint(self, gbar_all=None):
        if gbar_all is not None:
            self._put("gbar", 0, gbar_all)
        _check(self._f("adjoint")(self.h, None), self.api + "_adjoint", self.L)

    def append(self, loga, rows):
        """One more stored step per history (host arrays: log a per history, rows per item)."""
        import numpy as np
        la = np.ascontiguousarray(loga, dtype=np.float64)
        rw = np.ascontiguousarray(rows, dtype=np.float64)
        # staged in the batch's gbar (rows) and mnu_bar (log a) arrays, which this overwrites
        self._put("gbar", 0, rw)
        self._put("mnu_bar", 0, la)
        _check(self._f("append")(self.h, C.c_void_p(self.ptr["mnu_bar"]), C.c_void_p(self.ptr["gbar"]), None),
               self.api + "_append", self.L)

    def append_async(self, loga, rows):
        """The same step with no host synchronisation (capacity plan): the
        step is staged in the batch's own device buffers; a rejected history
        shows in the next status()."""
        import numpy as np
        self._stage_step(np.ascontiguousarray(loga, dtype=np.float64), np.ascontiguousarray(rows, dtype=np.float64))
        _check(self._f("append_async")(self.h, C.c_void_p(self.step_ptr[0]), C.c_void_p(self.step_ptr[1]), None),
               self.api + "_append_async", self.L)

    def _stage_step(self, la, rw):
        """Copies a step into device buffers the batch owns (torch tensors,
        allocated on first use; host tensors where there is no GPU, for the
        CPU emulation build)."""
        if not hasattr(self, "step_buf"):
            import torch
            dev = "cuda" if torch.cuda.is_available() else "cpu"
            self.step_buf = (torch.zeros(self.n, dtype=torch.float64, device=dev),
                             torch.zeros(int(self.off_k[self.n]), dtype=torch.float64, device=dev))
            self.step_ptr = (self.step_buf[0].data_ptr(), self.step_buf[1].data_ptr())
        self._mc(self.step_ptr[0], la.ctypes.data, la.nbytes, H2D)
        self._mc(self.step_ptr[1], rw.ctypes.data, rw.nbytes, H2D)

    def set_capacity_plan(self, on=True):
        _check(self._f("set_capacity_plan")(self.h, 1 if on else 0), self.api + "_set_capacity_plan", self.L)

    def capture(self, with_adjoint=True, with_append=False, stream=None):
        """Capture [append +] solve [+ adjoint]; with_append stages each
        replay's step with stage_step() before graph_launch()."""
        if with_append and not hasattr(self, "step_buf"):
            import numpy as np
            self._stage_step(np.zeros(self.n), np.zeros(int(self.off_k[self.n])))
        la = C.c_void_p(self.step_ptr[0]) if with_append else None
        rw = C.c_void_p(self.step_ptr[1]) if with_append else None
        _check(self._f("graph_capture")(self.h, 1 if with_adjoint else 0, la, rw, stream), self.api + "_graph_capture",
               self.L)

    def stage_step(self, loga, rows):
        import numpy as np
        self._stage_step(np.ascontiguousarray(loga, dtype=np.float64), np.ascontiguousarray(rows, dtype=np.float64))

    def graph_launch(self, stream=None):
        _check(self._f("graph_launch")(self.h, stream), self.api + "_graph_launch", self.L)

    def status(self):
        return self._f("status")(self.h)

    def subdivision(self, i):
        """History i's final QAG subdivisions: (ivl[nk, 200, 2], nint[nk])."""
        import numpy as np
        nk = self.hists[i].nk
        ivl = np.zeros((nk, 200, 2))
        nint = np.zeros(nk, dtype=np.int32)
        _check(self._f("subdivision")(self.h, i, ivl.ctypes.data_as(dp), nint.ctypes.data_as(C.POINTER(C.c_int))),
               self.api + "_subdivision", self.L)
        return ivl, nint

    def delta_nu(self, i=None):
        if i is None:
            return self._get("delta_nu", 0, self.off_k[self.n])
        return self._get("delta_nu", self.off_k[i], self.hists[i].nk)

    def dtot_bar(self, i, na):
        hs = self.hists[i]
        return self._get("dtot_bar", self.off_y[i], hs.nk * hs.na_max).reshape(hs.nk, hs.na_max)[:, :na]

    def init_bar(self, i):
        return self._get("init_bar", self.off_k[i], self.hists[i].nk)

    def mnu_bar(self):
        return self._get("mnu_bar", 0, self.n)

    def set_mass(self, i, mnu):
        _check(self._f("set_mass")(self.h, i, mnu), self.api + "_set_mass", self.L)

    def set_streams(self, n):
        """0 merged launches, n > 0 per-history sequences on n streams."""
        _check(self._f("set_streams")(self.h, n), self.api + "_set_streams", self.L)

    def streams(self):
        return self._f("streams")(self.h)

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def power_spectrum_function():
    """torch.autograd.Function over total_powerspectrum on a device grid:
    apply(field[dims, dims, dims/2+1, 2] float64 cuda, nrbins) -> power[n]
    (the compacted, normalised bins); the gradient comes from
    kn_bin_power_adjoint."""
    import numpy as np
    import torch

    class PowerSpectrumFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, field, nrbins):
            assert field.is_cuda and field.dtype == torch.float64 and field.is_contiguous()
            dims = field.shape[0]
            power, keffs = np.zeros(nrbins), np.zeros(nrbins)
            count = np.zeros(nrbins, dtype=np.int64)
            n = lib().total_powerspectrum(dims, field.data_ptr(), nrbins, 0, dims, power.ctypes.data_as(dp),
                                          count.ctypes.data_as(llp), keffs.ctypes.data_as(dp), 0)
            ctx.save_for_backward(field)
            ctx.nrbins, ctx.n = nrbins, n
            return torch.from_numpy(power[:n].copy()).to(field.device)

        @staticmethod
        def backward(ctx, g):
            (field,) = ctx.saved_tensors
            gb = np.ascontiguousarray(g.detach().cpu().numpy(), dtype=np.float64)
            fb = torch.empty_like(field)
            _check(lib().kn_bin_power_adjoint(field.data_ptr(), field.shape[0], ctx.nrbins, 0, field.shape[0],
                                              gb.ctypes.data_as(dp), ctx.n, fb.data_ptr(), None),
                   "kn_bin_power_adjoint")
            return fb, None

    return PowerSpectrumFunction


def delta_nu_function():
    """torch.autograd.Function over get_delta_nu of one table:
    apply(delta_tot[nk, na], delta_nu_init[nk], mnu[()], state) -> delta_nu[nk],
    where state = DeltaNuState(table, a, wavenum).  Gradients come from the
    C-ABI adjoint (kn_group_adjoint) with the forward's QAG subdivision."""
    import torch

    class DeltaNuFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, delta_tot, delta_nu_init, mnu, state):
            out = state.run(delta_tot, delta_nu_init, float(mnu))
            ctx.state = state
            ctx.save_for_backward(delta_tot, delta_nu_init, mnu)
            return out

        @staticmethod
        def backward(ctx, g):
            delta_tot, delta_nu_init, mnu = ctx.saved_tensors
            db, ib, mb = ctx.state.grad(delta_tot, delta_nu_init, float(mnu), g)
            return db, ib, torch.tensor(mb, dtype=mnu.dtype, device=mnu.device), None

    return DeltaNuFunction


def table_views(table):
    """numpy views of a _delta_tot_table's history block: (scalefact[namax], delta_tot[nk_allocated][namax],
    delta_nu_init[nk])."""
    import numpy as np
    namax, nka = table.namax, table.nk_allocated
    blk = np.ctypeslib.as_array(table.scalefact, shape=(namax * (nka + 1),))
    return blk[:namax], blk[namax:].reshape(nka, namax), np.ctypeslib.as_array(table.delta_nu_init,
                                                                               shape=(table.nk,))


def delta_nu_plain_function():
    """torch.autograd.Function over get_delta_nu on a table (per-call path):
    apply(delta_tot[nk, ia], delta_nu_init[nk], mnu[()], table, a, wavenum) ->
    delta_nu[nk], CPU float64 tensors as the reference's caller holds them.
    Forward: get_delta_nu; backward: kn_get_delta_nu_adjoint with the same
    inputs written back into the table."""
    import numpy as np
    import torch

    def load(table, dt, di):
        _, rows, init = table_views(table)
        rows[:table.nk, :table.ia] = dt.detach().numpy()
        init[:] = di.detach().numpy()

    class DeltaNuPlainFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, delta_tot, delta_nu_init, mnu, table, a, wavenum):
            load(table, delta_tot, delta_nu_init)
            kk = np.ascontiguousarray(wavenum, dtype=np.float64)
            out = np.zeros(table.nk)
            lib().get_delta_nu(C.byref(table), a, kk.ctypes.data_as(dp), out.ctypes.data_as(dp), float(mnu))
            ctx.save_for_backward(delta_tot, delta_nu_init, mnu)
            ctx.table, ctx.a, ctx.k = table, a, kk
            return torch.from_numpy(out)

        @staticmethod
        def backward(ctx, g):
            delta_tot, delta_nu_init, mnu = ctx.saved_tensors
            t = ctx.table
            load(t, delta_tot, delta_nu_init)
            gb = np.ascontiguousarray(g.detach().numpy(), dtype=np.float64)
            db, ib, mb = np.zeros((t.nk, t.ia)), np.zeros(t.nk), C.c_double()
            _check(lib().kn_get_delta_nu_adjoint(C.byref(t), ctx.a, ctx.k.ctypes.data_as(dp), float(mnu),
                                                 gb.ctypes.data_as(dp), None, db.ctypes.data_as(dp),
                                                 ib.ctypes.data_as(dp), C.byref(mb)), "kn_get_delta_nu_adjoint")
            return (torch.from_numpy(db), torch.from_numpy(ib), torch.tensor(mb.value, dtype=mnu.dtype),
                    None, None, None)

    return DeltaNuPlainFunction


class DeltaNuState:
    """A one-history group for get_delta_nu(table, a, wavenum, .) whose inputs
    come from torch tensors (device memory, float64)."""

    def __init__(self, table, a, wavenum, mnu, L=None, memcpy=None):
        import numpy as np
        L = L or lib()
        self._mc = memcpy or _memcpy
        hh = BatchHistory()
        _check(L.kn_batch_history_from_table(C.byref(table), a, mnu, C.byref(hh)), "history_from_table", L)
        self.nk, self.na = hh.nk, hh.na
        self.batch = Batch([hh], L=L, memcpy=memcpy)
        logs = np.ctypeslib.as_array(table.scalefact, shape=(self.na,)).copy()
        self.batch._put("scalefact", 0, logs)
        self.batch._put("wavenum", 0, np.asarray(wavenum, dtype=np.float64))

    def _load(self, delta_tot, delta_nu_init, mnu):
        assert delta_tot.is_contiguous() and tuple(delta_tot.shape) == (self.nk, self.na)
        assert delta_nu_init.is_contiguous() and tuple(delta_nu_init.shape) == (self.nk,)
        b = self.batch
        b.set_mass(0, mnu)
        self._mc(b.ptr["delta_tot"], delta_tot.data_ptr(), 8 * self.nk * self.na, D2D)
        self._mc(b.ptr["delta_nu_init"], delta_nu_init.data_ptr(), 8 * self.nk, D2D)
        b.solve()

    def run(self, delta_tot, delta_nu_init, mnu):
        import torch
        self._load(delta_tot, delta_nu_init, mnu)
        out = torch.empty(self.nk, dtype=torch.float64, device=delta_tot.device)
        self._mc(out.data_ptr(), self.batch.ptr["delta_nu"], 8 * self.nk, D2D)
        return out

    def grad(self, delta_tot, delta_nu_init, mnu, g):
        import torch
        self._load(delta_tot, delta_nu_init, mnu)
        gc = g.detach().to(torch.float64).contiguous()
        b = self.batch
        self._mc(b.ptr["gbar"], gc.data_ptr(), 8 * self.nk, D2D)
        b.adjoint()
        db = torch.empty((self.nk, self.na), dtype=torch.float64, device=delta_tot.device)
        ib = torch.empty(self.nk, dtype=torch.float64, device=delta_tot.device)
        self._mc(db.data_ptr(), b.ptr["dtot_bar"], 8 * self.nk * self.na, D2D)
        self._mc(ib.data_ptr(), b.ptr["init_bar"], 8 * self.nk, D2D)
        mb = b.mnu_bar()[0]
        return db, ib, mb
