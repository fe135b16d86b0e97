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
    "particle_nu_active": (C.c_int, [C.c_double]),
    "add_nu_power_to_rhogrid": (None, [C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "kn_set_hubble_function": (None, [C.c_void_p]),
    "kn_synth_grid": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_ulonglong]),
    "kn_bin_power_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "kn_rescale_device": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                    C.c_int, C.c_double, C.c_void_p]),
    "kn_last_error": (C.c_char_p, []),
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
