"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, keeps the reference struct layouts, and its host-only
functions (no kernel) agree with the oracle."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

import knhip
import oracle_py as O
from conftest import REFDATA, ROOT, UNIT_LENGTH, UNIT_TIME

HEADER = os.path.join(ROOT, "include", "kspace_neutrinos_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", src, re.M):
        name = m.group(1)
        if name not in ("if", "while", "sizeof", "struct", "typedef"):
            names.add(name)
    host_imports = {"hubble_function", "terminate", "message", "mymalloc_fullinfo", "myfree_fullinfo"}
    return sorted(names - host_imports)


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref in ["total_powerspectrum", "get_delta_nu", "get_delta_nu_update", "delta_tot_init", "update_delta_tot",
                "allocate_delta_tot_table", "get_delta_nu_combined", "add_nu_power_to_rhogrid", "init_delta_pow",
                "get_dnudcdm_powerspec", "specialJ", "fslength", "get_delta_tot", "compute_neutrino_power_from_cdm"]:
        assert ref in names


def test_library_exports_every_declared_symbol(kn):
    out = subprocess.run(["nm", "-D", "--defined-only", knhip.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert "kspace_params" in exported
    for name in knhip.exported_symbols():
        assert hasattr(kn, name)


def test_host_services_are_weak_imports():
    out = subprocess.run(["nm", "-D", knhip.LIB_PATH], capture_output=True, text=True, check=True).stdout
    kinds = {line.split()[-1]: line.split()[-2] for line in out.splitlines() if len(line.split()) >= 2}
    for name in ["hubble_function", "terminate", "message", "mymalloc_fullinfo", "myfree_fullinfo"]:
        assert kinds.get(name) == "w", name


def test_struct_layouts_match_reference():
    # field order and types of delta_tot_table.h:17-58 on LP64
    assert C.sizeof(knhip.DeltaTotTable) == 4 * 5 + 4 + 8 + 4 + 4 + 8 * 6 + 8 * 3
    assert knhip.DeltaTotTable.delta_tot.offset == 40
    assert C.sizeof(knhip.HybridNu) == 48
    assert C.sizeof(knhip.OmegaNu) == 8 * 3 + 4 * 3 + 4 + 8 * 3 + 48
    assert C.sizeof(knhip.DeltaPow) == 8 * 4 + 8 + 8
    assert C.sizeof(knhip.TransferInitTable) == 24


@pytest.fixture(scope="module")
def pair(kn, host, orc):
    masses = [0.15, 0.15, 0.15]
    om = knhip.OmegaNu()
    kn.init_omega_nu(C.byref(om), (C.c_double * 3)(*masses), 0.01, 0.7, 2.7255)
    host.kn_test_init_hubble(C.byref(om), 0.2793, UNIT_TIME)
    oom = O.omega_nu(masses)
    orc.orc_init_hubble(C.byref(oom), 0.2793, UNIT_TIME)
    return om, oom


def test_background_matches_oracle_exactly(kn, orc, pair):
    om, oom = pair
    assert list(om.nu_degeneracies) == list(oom.deg)
    for a in [0.005, 0.01, 0.07, 0.3, 0.999, 1.0, 3.0]:
        assert kn.get_omega_nu(C.byref(om), a) == orc.orc_get_omega_nu(C.byref(oom), a)
        assert kn.get_omegag(C.byref(om), a) == orc.orc_get_omegag(C.byref(oom), a)
        assert kn.omega_nu_single(C.byref(om), a, 1) == orc.orc_omega_nu_single(C.byref(oom), a, 1)
    for qc in [0.0, 0.3, 1.0, 4.0]:
        assert kn.nufrac_low(qc) == orc.orc_nufrac_low(qc)


def test_special_functions_match_oracle_exactly(kn, orc, pair):
    for x in [0.0, 1e-3, 0.3, 1.0, 2.5, 17.0]:
        assert kn.specialJ(x, -1, 0) == orc.orc_specialJ(x, -1, 0)
        for qc, f in [(0.05, 0.001), (1.0, 0.06), (3.0, 0.4)]:
            assert kn.specialJ(x, qc, f) == orc.orc_specialJ(x, qc, f)
    for xi, xf in [(math.log(0.01), math.log(0.5)), (math.log(0.3), 0.0), (0.0, -1.0)]:
        assert kn.fslength(xi, xf, 299792.) == orc.orc_fslength(xi, xf, 299792.)
    assert kn.get_delta_tot(0.1, 2.0, 3e-3, 0.27, 6e-3, 0.02) == orc.orc_get_delta_tot(0.1, 2.0, 3e-3, 0.27, 6e-3, 0.02)


def test_transfer_table_matches_oracle(kn):
    f = os.path.join(REFDATA, "ics_transfer_99.dat")
    t = knhip.TransferInitTable()
    kn.allocate_transfer_init_table(C.byref(t), 512000, UNIT_LENGTH, UNIT_LENGTH * 1e3, f.encode())
    rc, ot = O.load_transfer(f, 512000, UNIT_LENGTH, UNIT_LENGTH * 1e3)
    assert t.NPowerTable == ot.NPowerTable == 271
    assert all(t.logk[i] == ot.logk[i] and t.T_nu[i] == ot.T_nu[i] for i in range(271))
    kn.free_transfer_init_table(C.byref(t))


def test_delta_pow_host_interpolation(kn):
    logkk = np.linspace(-9.0, -4.0, 50)
    ratio = np.cos(logkk)
    d = knhip.DeltaPow()
    kn.init_delta_pow(C.byref(d), logkk.ctypes.data_as(knhip.dp), ratio.ctypes.data_as(knhip.dp), 50, 0.7)
    for x in [-9.5, -9.0, -8.123, -6.0, -4.0, -3.0]:
        assert kn.get_dnudcdm_powerspec(C.byref(d), x) == O.get_dnudcdm(logkk, ratio, 0.7, x)


def _run_child(code):
    env = dict(os.environ)
    return subprocess.run(["python3", "-c", code], capture_output=True, text=True, env=env, cwd=ROOT, timeout=120)


def test_errors_reach_host_terminate_with_reference_codes():
    prelude = ("import sys, ctypes as C; sys.path[:0]=['kspace-neutrinos_amd','tests'];"
               "import conftest; h=conftest.load_testhost(); import knhip; L=knhip.lib();")
    # transfer file missing -> terminate(2019) (transfer_init.c:219-221)
    r = _run_child(prelude + "t=knhip.TransferInitTable(); L.allocate_transfer_init_table(C.byref(t),1.,1.,1.,b'/nonexistent')")
    assert "terminate(2019)" in r.stderr and r.returncode == 2019 % 256, r.stderr
    # nk mismatch -> 2002 (delta_tot_table.c:203-204)
    r = _run_child(prelude + (
        "import numpy as np; om=knhip.OmegaNu(); L.init_omega_nu(C.byref(om),(C.c_double*3)(0.1,0.1,0.1),0.01,0.7,2.7255);"
        "h.kn_test_init_hubble(C.byref(om),0.3,3.085678e16); d=knhip.DeltaTotTable();"
        "L.allocate_delta_tot_table(C.byref(d),8,0.01,1.0,0.3,C.byref(om),3.085678e16,3.085678e21,0);"
        "d.delta_tot_init_done=1; d.ia=1; k=np.ones(4); o=np.zeros(4);"
        "L.get_delta_nu_update(C.byref(d),0.5,4,k.ctypes.data_as(knhip.dp),k.ctypes.data_as(knhip.dp),o.ctypes.data_as(knhip.dp),None)"))
    assert "terminate(2002)" in r.stderr and r.returncode == 2002 % 256, r.stderr
