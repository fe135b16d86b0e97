"""State files of the integrator (delta_tot_table.c:254-374, interface_common.c:41-52,
:93-95): the history table written and read in the reference's text format,
the .bak kept on save, the nk taken from a short first line, the error codes,
and the P_nu(k) output.  Host code only (no GPU).

The reference's own state file (testdata/delta_tot_nu.txt, 25 stored steps x
300 k, written by the reference's save_all_nu_state) is read and written back
here: the rewrite is byte-identical, so files pass both ways between the two
implementations.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import knhip
from conftest import REFDATA, ROOT, UNIT_LENGTH, UNIT_TIME

STATE = os.path.join(REFDATA, "delta_tot_nu.txt")


def _table(kn, nk):
    om = knhip.OmegaNu()
    kn.init_omega_nu(C.byref(om), (C.c_double * 3)(0.15, 0.15, 0.15), 0.01, 0.7, 2.7255)
    d = knhip.DeltaTotTable()
    kn.allocate_delta_tot_table(C.byref(d), nk, 0.01, 1.0, 0.2793, C.byref(om), UNIT_TIME, UNIT_LENGTH, 0)
    return om, d


def _rows(path):
    with open(path) as fh:
        return [[float(v) for v in line[1:].split()] for line in fh if line.startswith("#")]


def test_read_reference_state_file(kn):
    """delta_tot_table_test.c:95-102: 25 steps, starting at log 0.01, increasing, positive."""
    om, d = _table(kn, 300)
    kn.read_all_nu_state(C.byref(d), STATE.encode())
    rows = _rows(STATE)
    assert d.ia == 25 == len(rows) and d.nk == 300
    sf, dt, _ = knhip.table_views(d)
    assert abs(sf[0] / math.log(0.01) - 1) < 1e-5
    for i, r in enumerate(rows):
        assert sf[i] == r[0]
        assert np.array_equal(dt[:300, i], np.array(r[1:]))
    assert np.all(np.diff(sf[:25]) > 0) and np.all(dt[:300, 1:25] > 0)
    kn.free_delta_tot_table(C.byref(d))


def test_save_is_byte_identical_and_reads_back(kn, tmp_path):
    """delta_tot_table_test.c:112-123: save, read into a fresh table, equal;
    the text equals the reference-written file line for line."""
    om, d = _table(kn, 300)
    kn.read_all_nu_state(C.byref(d), STATE.encode())
    out = str(tmp_path / "state.txt")
    kn.save_all_nu_state(C.byref(d), out.encode())
    assert open(out).read() == open(STATE).read()
    om2, d2 = _table(kn, 300)
    kn.read_all_nu_state(C.byref(d2), out.encode())
    assert d2.ia == d.ia and d2.nk == d.nk
    a, b = knhip.table_views(d), knhip.table_views(d2)
    assert np.array_equal(a[0][:25], b[0][:25]) and np.array_equal(a[1][:300, :25], b[1][:300, :25])
    # a second save keeps the first file as .bak
    d.ia = 3
    kn.save_all_nu_state(C.byref(d), out.encode())
    assert open(out + ".bak").read() == open(STATE).read()
    assert len(_rows(out)) == 3
    for t in (d, d2):
        kn.free_delta_tot_table(C.byref(t))


def test_missing_file_and_short_first_line(kn, tmp_path):
    """No file: nothing changes.  A first line shorter than nk sets nk."""
    om, d = _table(kn, 8)
    kn.read_all_nu_state(C.byref(d), str(tmp_path / "none.txt").encode())
    assert d.ia == 0 and d.nk == 8
    p = tmp_path / "short.txt"
    p.write_text("# %le 1.0 2.0 3.0 4.0 5.0 \n# %le 6.0 7.0 8.0 9.0 10.0 \n" % (math.log(0.01), math.log(0.02)))
    kn.read_all_nu_state(C.byref(d), str(p).encode())
    sf, dt, _ = knhip.table_views(d)
    assert d.nk == 5 and d.ia == 2
    assert list(dt[:5, 1]) == [6.0, 7.0, 8.0, 9.0, 10.0] and sf[1] == float("%e" % math.log(0.02))
    kn.free_delta_tot_table(C.byref(d))


def _child(code):
    pre = ("import sys, math, ctypes as C; sys.path[:0]=['kspace-neutrinos_amd','tests'];"
           "import conftest; conftest.load_testhost(); import knhip; L=knhip.lib();"
           "om=knhip.OmegaNu(); L.init_omega_nu(C.byref(om),(C.c_double*3)(0.1,0.1,0.1),0.01,0.7,2.7255);"
           "d=knhip.DeltaTotTable();"
           "L.allocate_delta_tot_table(C.byref(d),4,0.01,1.0,0.3,C.byref(om),3.085678e16,3.085678e21,0);")
    return subprocess.run(["python3", "-c", pre + code], capture_output=True, text=True, cwd=ROOT, timeout=120)


def test_state_file_errors_reach_terminate(tmp_path):
    """A short line after the first -> terminate(2006); a history that does not
    start at TimeTransfer -> terminate(2007) (delta_tot_table.c:280-291)."""
    short = tmp_path / "short2.txt"
    short.write_text("# %le 1 2 3 4 \n# %le 5 6 \n" % (math.log(0.01), math.log(0.02)))
    r = _child("L.read_all_nu_state(C.byref(d), %r.encode())" % str(short))
    assert "terminate(2006)" in r.stderr and r.returncode == 2006 % 256, r.stderr
    late = tmp_path / "late.txt"
    late.write_text("# %le 1 2 3 4 \n" % math.log(0.02))
    r = _child("L.read_all_nu_state(C.byref(d), %r.encode())" % str(late))
    assert "terminate(2007)" in r.stderr and r.returncode == 2007 % 256, r.stderr


def test_save_nu_power_format(kn, tmp_path):
    """delta_tot_table.c:356-374: header, then k and delta_nu_last^2 per bin ("%g %g")."""
    om, d = _table(kn, 4)
    for i in range(4):
        d.wavenum[i] = 0.01 * (i + 1)
        d.delta_nu_last[i] = 1.5 + i
    assert kn.save_nu_power(C.byref(d), 0.5, 7, str(tmp_path).encode()) == 0
    text = (tmp_path / "powerspec_nu_007.txt").read_text().splitlines()
    assert text[:3] == ["# k P_nu(k)", "# a = 0.5", "# nbins = 4"]
    assert text[3:] == ["%g %g" % (0.01 * (i + 1), (1.5 + i) ** 2) for i in range(4)]
    assert kn.save_nu_power(C.byref(d), 0.5, 7, str(tmp_path / "missing" / "dir").encode()) == -1
    kn.free_delta_tot_table(C.byref(d))


def test_module_resume_and_save(tmp_path):
    """allocate_kspace_memory with a saved history resumes from it
    (interface_common.c:93-95); save_nu_state writes it back unchanged and
    save_neutrino_power writes P_nu(k) (interface_common.c:41-52)."""
    out = tmp_path / "state.txt"
    code = (
        "import sys, ctypes as C; sys.path[:0]=['kspace-neutrinos_amd','tests'];"
        "import conftest; conftest.load_testhost(); import knhip; L=knhip.lib();"
        "kp=knhip.KspaceParams.in_dll(L,'kspace_params');"
        "kp.KspaceTransferFunction=%r.encode(); kp.TimeTransfer=0.01;"
        "kp.InputSpectrum_UnitLength_in_cm=conftest.UNIT_LENGTH*1e3;"
        "kp.MNu[0]=kp.MNu[1]=kp.MNu[2]=0.15; kp.hybrid_neutrinos_on=0;"
        "L.InitOmegaNu(0.7,2.7255,0);"
        "L.allocate_kspace_memory(300,0,512000.0,conftest.UNIT_TIME,conftest.UNIT_LENGTH,0.2793,%r.encode(),1.0,0);"
        "sf=C.POINTER(C.c_double)(); dt=C.POINTER(C.c_double)(); nk=C.c_size_t(); ia=C.c_size_t();"
        "L.get_nu_state(C.byref(sf),C.byref(dt),C.byref(nk),C.byref(ia)); print('state', nk.value, ia.value);"
        "L.save_nu_state(%r.encode());"
        "print('power', L.save_neutrino_power(0.5, 3, %r.encode()))"
        % (os.path.join(REFDATA, "ics_transfer_99.dat"), STATE, str(out), str(tmp_path)))
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "state 300 25" in r.stdout and "power 0" in r.stdout
    assert out.read_text() == open(STATE).read()
    assert (tmp_path / "powerspec_nu_003.txt").read_text().startswith("# k P_nu(k)\n# a = 0.5\n# nbins = 300\n")
