"""The Fortran host program (mcv_sw_amd/host/mcv_sw_driver.F90) driving the
C-ABI through iso_c_binding, as the reference's MCV_SW.F90 would.

On the CPU it is linked against the emulated product library; on the GPU
against the HIP build.  Its final state must equal, bit for bit, the same
solve driven from Python through the same C-ABI, and it prints the
reference's history lines (MCV_SW.F90:51,72)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from mcv_sw_amd import build, capi, hostlib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

NAMELIST = """&time_settings
run_days = 0
run_hours = {hours}
dt = {dt}
history_interval = 3600
integral_scheme = '{scheme}'
/
&case_select
case_num = {case}
/
&domain
dx = {dx}
dy = {dx}
/
&gpu_settings
dof = {dof}
precision = '{prec}'
dump_file = 'final.bin'
/
&jet_settings
u0 = 30.0
h0 = 7000.0
lmax = 6
seed = 4
rms_h = 40.0
/
"""


def _run_driver(exe, wd, dof, n, dt, hours, scheme, case, prec):
    os.makedirs(wd, exist_ok=True)
    with open(os.path.join(wd, "namelist.input"), "w") as f:
        f.write(NAMELIST.format(hours=hours, dt=dt, scheme=scheme, case=case, dx=90.0 / n,
                                dof=dof, prec=prec))
    p = subprocess.run([exe], cwd=wd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    h = dof - 1
    nh = n * h + 1 + 2 * h
    raw = np.fromfile(os.path.join(wd, "final.bin"), dtype=np.float64)
    phi, u, v = raw.reshape(3, 6, nh, nh)
    return p.stdout, {"phi": phi, "u": u, "v": v}


def _python_path(lib, dof, n, dt, hours, scheme, case, prec):
    m = hostlib.mesh(dof, n)
    st = (hostlib.jet(m, u0=30.0, h0=7000.0, lmax=6, seed=4, rms_h=40.0) if case == 0
          else hostlib.williamson(m, case))
    with capi.Solver(dof, n, dt, m, st, scheme=scheme, precision=prec, lib=lib) as s:
        s.step(int(hours * 3600 / dt))
        return s.sync(fields=("phi", "u", "v"))


CASES = [(3, 6, 600.0, 2, "RK4", 2, capi.FP64), (4, 5, 400.0, 2, "RK3_TVD", 5, capi.FP64),
         (3, 5, 600.0, 1, "RK4", 0, capi.FP32C)]


def _check(lib, exe, tmp_path, spec):
    dof, n, dt, hours, scheme, case, prec = spec
    out, fin = _run_driver(exe, str(tmp_path / f"run{dof}{n}{case}"), dof, n, dt, hours, scheme,
                           case, "FP32C" if prec == capi.FP32C else "FP64")
    lines = [ln for ln in out.splitlines() if "output index/total" in ln]
    assert len(lines) == hours + 1, out
    ref = _python_path(lib, dof, n, dt, hours, scheme, case, prec)
    for k in ("phi", "u", "v"):
        assert np.array_equal(fin[k], ref[k]), k


@pytest.fixture(scope="module")
def emu_driver(emu_lib):
    out = os.path.join(ROOT, "tests", "emu", "_build", "mcv_sw_gpu_emu")
    return build.build_driver(hip_lib=emu_lib, out=out)


@pytest.mark.parametrize("spec", CASES)
def test_emulated_fortran_host_matches_python_path(emu_lib, emu_driver, tmp_path, spec):
    _check(emu_lib, emu_driver, tmp_path, spec)


def test_fortran_host_reports_library_errors(emu_driver, tmp_path):
    # dt = 0 passes the namelist but the library refuses it with a message
    out = subprocess.run([emu_driver], cwd=tmp_path, capture_output=True, text=True)
    assert out.returncode != 0  # no namelist.input
    wd = tmp_path / "bad"
    os.makedirs(wd)
    with open(wd / "namelist.input", "w") as f:
        f.write(NAMELIST.format(hours=1, dt=-60.0, scheme="RK4", case=2, dx=15.0, dof=3,
                                prec="FP64"))
    p = subprocess.run([emu_driver], cwd=wd, capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "mcv_hip_init failed: dt must be > 0" in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("spec", CASES)
def test_fortran_host_matches_python_path(hip_lib, tmp_path, spec):
    print(f"[fortran host] {spec}", flush=True)
    exe = build.DRIVER
    assert os.path.exists(exe), "build the host program with __graft_entry__.build()"
    _check(hip_lib, exe, tmp_path, spec)
