"""The C-ABI boundary: exported symbols, argument checking and error codes.

Symbol checks load the real HIP library (no GPU call is made).  Error paths
that need a compute call run on the emulated build of the same sources.
"""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import cases
from mcv_sw_amd import capi, hostlib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(mcv_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = declared("mcv_hip.h")
    for n in ("mcv_hip_init", "mcv_hip_step", "mcv_hip_sync", "mcv_hip_free",
              "mcv_hip_last_error", "mcv_hip_tendency", "mcv_hip_sphere_winds"):
        assert n in names


def test_hip_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared("mcv_hip.h") if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_exports():
    lib = ctypes.CDLL(hostlib.LIB_PATH)
    for n in ("mcv_host_mesh", "mcv_host_ghost", "mcv_host_case", "mcv_host_jet",
              "mcv_host_handoff"):
        assert hasattr(lib, n), n


def test_no_error_before_first_call():
    lib = capi.load()
    assert lib.mcv_hip_last_error() in (b"",) or isinstance(lib.mcv_hip_last_error(), bytes)


def _inputs(dof=3, n=4):
    m = hostlib.mesh(dof, n)
    return m, hostlib.williamson(m, 2)


@pytest.mark.parametrize("field,value,msg", [
    ("dof", 5, "dof must be 3 or 4"),
    ("n", 1, "n must be in"),
    ("scheme", 7, "Unknown integral scheme"),
    ("dt", -1.0, "dt must be > 0"),
    ("dt", float("nan"), "dt must be > 0"),
    ("precision", 9, "precision"),
])
def test_invalid_config_is_rejected(emu_lib, field, value, msg):
    m, st = _inputs()
    kw = dict(dof=3, n=4, dt=600.0, scheme="RK4", precision=capi.FP64)
    kw[field] = value
    if field in ("dof", "n"):
        with pytest.raises(capi.MCVError) as e:
            capi.Solver(kw["dof"], kw["n"], kw["dt"], m, st, scheme=kw["scheme"],
                        precision=kw["precision"], lib=emu_lib)
    else:
        with pytest.raises(capi.MCVError) as e:
            capi.Solver(3, 4, kw["dt"], m, st, scheme=kw["scheme"], precision=kw["precision"],
                        lib=emu_lib)
    assert e.value.code == capi.MCV_EINVAL
    assert msg in str(e.value)


def test_asymmetric_metric_is_rejected(emu_lib):
    m, st = _inputs()
    m["IG"] = m["IG"].copy()
    m["IG"][0, 3, 3, 0, 1] *= 1.0 + 1e-15
    with pytest.raises(capi.MCVError, match="not symmetric"):
        capi.Solver(3, 4, 600.0, m, st, lib=emu_lib)


def test_inconsistent_phiG_is_rejected(emu_lib):
    m, st = _inputs()
    st["phiG"] = st["phiG"] * (1.0 + 1e-12)
    with pytest.raises(capi.MCVError, match="phiG"):
        capi.Solver(3, 4, 600.0, m, st, lib=emu_lib)


def test_bad_ghost_table_is_rejected(emu_lib):
    m, st = _inputs()
    m["iref"] = m["iref"].copy()
    m["iref"][0, 0] = 0
    with pytest.raises(capi.MCVError, match="ghost_iref"):
        capi.Solver(3, 4, 600.0, m, st, lib=emu_lib)


def test_negative_step_count_is_rejected(emu_lib):
    m, st = _inputs()
    with capi.Solver(3, 4, 600.0, m, st, lib=emu_lib) as s:
        with pytest.raises(capi.MCVError) as e:
            s.step(-1)
        assert e.value.code == capi.MCV_EINVAL
        s.step(0)
        assert s.steps_done == 0


def test_nan_in_initial_state_is_reported_by_the_integrals(emu_lib):
    m, st = _inputs()
    st["phi"] = st["phi"].copy()
    st["phi"][2, 5, 5] = np.nan
    st["phiG"] = st["phi"] * m["sqrtG"]
    with capi.Solver(3, 4, 600.0, m, st, lib=emu_lib) as s:
        s.step(1)
        with pytest.raises(capi.MCVError) as e:
            s.sync()
        assert e.value.code == capi.MCV_ENONFINITE


def test_null_arguments_are_rejected(emu_lib):
    L = capi.load(emu_lib)
    h = ctypes.c_void_p()
    assert L.mcv_hip_init(None, None, None, ctypes.byref(h)) == capi.MCV_EINVAL
    assert L.mcv_hip_step(None, 1) == capi.MCV_EINVAL
    assert L.mcv_hip_sync(None, None, None, None) == capi.MCV_EINVAL
    assert L.mcv_hip_free(None) == capi.MCV_OK
    assert b"NULL" in L.mcv_hip_last_error()


def test_sync_with_partial_outputs(emu_lib):
    m, st = _inputs()
    with capi.Solver(3, 4, 600.0, m, st, lib=emu_lib) as s:
        s.step(1)
        out = s.sync(fields=("phi",), integrals=False)
        assert set(out) == {"phi"}


def test_asan_build_of_every_stage_kind():
    """Every stage kind, both orders, partial tiles: no out-of-bounds access."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))
    import build_emu

    exe = build_emu.build_asan()
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]


def test_tsan_build_finds_no_race_within_a_block():
    """Each block run by one host thread per GPU thread with real barriers,
    under ThreadSanitizer: no two threads of a block touch the same LDS or
    global word without a __syncthreads() between them, in any stage kind,
    tile shape, boundary, diagnostic or batched kernel (a removed barrier in
    the stage kernel is reported as a data race by this build)."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))
    import build_emu

    exe = build_emu.build_tsan()
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    p = subprocess.run([exe, "threads"], capture_output=True, text=True, timeout=900, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-6000:]


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
