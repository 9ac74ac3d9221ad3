"""mcv_hip_tendency_noalloc: the tendency seam without allocation.

The entry point runs the same stage kernel as mcv_hip_tendency on the
library's static per-device scratch array, so its outputs must equal that
entry point's bit for bit.  Two calls in a row, the second on a smaller
mesh, check that the scratch left by one call does not leak into the next;
a mesh beyond the scratch is refused with MCV_EINVAL.
"""
import numpy as np
import pytest

import cases
from mcv_sw_amd import capi


def _state(dof, n, amp):
    m = cases.mesh(dof, n)
    return m, cases.bumped_jet(m, u0=25.0, h0=6000.0, amp=amp)


def _both(lib, dof, n, amp):
    m, st = _state(dof, n, amp)
    with capi.Solver(dof, n, 300.0, m, st, lib=lib) as s:
        a = s.tendency(st, noalloc=True)
        b = s.tendency(st)
    return a, b


def _check(lib):
    # larger mesh first, then a smaller one on the same device's scratch
    for dof, n, amp in ((4, 40, 50.0), (3, 9, 20.0), (4, 40, 35.0)):
        a, b = _both(lib, dof, n, amp)
        for k in ("phiG", "u", "v"):
            assert np.array_equal(a[k], b[k]), (dof, n, k)


def _too_large(lib):
    dof, n = 4, 130  # (3n+7)^2 * 6 points: 16 B + 24 B each exceed 32 MiB
    m, st = _state(dof, n, 10.0)
    with capi.Solver(dof, n, 60.0, m, st, lib=lib) as s:
        with pytest.raises(capi.MCVError) as e:
            s.tendency(st, noalloc=True)
        assert e.value.code == capi.MCV_EINVAL
        assert "scratch" in str(e.value)


def test_emulated_noalloc_tendency_equals_tendency(emu_lib):
    _check(emu_lib)


def test_emulated_noalloc_tendency_refuses_meshes_beyond_its_scratch(emu_lib):
    _too_large(emu_lib)


@pytest.mark.gpu
def test_noalloc_tendency_equals_tendency(hip_lib):
    print("[noalloc] tendency seam, two meshes in a row", flush=True)
    _check(hip_lib)


@pytest.mark.gpu
def test_noalloc_tendency_refuses_meshes_beyond_its_scratch(hip_lib):
    print("[noalloc] refusal beyond the scratch", flush=True)
    _too_large(hip_lib)
