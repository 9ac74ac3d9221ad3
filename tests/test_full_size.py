"""Size-independent properties of the single large solve at its full size
(MCV4, 6 x 384 x 384, RK4, dt = 10 s, the balanced jet with its degree-16
perturbation), where the oracle cannot follow for long:

  - total mass is conserved to round-off (the scheme's flux form telescopes;
    the reference's plain sum drifts by ~1e-13, the device's double-double
    integral reads 0 to 1e-14);
  - stepping is deterministic and splits: 300 steps in one call equal
    150 + 150, bit for bit, and equal the same steps with another stage-kernel
    tile shape;
  - the FP32-compensated form stays within its stated distance of the FP64
    trajectory (h relative L2 <= 1e-6 after 300 steps) and conserves mass to
    1e-11.
The first 3 steps at this size are pinned against the reference itself
(tests/test_golden.py)."""
import numpy as np
import pytest

from mcv_sw_amd import capi, hostlib

N, DOF, DT, STEPS = 384, 4, 10.0, 300


@pytest.fixture(scope="module")
def case():
    print("[full size] building the 6x384x384 mesh and jet", flush=True)
    m = hostlib.mesh(DOF, N)
    st = hostlib.jet(m, u0=40.0, h0=8000.0, lmax=16, seed=0, rms_h=60.0)
    return m, st


def _owned(a):
    h = DOF - 1
    npv = N * h + 1
    return a[:, h:h + npv, h:h + npv]


@pytest.mark.gpu
def test_full_size_mass_and_determinism(case, hip_lib):
    m, st = case
    print("[full size] FP64, 300 steps, three ways", flush=True)
    with capi.Solver(DOF, N, DT, m, st, lib=hip_lib) as s:
        m0 = s.sync(fields=())["mass"]
        s.step(STEPS)
        a = s.sync(fields=("phi", "u", "v"))
    assert abs(a["mass"] - m0) <= 1e-14 * abs(m0)
    with capi.Solver(DOF, N, DT, m, st, lib=hip_lib) as s:
        s.step(STEPS // 2)
        s.sync(fields=(), integrals=False)
        s.step(STEPS - STEPS // 2)
        b = s.sync(fields=("phi", "u", "v"))
    with capi.Solver(DOF, N, DT, m, st, lib=hip_lib) as s:
        s.set_tile(3)
        s.step(STEPS)
        c = s.sync(fields=("phi", "u", "v"))
    for k in ("phi", "u", "v"):
        assert np.array_equal(_owned(a[k]), _owned(b[k])), k
        assert np.array_equal(_owned(a[k]), _owned(c[k])), k


@pytest.mark.gpu
def test_full_size_fp32c_distance(case, hip_lib):
    m, st = case
    print("[full size] FP32C vs FP64, 300 steps", flush=True)
    out = []
    for prec in (capi.FP64, capi.FP32C):
        with capi.Solver(DOF, N, DT, m, st, precision=prec, lib=hip_lib) as s:
            m0 = s.sync(fields=())["mass"]
            s.step(STEPS)
            r = s.sync(fields=("phi", "u", "v"))
            r["mass0"] = m0
            out.append(r)
    ref, got = out
    x, y = _owned(ref["phi"]), _owned(got["phi"])
    assert np.linalg.norm(y - x) / np.linalg.norm(x) <= 1e-6
    for k in ("u", "v"):
        x, y = _owned(ref[k]), _owned(got[k])
        assert np.linalg.norm(y - x) / np.linalg.norm(x) <= 1e-5, k
    assert abs(got["mass"] - got["mass0"]) <= 1e-11 * abs(got["mass0"])
