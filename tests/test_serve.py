"""The continuous-batching request loop: every request's result is the same
alone and in the mix (bit for bit, FP64 and FP32C alike) and matches the
oracle (FP64 bit for bit, FP32C within tolerance), admission respects
the HBM budget and the slot count, and the report is consistent."""
import numpy as np
import pytest

import cases
from mcv_sw_amd import capi, hostlib, serve
from oracle import oracle


def _requests(sizes, hours=(0.5, 2.0), rate=400.0, count=10, seed=3):
    reqs = serve.draw_requests(rate, 10.0, seed=seed, hours=hours, sizes=sizes,
                               weights=[1.0] * len(sizes))[:count]
    assert len(reqs) == count
    return reqs


def _alone(lib, r):
    m = hostlib.mesh(r.dof, r.n)
    st = hostlib.jet(m, u0=r.u0, h0=r.h0, lmax=r.lmax, seed=r.seed, rms_h=r.rms_h)
    with capi.Solver(r.dof, r.n, r.dt, m, st, scheme=r.scheme, precision=r.precision,
                     lib=lib) as s:
        s.step(r.nsteps)
        out = s.sync(fields=("phi", "u", "v"))
    # the oracle's FP64 step on the same inputs (time_scheme_mod.F90:13-142)
    out["oracle"] = oracle.run(r.dof, r.n, 4 if r.scheme == "RK4" else 3, r.dt, m, st, r.nsteps)
    return out


def _check_serve(lib, reqs, **kw):
    srv = serve.Server(lib=lib, **kw)
    rep, res = srv.run(reqs)
    assert rep["completed"] == len(reqs)
    assert rep["latency_p50_s"] <= rep["latency_p99_s"] <= rep["latency_max_s"]
    for row in rep["per_request"]:
        assert row["arrival"] <= row["admit"] <= row["done"]
    for r in reqs:
        a = _alone(lib, r)
        for k in ("phi", "u", "v"):
            assert np.array_equal(res[r.rid][k], a[k]), (r, k)
        assert res[r.rid]["mass"] == a["mass"], r
        ref, got = a["oracle"], res[r.rid]
        if r.precision == capi.FP64:  # bit for bit on every defined point
            msk = cases.masks(r.dof, r.n)["defined"]
            for k in ("phi", "u", "v"):
                assert np.array_equal(got[k][:, msk], ref[k][:, msk]), (r, k)
        else:  # FP32C within tolerance of the FP64 trajectory
            msk = cases.masks(r.dof, r.n)["interior"]
            for k, tol in (("phi", 1e-6), ("u", 1e-5), ("v", 1e-5)):
                assert cases.rel_l2(got[k][:, msk], ref[k][:, msk]) <= tol, (r, k)
    return rep


def test_emulated_serve_results_equal_alone(emu_lib):
    reqs = _requests(sizes=(4, 5, 6))
    assert {r.precision for r in reqs} == {capi.FP64, capi.FP32C}
    assert {r.dof for r in reqs} == {3, 4}
    _check_serve(emu_lib, reqs, capacity=4, stages_per_poll=5, workers=2)


def test_emulated_serve_admission_by_memory(emu_lib):
    reqs = _requests(sizes=(4, 6), count=6, seed=5)
    big = max(capi.footprint(r.dof, r.n, precision=r.precision, lib=emu_lib) for r in reqs)
    budget = int(1.5 * big)
    rep = _check_serve(emu_lib, reqs, capacity=8, hbm_budget=budget, stages_per_poll=3,
                       workers=2)
    fp = {(r.dof, r.n, r.precision): capi.footprint(r.dof, r.n, precision=r.precision,
                                                     lib=emu_lib) for r in reqs}
    rows = rep["per_request"]
    by_rid = {r.rid: r for r in reqs}
    for row in rows:  # the solves resident when this one was admitted
        t = row["admit"]
        live = [x for x in rows if x["admit"] <= t < x["done"]]
        need = sum(fp[(by_rid[x["rid"]].dof, by_rid[x["rid"]].n, by_rid[x["rid"]].precision)]
                   for x in live)
        assert len(live) == 1 or need <= budget, (t, need, budget)


def test_cfl_dt_scales_with_resolution():
    a = serve.cfl_dt(3, 32, 20.0, 5960.0)
    b = serve.cfl_dt(3, 64, 20.0, 5960.0)
    c = serve.cfl_dt(4, 32, 20.0, 5960.0)
    assert abs(a / b - 2.0) < 1e-12 and abs(a / c - 1.5) < 1e-12


def test_draw_requests_follows_the_distribution():
    reqs = serve.draw_requests(20.0, 600.0, seed=0)
    assert 11000 < len(reqs) < 13000
    n = np.array([r.n for r in reqs])
    assert abs(np.mean(n == 32) - 0.3) < 0.02 and abs(np.mean(n == 384) - 0.05) < 0.01
    assert abs(np.mean([r.dof == 4 for r in reqs]) - 0.5) < 0.02
    assert abs(np.mean([r.precision == capi.FP32C for r in reqs]) - 0.5) < 0.02
    h = np.array([r.hours for r in reqs])
    assert h.min() >= 6.0 and h.max() <= 48.0


@pytest.mark.gpu
def test_serve_results_equal_alone(hip_lib):
    print("[serve] mixed requests", flush=True)
    reqs = _requests(sizes=(8, 16, 24, 32), hours=(2.0, 8.0), rate=50.0, count=16, seed=11)
    _check_serve(hip_lib, reqs, capacity=8, stages_per_poll=12, workers=4)
