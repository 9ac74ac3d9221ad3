"""The batched and served workloads at their configured sizes, anchored to
the oracle.

Each test runs one of the many-solve workloads on the MI355X through the
batch C-ABI, with graph replay on, and checks its solves against the oracle's
FP64 restatement of the reference step (time_scheme_mod.F90:13-142, one solve
per process in the reference, MCV_SW.F90:48-77) on identical inputs:
  - FP64 solves equal the oracle's arrays bit for bit;
  - FP32C solves are within the stated tolerance of the oracle's FP64
    trajectory (relative L2 over the panels' own points: h 1e-6, u and v
    1e-5) with their FP64 mass integral within 1e-9 of the initial mass (the
    FP32 flux arithmetic conserves mass to its own rounding, not FP64's).
Solves too large for the oracle to finish in seconds on the host are checked
against the FP64 drop-in run alone on the GPU (itself bit-identical to the
oracle and, at 6x384x384, to the reference build).

Time steps: the configured 300 s (ensemble) is unstable for the reference
scheme at 6x90x90 (point Courant ~1.4); 60 s is used.  Uneven batches and
served requests take dt at Courant 0.5 (serve.cfl_dt).  Runs are a few steps
long (at least one 12-stage graph replay), not the configured days.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import cases
from mcv_sw_amd import capi, hostlib, serve
from oracle import oracle

FIELDS = ("phi", "phiG", "u", "v", "contraU", "contraV")
TOL = {"phi": 1e-6, "u": 1e-5, "v": 1e-5}
WORKERS = 16
# workload sizes: (members, n, steps) of the ensemble; panel sizes, count and
# steps of the uneven batch; the largest size checked against the oracle
ENSEMBLE = (64, 90, 8)
UNEVEN = ((32, 48, 90, 128, 180, 256), 128, 3)
ORACLE_MAX_N = 90
# served mix: panel sizes (None: the configured {32, 64, 128, 256, 384}), the
# size of the two large requests added to it, the largest size checked
# against the oracle
SERVE = (None, 384, 64)


def _oracle_runs(jobs):
    """oracle.run over a thread pool (the C library runs without the GIL);
    jobs: list of (dof, n, scheme, dt, mesh, state, nsteps)."""
    with cf.ThreadPoolExecutor(WORKERS) as pool:
        return list(pool.map(lambda j: oracle.run(j[0], j[1], 4 if j[2] == "RK4" else 3,
                                                  *j[3:]), jobs))


def _bitwise(got, ref, dof, n):
    msk = cases.masks(dof, n)["defined"]
    bad = [k for k in FIELDS if not np.array_equal(got[k][:, msk], ref[k][:, msk])]
    return bad


def _fp32c_err(got, ref, dof, n):
    msk = cases.masks(dof, n)["interior"]
    return {k: cases.rel_l2(got[k][:, msk], ref[k][:, msk]) for k in TOL}


def _check_fp32c(tag, got, ref, dof, n, mass0):
    e = _fp32c_err(got, ref, dof, n)
    drift = abs(got["mass"] - mass0) / abs(mass0)
    assert all(e[k] <= TOL[k] for k in TOL), (tag, e)
    assert drift <= 1e-9, (tag, drift)
    return e


def _drain(b, slots, chunk=12, limit=400):
    for _ in range(limit):
        left, _ = b.poll()
        if all(left[s] == 0 for s in slots):
            return
        b.run(chunk)
    raise AssertionError("solves did not finish")


@pytest.mark.gpu
def test_ensemble_64_members_mcv3_n90_against_oracle(hip_lib):
    """The uniform ensemble: 64 FP32C members of MCV3 on 6x90x90, RK3_TVD,
    member k perturbed with seed k (degree 12, 20 m rms) on the balanced jet
    u0 = 20 m/s, h0 = 5960 m, in one batch with one fused stage launch per RK
    stage, replayed from a hipGraph; beside them the FP64 twins of members 1
    and 64.  Every member is compared with the oracle after 8 steps."""
    members, n, steps = ENSEMBLE
    dof, dt, scheme = 3, 60.0, "RK3_TVD"
    print(f"[cfg ensemble] preparing {members} members", flush=True)
    m = hostlib.mesh(dof, n)
    with cf.ThreadPoolExecutor(WORKERS) as pool:
        states = list(pool.map(lambda k: hostlib.jet(dict(m), u0=20.0, h0=5960.0, lmax=12,
                                                     seed=k, rms_h=20.0),
                               range(1, members + 1)))
    twins = (0, members - 1)
    with capi.Batch(members + 2, lib=hip_lib) as b:
        b.set_graphs(True)
        slots = [b.add(dof, n, dt, m, st, steps, scheme=scheme, precision=capi.FP32C)
                 for st in states]
        tw = [b.add(dof, n, dt, m, states[k], steps, scheme=scheme) for k in twins]
        mass0 = [b.sync(s, fields=())["mass"] for s in slots]
        print("[cfg ensemble] running", flush=True)
        _drain(b, slots + tw)
        got = [b.sync(s) for s in slots]
        got_tw = [b.sync(s) for s in tw]
    print("[cfg ensemble] oracle", flush=True)
    ref = _oracle_runs([(dof, n, scheme, dt, m, st, steps) for st in states])
    for k, t in zip(twins, got_tw):
        assert not _bitwise(t, ref[k], dof, n), f"FP64 twin of member {k + 1}"
    worst = {f: 0.0 for f in TOL}
    for k in range(members):
        e = _check_fp32c(f"member {k + 1}", got[k], ref[k], dof, n, mass0[k])
        worst = {f: max(worst[f], e[f]) for f in TOL}
    print(f"[cfg ensemble] FP32C worst rel L2 vs oracle after {steps} steps: {worst}", flush=True)


def _uneven_specs(sizes, count, seed=0):
    """The uneven batch: MCV3/MCV4 50/50, n uniform in `sizes` ({32, 48, 90,
    128, 180, 256}), u0 ~ U(10, 40) m/s, perturbation rms ~ U(10, 50) m,
    h0 = 5960 m."""
    rng = np.random.default_rng(seed)
    return [dict(dof=4 if rng.random() < 0.5 else 3, n=int(rng.choice(sizes)),
                 u0=float(rng.uniform(10.0, 40.0)), rms=float(rng.uniform(10.0, 50.0)),
                 seed=2000 + k) for k in range(count)]


@pytest.mark.gpu
def test_uneven_batch_128_solves_against_oracle(hip_lib):
    """128 FP32C RK4 solves of mixed order and size (n up to 256) in one batch
    with CFL time steps, graph replay on, plus FP64 twins of the smallest and
    the largest solve of each order.  After 3 steps: the twins equal the
    oracle bit for bit; FP32C solves with n <= 90 are within tolerance of the
    oracle; the larger ones within tolerance of the FP64 drop-in run alone."""
    sizes, count, steps = UNEVEN
    specs = _uneven_specs(sizes, count)
    print(f"[cfg uneven] preparing {count} solves", flush=True)
    meshes = {key: hostlib.mesh(*key) for key in sorted({(s["dof"], s["n"]) for s in specs})}

    def prep(sp):
        return hostlib.jet(dict(meshes[(sp["dof"], sp["n"])]), u0=sp["u0"], h0=5960.0, lmax=8,
                           seed=sp["seed"], rms_h=sp["rms"])

    with cf.ThreadPoolExecutor(WORKERS) as pool:
        states = list(pool.map(prep, specs))
    dts = [serve.cfl_dt(sp["dof"], sp["n"], sp["u0"], 5960.0) for sp in specs]
    twins = []
    for dof in (3, 4):
        ks = [k for k, sp in enumerate(specs) if sp["dof"] == dof]
        twins += [min(ks, key=lambda k: specs[k]["n"]), max(ks, key=lambda k: specs[k]["n"])]
    assert {specs[k]["n"] for k in twins} >= {min(sizes), max(sizes)}
    with capi.Batch(len(specs) + len(twins), lib=hip_lib) as b:
        b.set_graphs(True)
        slots = [b.add(sp["dof"], sp["n"], dt, meshes[(sp["dof"], sp["n"])], st, steps,
                       scheme="RK4", precision=capi.FP32C)
                 for sp, st, dt in zip(specs, states, dts)]
        tw = [b.add(specs[k]["dof"], specs[k]["n"], dts[k], meshes[(specs[k]["dof"],
                                                                   specs[k]["n"])],
                    states[k], steps, scheme="RK4") for k in twins]
        mass0 = [b.sync(s, fields=())["mass"] for s in slots]
        tile, _ = b.autotune(reps=2)
        print(f"[cfg uneven] running (tile variant {tile})", flush=True)
        _drain(b, slots + tw)
        got = [b.sync(s, fields=("phi", "u", "v")) for s in slots]
        got_tw = [b.sync(s) for s in tw]
    small = [k for k, sp in enumerate(specs) if sp["n"] <= ORACLE_MAX_N]
    print(f"[cfg uneven] oracle: {len(twins)} twins, {len(small)} small solves", flush=True)
    jobs = [(specs[k]["dof"], specs[k]["n"], "RK4", dts[k], meshes[(specs[k]["dof"],
                                                                    specs[k]["n"])],
             states[k], steps) for k in twins + small]
    ref = _oracle_runs(jobs)
    for i, k in enumerate(twins):
        sp = specs[k]
        assert not _bitwise(got_tw[i], ref[i], sp["dof"], sp["n"]), f"FP64 twin {sp}"
    ref_of = dict(zip(twins + small, ref))
    for k, sp in enumerate(specs):
        if k not in ref_of:  # too large for the oracle here: the FP64 drop-in alone
            with capi.Solver(sp["dof"], sp["n"], dts[k], meshes[(sp["dof"], sp["n"])],
                             states[k], scheme="RK4", lib=hip_lib) as s:
                s.step(steps)
                ref_of[k] = s.sync(fields=("phi", "u", "v"))
        _check_fp32c(f"solve {k} {sp}", got[k], ref_of[k], sp["dof"], sp["n"], mass0[k])


@pytest.mark.gpu
def test_serving_configured_mix_against_oracle(hip_lib):
    """The serving loop on the configured request mix (order MCV3/MCV4,
    n in {32, 64, 128, 256, 384} with weights 0.3/0.3/0.2/0.15/0.05, FP64 or
    FP32C), shortened to 0.05-0.3 simulated hours per request and with
    arrivals compressed so that requests join a batch that is mid-run, plus
    one MCV4 6x384x384 request of each precision.  Every served result: FP64
    equal to the oracle (n <= 64) or to the drop-in run alone (larger), FP32C
    within tolerance of the same FP64 trajectory."""
    mix, big_n, oracle_n = SERVE
    kw = {} if mix is None else {"sizes": mix, "weights": [1.0] * len(mix)}
    reqs = serve.draw_requests(200.0, 10.0, seed=21, hours=(0.05, 0.3), **kw)[:22]
    big = [serve.Request(rid=1000 + i, t_arrival=reqs[-1].t_arrival, dof=4, n=big_n,
                         hours=0.05, precision=p, seed=77, u0=25.0, rms_h=30.0)
           for i, p in enumerate((capi.FP64, capi.FP32C))]
    reqs = reqs + big
    sizes = {r.n for r in reqs}
    assert len(sizes) >= 3 and big_n in sizes
    assert {r.precision for r in reqs} == {capi.FP64, capi.FP32C}
    print(f"[cfg serve] {len(reqs)} requests, sizes {sorted(sizes)}", flush=True)
    srv = serve.Server(capacity=16, stages_per_poll=12, workers=WORKERS, lib=hip_lib)
    rep, res = srv.run(reqs, keep=FIELDS)
    assert rep["completed"] == len(reqs) and not rep["non_finite"]
    for row in rep["per_request"]:
        assert row["arrival"] <= row["admit"] <= row["done"]
    # several requests were on the GPU together
    rows = rep["per_request"]
    assert max(sum(1 for y in rows if y["admit"] <= x["admit"] < y["done"]) for x in rows) >= 2
    meshes, states, ref_of = {}, {}, {}
    for r in reqs:
        key = (r.dof, r.n)
        if key not in meshes:
            meshes[key] = hostlib.mesh(*key)
        states[r.rid] = hostlib.jet(dict(meshes[key]), u0=r.u0, h0=r.h0, lmax=r.lmax,
                                    seed=r.seed, rms_h=r.rms_h)
    small = [r for r in reqs if r.n <= oracle_n]
    print(f"[cfg serve] oracle on {len(small)} requests", flush=True)
    for r, ref in zip(small, _oracle_runs([(r.dof, r.n, r.scheme, r.dt, meshes[(r.dof, r.n)],
                                            states[r.rid], r.nsteps) for r in small])):
        ref_of[r.rid] = ref
    for r in reqs:
        if r.rid not in ref_of:
            with capi.Solver(r.dof, r.n, r.dt, meshes[(r.dof, r.n)], states[r.rid],
                             scheme=r.scheme, lib=hip_lib) as s:
                s.step(r.nsteps)
                ref_of[r.rid] = s.sync()
        got, ref = res[r.rid], ref_of[r.rid]
        if r.precision == capi.FP64:
            assert not _bitwise(got, ref, r.dof, r.n), r
        else:
            with capi.Solver(r.dof, r.n, r.dt, meshes[(r.dof, r.n)], states[r.rid],
                             scheme=r.scheme, precision=capi.FP32C, lib=hip_lib) as s:
                mass0 = s.sync(fields=())["mass"]
            _check_fp32c(f"request {r.rid}", got, ref, r.dof, r.n, mass0)
