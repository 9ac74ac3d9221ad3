#!/usr/bin/env python3
"""Benchmark of the MCV_SW time step on one MI355X.

Headline workload (value): the single large solve, MCV4 on the 6 x 384 x 384
gnomonic cubed sphere (~0.23 deg), RK4, FP64 drop-in (bit-identical to the
reference), balanced jet (u0 = 40 m/s, h0 = 8000 m) plus a seeded (seed 0)
degree-16 spherical-harmonic height perturbation at 60 m rms, dt = 10 s.
One step = one full RK4 time step of all 884,736 cells; the state is
resident in HBM before the timed region starts.

Also reported on the same line:
  fp32c        the FP32-compensated form on the same solve and step count:
               cell-steps/s, and its h/u/v relative L2 and Linf distance and
               mass drift from the FP64 trajectory after the same steps
  batched      the uniform ensemble: 64 members of MCV3 on 6 x 90 x 90, RK3,
               dt = 60 s, FP32C, member k perturbed with seed k (degree 12,
               20 m rms) on the balanced jet u0 = 20 m/s, h0 = 5960 m; one
               fused stage launch + one fused boundary launch per RK stage for
               all members, replayed from a hipGraph
  roofline     the FP64 stage kernel's achieved algorithmic HBM bandwidth
               (51 FP64 words per point per RK4 step, 9 points per MCV4 cell)
               from HIP events on the solver's stream, against 8 TB/s
  batched_uneven  128 solves of mixed order and size (n in 32..256) in one
               batch, FP32C, RK4, CFL time steps; cell-steps/s of all solves
  serve        the serving loop: Poisson arrivals of seeded solves of mixed
               order, size, length and precision (the configured mix) at a
               rate one GPU sustains (default 0.5 req/s for 60 s), continuous
               batching with admission by HBM footprint; offered and served
               cell-steps/s, steady-state p50/p99 latency, and the offered
               load of the configured 20 req/s
  cpu_baseline the reference Fortran (oracle/_ref, built from the reference's
               own sources) on this host's cores, on a 2-step sample of the
               headline workload
Progress goes to stderr.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
DOF, N, SCHEME = 4, 384, "RK4"
DT = 10.0  # s; the configured 60 s is unstable for the reference scheme (Courant ~2.2)
U0, H0, LMAX, SEED, RMS = 40.0, 8000.0, 16, 0, 60.0
GRAVITY = 9.80616
# uniform ensemble
# dt = 60 s: the configured 300 s (point Courant ~1.4) blows up within a day
E_MEMBERS, E_DOF, E_N, E_DT, E_SCHEME = 64, 3, 90, 60.0, "RK3_TVD"
E_U0, E_H0, E_LMAX, E_RMS = 20.0, 5960.0, 12, 20.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def words_per_point(scheme):
    # compulsory FP64 prognostic-state traffic per unique point per RK step
    return 51 if scheme == "RK4" else 24


def fp32c_bytes_per_point_step(scheme):
    """Algorithmic HBM bytes per unique point per time step of the FP32C stage
    kernels: FP32 stage input, base hi (stages 2..), base lo (last stage),
    RK4 accumulator, output and (last stage) output residue, plus per stage
    the FP64 phi_ref*sqrtG and one 16-byte FP32 record of the reference
    gradient (2 words) and the Coriolis parameter."""
    if scheme == "RK4":
        # in 4x12, hi 3x12, lo 12, acc 12+24+24+12, out 4x12 + lo 12, static 4 x 24
        return 48 + 36 + 12 + 72 + 48 + 12 + 4 * 24
    # RK3: in 3x12, hi 2x12, lo 12, out 3x12 + lo 12, static 3 x 24
    return 36 + 24 + 12 + 36 + 12 + 3 * 24


def measured_traffic(path):
    """Calibrated HBM bytes per stage launch from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/), or None."""
    import re

    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"Per RK4 step: (\d+(?:\.\d+)?) MB measured", text)
    return float(m.group(1)) * 1e6 / 4 if m else None


def owned(a, dof, n):
    h = dof - 1
    npv = n * h + 1
    nh = npv + 2 * h
    return a.reshape(6, nh, nh)[:, h:h + npv, h:h + npv]


def distance(ref, out, dof, n, mass0):
    """Relative L2 and Linf distance of h, u, v, and the relative mass drift."""
    r = {}
    for k, name in (("phi", "h"), ("u", "u"), ("v", "v")):
        x, y = owned(ref[k], dof, n), owned(out[k], dof, n)
        r[f"{name}_rel_l2"] = float(np.linalg.norm(y - x) / np.linalg.norm(x))
        r[f"{name}_rel_linf"] = float(np.abs(y - x).max() / np.abs(x).max())
    r["mass_drift"] = float((out["mass"] - mass0) / mass0)
    r["mass_drift_fp64"] = float((ref["mass"] - mass0) / mass0)
    return r


def cpu_baseline(mesh_fn, steps=2):
    """Reference Fortran (oracle/_ref) on this host, `steps` steps of the workload;
    the oracle's C restatement (one core) if the reference build cannot run."""
    from mcv_sw_amd import hostlib
    from oracle import ref_runner

    threads = max(1, min(16, os.cpu_count() or 1))
    m = mesh_fn()
    st = hostlib.jet(m, u0=U0, h0=H0, lmax=LMAX, seed=SEED, rms_h=RMS)
    cells = 6 * N * N
    if ref_runner.available(DOF):
        ic = {"phi": st["phi"], "zonal": st["zonal"], "merid": st["merid"], "phi_s": m["phi_s"]}
        log(f"cpu baseline: reference Fortran, {steps} steps, {threads} threads")
        try:
            with tempfile.TemporaryDirectory(prefix="mcvbench_") as wd:
                r = ref_runner.run(DOF, N, DT, steps, scheme=SCHEME, ic=ic, threads=threads,
                                   workdir=wd, dump_mesh=False, dump_states=False, timeout=900)
            return {"value": cells * steps / r["seconds"], "unit": "cell-steps/s",
                    "cores": threads, "kind": "reference",
                    "sample": f"{steps} RK4 steps of the same 6x{N}x{N} MCV4 solve; reference "
                              f"Fortran (amdflang -O3 -fopenmp -fdefault-real-8), "
                              f"OMP_NUM_THREADS={threads}, time-loop wall clock only "
                              f"({r['seconds']:.2f} s)"}
        except Exception as e:
            log(f"reference build failed ({e}); timing the C restatement instead")
    from oracle import oracle

    om = dict(m)
    t = time.perf_counter()
    oracle.run(DOF, N, 4, DT, om, st, 1)
    sec = time.perf_counter() - t
    return {"value": cells / sec, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"1 RK4 step of the same 6x{N}x{N} MCV4 solve with the oracle's C "
                      f"restatement (gcc -O2, one core), {sec:.2f} s"}


def time_history(capi, hostlib, s, m):
    """The output block of MCV_SW.F90:62-70 on the device at full size:
    cell averages alone, the record fields (device + read-back), and one
    whole history record written to a file (history_init first)."""
    r = {}
    t = time.perf_counter()
    s.cell_averages()
    r["cell_averages_ms"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    s.history_fields()
    r["history_fields_ms"] = (time.perf_counter() - t) * 1e3
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "mcv_output.nc")
        capi.history_init(path, DOF, N, DT, hostlib.cells(DOF, N), m["lonP"], m["latP"],
                          m["phi_s"])
        t = time.perf_counter()
        s.history_write(path, 2)
        r["history_record_write_ms"] = (time.perf_counter() - t) * 1e3
        r["history_record_bytes"] = os.path.getsize(path)
    return r


TILE = {}  # precision -> fixed tile shape (--tile64/--tile32), else autotuned


def run_single(capi, m, st, precision, steps, warmup, history=None):
    """Times `steps` steps after warm-up and autotuning; returns the timing
    record and the final state after warmup + 2 + steps + nt steps."""
    cells = 6 * N * N
    with capi.Solver(DOF, N, DT, m, st, scheme=SCHEME, precision=precision) as s:
        mass0 = s.sync(fields=())["mass"]
        s.step(warmup)
        if precision in TILE:
            tile, tile_ms = TILE[precision], None
            s.set_tile(tile)
        else:
            tile, tile_ms = s.autotune(reps=5)
        s.step(2)
        s.sync(fields=(), integrals=False)
        t = time.perf_counter()
        s.step(steps)
        s.sync(fields=(), integrals=False)
        wall = time.perf_counter() - t
        nt = max(10, steps // 5)
        stage_ms, bnd_ms = s.step_timed(nt)
        out = s.sync(fields=("phi", "u", "v"))
        out["mass0"] = mass0
        hist = history(s) if history else None
        rec = {"value": cells * steps / wall, "ms_per_step": wall * 1e3 / steps,
               "stage_kernel_ms_per_step": stage_ms / nt,
               "boundary_kernel_ms_per_step": bnd_ms / nt, "tile_variant": tile,
               "tile_variant_ms_per_stage": tile_ms, "steps_compared": warmup + 2 + steps + nt}
        if hist:
            rec["history"] = hist
    return rec, out


GRAPHS = True  # batch stage sequences replayed as hipGraphs (--no-graphs: direct launches)


def fp64_check(capi, b, slots, specs, steps_of):
    """FP32C members of a batch against the same solve run alone in FP64 for
    the same number of steps: h/u/v distance and mass drift.  `specs[slot]`
    is (dof, n, dt, mesh, state, scheme, mass0)."""
    out = []
    for slot in slots:
        dof, n, dt, m, st, scheme, mass0 = specs[slot]
        steps = steps_of(slot)
        got = b.sync(slot, fields=("phi", "u", "v"))
        with capi.Solver(dof, n, dt, m, st, scheme=scheme) as s:
            s.step(steps)
            ref = s.sync(fields=("phi", "u", "v"))
        e = distance(ref, got, dof, n, mass0)
        out.append({"slot": slot, "dof": dof, "n": n, "steps": int(steps), **e})
    return out


def run_ensemble(capi, hostlib, stages, warmup_stages):
    """The uniform ensemble in one batch, FP32C, and the same 64 members
    advanced in FP64 in a batch of their own; cell-steps/s of all members."""
    import concurrent.futures as cf

    m = hostlib.mesh(E_DOF, E_N)
    cells = E_MEMBERS * 6 * E_N * E_N
    nst = 3  # RK3 stages per step
    total_steps = (warmup_stages + stages + 2 * stages) // nst + 2
    t0 = time.time()
    with cf.ThreadPoolExecutor(16) as pool:
        states = list(pool.map(lambda k: hostlib.jet(dict(m), u0=E_U0, h0=E_H0, lmax=E_LMAX,
                                                     seed=k + 1, rms_h=E_RMS),
                               range(E_MEMBERS)))
    log(f"ensemble: {E_MEMBERS} initial states in {time.time() - t0:.1f} s")

    def timed(precision):
        with capi.Batch(E_MEMBERS) as b:
            b.set_graphs(GRAPHS)
            checked = (0, E_MEMBERS - 1)
            specs = {}
            for k, st in enumerate(states):
                slot = b.add(E_DOF, E_N, E_DT, m, st, total_steps, scheme=E_SCHEME,
                             precision=precision)
                if k in checked:
                    specs[slot] = (E_DOF, E_N, E_DT, m, st, E_SCHEME,
                                   b.sync(slot, fields=())["mass"])
            tile, tile_ms = b.autotune(reps=3)
            b.run(warmup_stages)
            b.poll()
            t = time.perf_counter()
            b.run(stages)
            b.poll()
            wall = time.perf_counter() - t
            stage_ms, bnd_ms = b.run_timed(stages)
            left, done = b.poll()
            errs = (fp64_check(capi, b, sorted(specs), specs, lambda sl: done[sl])
                    if precision == capi.FP32C else None)
        return wall, stage_ms, bnd_ms, tile, tile_ms, int(done[:E_MEMBERS].min()), errs

    wall, stage_ms, bnd_ms, tile, tile_ms, dmin, errs = timed(capi.FP32C)
    wall64, stage64, bnd64, tile64, _, _, _ = timed(capi.FP64)
    steps = stages / nst
    bpps = fp32c_bytes_per_point_step(E_SCHEME)
    pts = E_MEMBERS * 6 * (E_N * (E_DOF - 1) + 1) ** 2
    achieved = bpps * pts * steps / (stage_ms / 1e3) / 1e9
    ach64 = words_per_point(E_SCHEME) * 8 * pts * steps / (stage64 / 1e3) / 1e9
    return {"workload": f"{E_MEMBERS} members MCV3 6x{E_N}x{E_N}, RK3, dt={E_DT} s, FP32C, "
                        f"one fused launch per stage for all members, hipGraph replay",
            "value": cells * steps / wall, "unit": "cell-steps/s",
            "ms_per_step": wall * 1e3 / steps, "stages_timed": stages,
            "stage_kernel_ms_per_step": stage_ms / steps,
            "boundary_kernel_ms_per_step": bnd_ms / steps,
            "stage_kernel_GBs": achieved, "stage_kernel_frac": achieved / PEAK_HBM_GBS,
            "algorithmic_bytes_per_point_step": bpps,
            "tile_variant": tile, "tile_variant_ms_per_stage": tile_ms,
            "members_done_steps": dmin,
            "fp32c_error_vs_fp64_alone": errs,
            "fp64_same_members": {
                "value": cells * steps / wall64, "ms_per_step": wall64 * 1e3 / steps,
                "stage_kernel_ms_per_step": stage64 / steps,
                "boundary_kernel_ms_per_step": bnd64 / steps,
                "stage_kernel_GBs": ach64, "stage_kernel_frac": ach64 / PEAK_HBM_GBS,
                "tile_variant": tile64},
            "fp32c_speedup_same_members": wall64 / wall}


def run_uneven(capi, hostlib, count=128, stages=24, seed=0):
    """Batched uneven sizes: `count` solves in one batch, MCV3/MCV4 50/50, panel
    size uniform in {32, 48, 90, 128, 180, 256}, FP32C, RK4 with dt at Courant
    0.5 on the fastest gravity wave, balanced jets with u0 ~ U(10, 40) m/s and
    perturbation rms ~ U(10, 50) m; autotuned; cell-steps/s over all solves."""
    import concurrent.futures as cf
    from mcv_sw_amd import serve

    rng = np.random.default_rng(seed)
    sizes = (32, 48, 90, 128, 180, 256)
    specs = [dict(dof=4 if rng.random() < 0.5 else 3, n=int(rng.choice(sizes)),
                  u0=float(rng.uniform(10.0, 40.0)), rms=float(rng.uniform(10.0, 50.0)),
                  seed=2000 + k) for k in range(count)]
    t0 = time.time()
    meshes = {}
    for key in sorted({(sp["dof"], sp["n"]) for sp in specs}):
        meshes[key] = hostlib.mesh(*key)

    def prep(sp):
        m = dict(meshes[(sp["dof"], sp["n"])])
        return m, hostlib.jet(m, u0=sp["u0"], h0=5960.0, lmax=8, seed=sp["seed"],
                              rms_h=sp["rms"])

    with cf.ThreadPoolExecutor(16) as pool:
        states = list(pool.map(prep, specs))
    log(f"uneven batch: {count} initial states in {time.time() - t0:.1f} s")
    steps = stages // 4
    cells = 0
    # the smallest solve of each order is checked against FP64 alone
    chk = {}
    for k, sp in enumerate(specs):
        if sp["dof"] not in chk or sp["n"] < specs[chk[sp["dof"]]]["n"]:
            chk[sp["dof"]] = k
    with capi.Batch(count) as b:
        b.set_graphs(GRAPHS)
        cspecs, slot_of = {}, {}
        for k, (sp, (m, st)) in enumerate(zip(specs, states)):
            dt = serve.cfl_dt(sp["dof"], sp["n"], sp["u0"], 5960.0)
            slot = b.add(sp["dof"], sp["n"], dt, m, st, 1000, scheme="RK4", precision=capi.FP32C)
            cells += 6 * sp["n"] ** 2
            if k in chk.values():
                cspecs[slot] = (sp["dof"], sp["n"], dt, m, st, "RK4",
                                b.sync(slot, fields=())["mass"])
        del states
        tile, tile_ms = b.autotune(reps=3)
        b.run(8)
        b.poll()
        t = time.perf_counter()
        b.run(stages)
        b.poll()
        wall = time.perf_counter() - t
        stage_ms, bnd_ms = b.run_timed(stages)
        _, done = b.poll()
        errs = fp64_check(capi, b, sorted(cspecs), cspecs, lambda sl: done[sl])
    mix = {}
    for sp in specs:
        k = f"mcv{sp['dof']}_n{sp['n']}"
        mix[k] = mix.get(k, 0) + 1
    return {"workload": f"{count} solves in one batch: MCV3/MCV4 50/50, n uniform in "
                        f"{{32,48,90,128,180,256}}, FP32C, RK4, CFL dt (Courant 0.5), one fused "
                        f"launch per stage, autotuned tile shapes",
            "value": cells * steps / wall, "unit": "cell-steps/s", "cells": cells,
            "ms_per_step": wall * 1e3 / steps, "stages_timed": stages,
            "stage_kernel_ms_per_step": stage_ms / steps,
            "boundary_kernel_ms_per_step": bnd_ms / steps,
            "tile_variant": tile, "tile_variant_ms_per_stage": tile_ms, "mix": mix,
            "fp32c_error_vs_fp64_alone": errs}


def run_serve(capi, seconds, rate, seed=0, check=6, warm_frac=0.25):
    """The serving workload at a steady offered load: Poisson arrivals at
    `rate` req/s for `seconds`, one seeded solve per request (order, size,
    length and precision drawn as configured), continuous batching with
    admission by HBM footprint.  Reports offered against served cell-steps/s
    and the latency percentiles of the requests that arrived after the first
    `warm_frac` of the window (the queue's steady part); the configured
    20 req/s mix is characterised by its offered load only (one GPU serves a
    small fraction of it).  After the run, up to `check` of the smallest FP32C
    requests are re-run alone in FP64 and their h/u/v distance reported, and
    every request that went non-finite is re-run alone in FP64 to tell the
    scheme's blow-up from the form's."""
    from mcv_sw_amd import hostlib, serve

    reqs = serve.draw_requests(rate, seconds, seed=seed)
    srv = serve.Server(capacity=256, hbm_budget=200 << 30, stages_per_poll=12, workers=16)
    rep, res = srv.run(reqs, keep=("phi", "u", "v"))
    rows = rep["per_request"]
    by_rid = {r.rid: r for r in reqs}
    offered = sum(r.cells * r.nsteps for r in reqs) / seconds
    done_cs = sum(by_rid[x["rid"]].cells * by_rid[x["rid"]].nsteps for x in rows)
    steady = [x for x in rows if x["arrival"] >= warm_frac * seconds]
    lat = np.array([x["latency"] for x in steady]) if steady else np.array([np.nan])
    wait = np.array([x["admit"] - x["arrival"] for x in rows])
    cfg_reqs = serve.draw_requests(20.0, 600.0, seed=seed)
    cfg_offered = sum(r.cells * r.nsteps for r in cfg_reqs) / 600.0
    errs = []
    sample = sorted((r for r in reqs if r.precision == capi.FP32C and r.rid in res),
                    key=lambda r: r.cells)[:check]
    for r in sample:
        m = hostlib.mesh(r.dof, r.n)
        st = hostlib.jet(m, u0=r.u0, h0=r.h0, lmax=r.lmax, seed=r.seed, rms_h=r.rms_h)
        with capi.Solver(r.dof, r.n, r.dt, m, st, scheme=r.scheme) as s:
            s.step(r.nsteps)
            ref = s.sync(fields=("phi", "u", "v"))
        e = distance(ref, res[r.rid], r.dof, r.n, ref["mass"])
        errs.append({"rid": r.rid, "dof": r.dof, "n": r.n, "hours": round(r.hours, 2),
                     "h_rel_l2": e["h_rel_l2"], "u_rel_l2": e["u_rel_l2"],
                     "v_rel_l2": e["v_rel_l2"]})
    twins = []
    for x in rows:
        if x["status"] == "ok":
            continue
        r = by_rid[x["rid"]]
        m = hostlib.mesh(r.dof, r.n)
        st = hostlib.jet(m, u0=r.u0, h0=r.h0, lmax=r.lmax, seed=r.seed, rms_h=r.rms_h)
        with capi.Solver(r.dof, r.n, r.dt, m, st, scheme=r.scheme) as s:
            s.step(r.nsteps)
            try:
                s.sync(fields=())
                fin = True
            except capi.MCVError:
                fin = False
        twins.append({"rid": r.rid, "dof": r.dof, "n": r.n, "hours": round(r.hours, 2),
                      "precision": r.precision, "fp64_alone_finite": fin})
    sizes = {}
    for r in reqs:
        sizes[f"mcv{r.dof}_n{r.n}"] = sizes.get(f"mcv{r.dof}_n{r.n}", 0) + 1
    out = {k: v for k, v in rep.items() if k != "per_request"}
    out.update({
        "workload": f"Poisson {rate} req/s for {seconds} s of arrivals; MCV3/MCV4 p=0.5; n from "
                    f"{{32,64,128,256,384}} w 0.3/0.3/0.2/0.15/0.05; 6-48 simulated hours; "
                    f"FP64/FP32C p=0.5; RK4, CFL dt",
        "offered_req_per_s": rate, "window_s": seconds,
        "offered_cell_steps_per_s": offered,
        "served_cell_steps_per_s": done_cs / rep["makespan_s"] if rep["makespan_s"] else 0.0,
        "steady_sample": len(steady),
        "latency_p50_steady_s": float(np.percentile(lat, 50)),
        "latency_p99_steady_s": float(np.percentile(lat, 99)),
        "admission_wait_mean_s": float(wait.mean()) if len(wait) else None,
        "admission_wait_max_s": float(wait.max()) if len(wait) else None,
        "configured_20_req_per_s_offered_cell_steps_per_s": cfg_offered,
        "requests_by_size": sizes,
        "fp32c_error_vs_fp64_alone": errs,
        "non_finite_requests": [
            {k: x[k] for k in ("rid", "dof", "n", "precision", "hours", "u0", "rms_h")}
            for x in rows if x["status"] != "ok"],
        "non_finite_fp64_twins": twins,
        "note": "latency = arrival to read-back of the finished solve, including host "
                "preparation of the initial state; steady = requests arriving after the "
                f"first {warm_frac:.0%} of the window"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32c", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--ensemble-stages", type=int, default=180,
                    help="timed RK stages of the ensemble (a multiple of 3)")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="seconds of Poisson arrivals for the serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=0.5,
                    help="offered requests per second of the serving run (the configured "
                         "20 req/s mix offers about 18x what one GPU serves; its offered "
                         "load is reported beside)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the plain copy kernel on this many MiB (counter calibration)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="batches launch every kernel directly instead of replaying hipGraphs "
                         "(for kernel-trace profiling)")
    ap.add_argument("--tile64", type=int, default=None,
                    help="fixed stage-kernel tile shape of the FP64 solve (default: autotuned)")
    ap.add_argument("--tile32", type=int, default=None,
                    help="fixed stage-kernel tile shape of the FP32C solve (default: autotuned)")
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = not args.no_graphs
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the MCV step here is a "
              "single-GPU build.", file=sys.stderr)
        return 2

    from mcv_sw_amd import capi, hostlib

    if args.tile64 is not None:
        TILE[capi.FP64] = args.tile64
    if args.tile32 is not None:
        TILE[capi.FP32C] = args.tile32

    if args.copy_probe_mib:
        nb = args.copy_probe_mib << 20
        ms = capi.copy_probe(nb, reps=args.steps)
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "ms_per_copy": ms,
                          "GBs": 2 * nb / (ms / 1e3) / 1e9, "reps": args.steps}), flush=True)
        return 0

    t0 = time.time()
    m = hostlib.mesh(DOF, N)
    st = hostlib.jet(m, u0=U0, h0=H0, lmax=LMAX, seed=SEED, rms_h=RMS)
    log(f"mesh + initial state: {time.time() - t0:.1f} s")
    cells = 6 * N * N
    rec64, out64 = run_single(capi, m, st, capi.FP64, args.steps, args.warmup,
                              history=lambda s: time_history(capi, hostlib, s, m))
    log(f"FP64: {rec64['ms_per_step']:.3f} ms/step")
    nt_launches = 4 * max(10, args.steps // 5)
    bpcs = words_per_point(SCHEME) * 8 * (DOF - 1) ** 2  # bytes per cell-step
    bytes_per_launch = bpcs * cells / 4
    avg_launch_s = rec64["stage_kernel_ms_per_step"] / 4 / 1e3
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = measured_traffic(os.path.join(ROOT, "profiles", "stage_kernel_mcv4_n384_fp64.md"))
    res = {
        "metric": "SW cell-steps/s (plain FP64 vs batched FP32) at stated h err; serve p50/p99",
        "value": rec64["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rec64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: balanced jet u0=40 m/s h0=8000 m + seeded (seed 0) degree-16 "
                "spherical-harmonic perturbation, 60 m rms",
        "config": {"workload": f"single large solve: MCV4 6x{N}x{N}, RK4, FP64 drop-in "
                               f"(bit-identical to the reference), dt={DT} s",
                   "dof": DOF, "n": N, "cells": cells, "scheme": SCHEME, "dt_s": DT,
                   "h_rel_err_vs_fp64": 0.0},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic,
                     # the measured HBM bytes per launch over this run's mean launch time
                     "traffic_GBs": traffic / avg_launch_s / 1e9 if traffic else None,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "kernel": "stage_kernel (4 launches per RK4 step)",
                     "bytes_per_cell_step": bpcs,
                     "launches_timed": nt_launches,
                     "stage_kernel_ms_per_step": rec64["stage_kernel_ms_per_step"],
                     "tile_variant": rec64["tile_variant"],
                     "tile_variant_ms_per_stage": rec64["tile_variant_ms_per_stage"],
                     "boundary_kernel_ms_per_step": rec64["boundary_kernel_ms_per_step"]},
        "mass_drift_fp64": float((out64["mass"] - out64["mass0"]) / out64["mass0"]),
        "history": rec64.pop("history"),
    }
    if not args.no_fp32c:
        rec32, out32 = run_single(capi, m, st, capi.FP32C, args.steps, args.warmup)
        log(f"FP32C: {rec32['ms_per_step']:.3f} ms/step")
        err = distance(out64, out32, DOF, N, out32["mass0"])
        bpp = fp32c_bytes_per_point_step(SCHEME)
        pts = 6 * (N * (DOF - 1) + 1) ** 2
        ach32 = bpp * pts / (rec32["stage_kernel_ms_per_step"] / 1e3) / 1e9
        rec32.update({"speedup_vs_fp64": rec32["value"] / rec64["value"],
                      "error_vs_fp64": err,
                      "stage_kernel_GBs": ach32, "stage_kernel_frac": ach32 / PEAK_HBM_GBS,
                      "algorithmic_bytes_per_point_step": bpp})
        res["fp32c"] = rec32
        del out32
    del out64
    if not args.no_batch:
        res["batched"] = run_ensemble(capi, hostlib, stages=max(3, args.ensemble_stages // 3 * 3),
                                      warmup_stages=3 * 4)
        log(f"ensemble: {res['batched']['value']:.3e} cell-steps/s")
    if not args.no_batch:
        res["batched_uneven"] = run_uneven(capi, hostlib)
        log(f"uneven batch: {res['batched_uneven']['value']:.3e} cell-steps/s")
    if args.serve_seconds > 0:
        res["serve"] = run_serve(capi, args.serve_seconds, args.serve_rate)
        sv = res["serve"]
        log(f"serve: {sv['completed']} solves, offered {sv['offered_cell_steps_per_s']:.3e} "
            f"served {sv['served_cell_steps_per_s']:.3e} cell-steps/s, steady p50 "
            f"{sv['latency_p50_steady_s']:.2f} s p99 {sv['latency_p99_steady_s']:.2f} s")
    cp_ms = capi.copy_probe(1 << 30, reps=10)
    res["roofline"]["copy_kernel_GBs"] = 2 * (1 << 30) / (cp_ms / 1e3) / 1e9
    # the streaming-copy yardstick: the best of several loads in flight per
    # thread and non-temporal accesses, over 1 GiB and 4 GiB
    best = 0.0
    for gib in (1, 4):
        sp_ms, sp_v, sp_pcu = capi.stream_probe(gib << 30, reps=10)
        gbs = 2 * (gib << 30) / (sp_ms / 1e3) / 1e9
        res["roofline"][f"stream_copy_{gib}GiB_GBs"] = gbs
        best = max(best, gbs)
    res["roofline"]["frac_of_stream_copy"] = achieved / best
    if res["roofline"].get("traffic_GBs"):
        res["roofline"]["traffic_frac_of_stream_copy_4GiB"] = (
            res["roofline"]["traffic_GBs"] / res["roofline"]["stream_copy_4GiB_GBs"])
    if not args.no_cpu_baseline:
        try:
            res["cpu_baseline"] = cpu_baseline(lambda: hostlib.mesh(DOF, N))
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e}")
            res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
