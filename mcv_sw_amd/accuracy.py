"""Error report of the FP32-compensated form against the FP64 path.

Runs one initial state twice on the GPU, FP64 (the reference's arithmetic bit
for bit, so the FP64 trajectory is the oracle's) and FP32C, side by side, and
reports at the requested simulated days the relative L2 and Linf distance of
h, u and v over the owned points, the area-weighted L1/L2/LInf of the cell
averages (doc/orderCheck.m:60-82, device calc_VIA) of h, u and v, and both
runs' mass drift.

    python -m mcv_sw_amd.accuracy --dof 4 --n 384 --dt 10 --days 1,5,10

Default workload: MCV4 on 6 x 384 x 384, RK4, dt = 10 s, balanced jet
u0 = 40 m/s, h0 = 8000 m, seeded (seed 0) degree-16 perturbation at 60 m rms.
One JSON line per checkpoint on stdout; progress on stderr.

Yardstick (--yardstick): a third run, FP64 from the initial state rounded to
FP32 (u, v and phi; the same handoff otherwise).  Its distance from the FP64
trajectory is how far the flow itself carries differences of FP32 size; an
FP32C distance of the same order means the error is the flow's sensitivity,
not an accumulation of the form's rounding.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import capi, diagnostics, hostlib


def owned(a, dof, n):
    h = dof - 1
    npv = n * h + 1
    nh = npv + 2 * h
    return a.reshape(6, nh, nh)[:, h:h + npv, h:h + npv]


def compare(ref, out, dof, n):
    r = {}
    for k, name in (("phi", "h"), ("u", "u"), ("v", "v")):
        x, y = owned(ref[k], dof, n), owned(out[k], dof, n)
        r[f"{name}_rel_l2"] = float(np.linalg.norm(y - x) / np.linalg.norm(x))
        r[f"{name}_rel_linf"] = float(np.abs(y - x).max() / np.abs(x).max())
    return r


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--dof", type=int, default=4)
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--dt", type=float, default=10.0)
    ap.add_argument("--scheme", default="RK4")
    ap.add_argument("--days", default="1,5,10")
    ap.add_argument("--u0", type=float, default=40.0)
    ap.add_argument("--h0", type=float, default=8000.0)
    ap.add_argument("--lmax", type=int, default=16)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rms", type=float, default=60.0)
    ap.add_argument("--chunk-hours", type=float, default=6.0)
    ap.add_argument("--yardstick", action="store_true")
    ap.add_argument("--lib", default=None, help="C-ABI library (default: the in-tree HIP build)")
    args = ap.parse_args(argv)
    lib = capi.load(args.lib)

    def log(msg):
        print(f"[accuracy] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    m = hostlib.mesh(args.dof, args.n)
    area = hostlib.cells(args.dof, args.n)["areaCell"]
    st = hostlib.jet(m, u0=args.u0, h0=args.h0, lmax=args.lmax, seed=args.seed, rms_h=args.rms)
    log(f"setup {time.time() - t0:.1f} s")
    days = [float(d) for d in args.days.split(",")]
    targets = [int(round(d * 86400.0 / args.dt)) for d in days]
    chunk = max(1, int(round(args.chunk_hours * 3600.0 / args.dt)))
    solvers = [capi.Solver(args.dof, args.n, args.dt, m, st, scheme=args.scheme, lib=lib),
               capi.Solver(args.dof, args.n, args.dt, m, st, scheme=args.scheme,
                           precision=capi.FP32C, lib=lib)]
    if args.yardstick:
        r32 = {k: st[k].astype(np.float32).astype(np.float64) for k in ("phi", "zonal", "merid")}
        sy = hostlib.handoff(m, r32["phi"], r32["zonal"], r32["merid"])
        solvers.append(capi.Solver(args.dof, args.n, args.dt, m, sy, scheme=args.scheme,
                                   lib=lib))
        del sy
    del st
    try:
        for s in solvers:
            s.autotune(reps=3)
        mass0 = [s.sync(fields=())["mass"] for s in solvers]
        done = 0
        for day, target in zip(days, targets):
            while done < target:
                k = min(chunk, target - done)
                for s in solvers:
                    s.step(k)
                for s in solvers:
                    s.sync(fields=(), integrals=False)
                done += k
                log(f"{done} / {targets[-1]} steps ({done * args.dt / 86400.0:.2f} days), "
                    f"{time.time() - t0:.0f} s")
            out = [s.sync(fields=("phi", "u", "v")) for s in solvers]
            rec = {"day": day, "steps": done, "dof": args.dof, "n": args.n, "dt_s": args.dt,
                   "scheme": args.scheme}
            rec.update(compare(out[0], out[1], args.dof, args.n))
            cav = [s.cell_averages() for s in solvers[:2]]  # (uC, vC, phiC) each
            rec["ordercheck_fp32c"] = {
                name: dict(zip(("L1", "L2", "LInf"),
                               diagnostics.norms(cav[1][k], cav[0][k], area)))
                for name, k in (("h", 2), ("u", 0), ("v", 1))}
            rec["mass_drift_fp32c"] = (out[1]["mass"] - mass0[1]) / mass0[1]
            rec["mass_drift_fp64"] = (out[0]["mass"] - mass0[0]) / mass0[0]
            if args.yardstick:
                y = compare(out[0], out[2], args.dof, args.n)
                rec["yardstick_fp64_from_fp32_rounded_ic"] = {
                    k: y[k] for k in ("h_rel_l2", "u_rel_l2", "v_rel_l2")}
            print(json.dumps(rec), flush=True)
    finally:
        for s in solvers:
            s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
