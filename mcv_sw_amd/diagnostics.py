"""Error norms and convergence orders of doc/orderCheck.m, on the history
file the library writes (the reference's mcv_output.nc) or on cell averages
taken from the device (Solver.cell_averages).

orderCheck.m:60-82 compares a cell-averaged field (phiC) at two time slots
over the interior cells (its..ite, jts..jte: cells 1..n) with areaCell
weights:
    L1   = sum(|m - r| A) / sum(|r| A)
    L2   = sqrt(sum((m - r)^2 A) / sum(r^2 A))
    LInf = max(|m - r| A) / max(|r| A)
and, over runs at several resolutions, the order between consecutive runs
log(L(k) / L(k-1)) / log(dx(k) / dx(k-1))  (orderCheck.m:54-58).

    python -m mcv_sw_amd.diagnostics run1/mcv_output.nc [run2/mcv_output.nc ...]
        [--field phiC] [--start 1] [--end -1]
prints one JSON line per file and one with the orders.
"""
import argparse
import json
import math
import sys

import numpy as np


def interior(a):
    """Cells 1..n of a (..., 6, n+2, n+2) cell field (one halo cell)."""
    return a[..., 1:-1, 1:-1]


def norms(model, ref, area):
    """(L1, L2, LInf) of orderCheck.m:60-82 on interior cells; arrays
    (6, n+2, n+2) as the history file and Solver.cell_averages hold them."""
    m, r, a = interior(np.asarray(model)), interior(np.asarray(ref)), interior(np.asarray(area))
    d = m - r
    l1 = np.sum(np.abs(d) * a) / np.sum(np.abs(r) * a)
    l2 = math.sqrt(np.sum(d * d * a) / np.sum(r * r * a))
    linf = np.max(np.abs(d) * a) / np.max(np.abs(r) * a)
    return float(l1), float(l2), float(linf)


def orders(errors, dxs):
    """Convergence orders between consecutive resolutions (orderCheck.m:54-58)."""
    return [math.log(errors[k] / errors[k - 1]) / math.log(dxs[k] / dxs[k - 1])
            for k in range(1, len(errors))]


def file_norms(path, field="phiC", start=1, end=-1):
    """Norms of `field` at time slot `end` against slot `start` (1-based;
    negative counts from the last record) of a history file."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as f:
        data = f.variables[field].data
        area = f.variables["areaCell"].data.copy()
        nrec = data.shape[0]
        s = start - 1 if start > 0 else nrec + start
        e = end - 1 if end > 0 else nrec + end
        model, ref = data[e].copy(), data[s].copy()
        dx = float(np.asarray(f._attributes["dx"]))
    return {"file": path, "field": field, "start": s + 1, "end": e + 1, "dx_deg": dx,
            **dict(zip(("L1", "L2", "LInf"), norms(model, ref, area)))}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("files", nargs="+")
    ap.add_argument("--field", default="phiC")
    ap.add_argument("--start", type=int, default=1)
    ap.add_argument("--end", type=int, default=-1)
    args = ap.parse_args(argv)
    recs = [file_norms(p, args.field, args.start, args.end) for p in args.files]
    for r in recs:
        print(json.dumps(r))
    if len(recs) > 1:
        dxs = [r["dx_deg"] for r in recs]
        print(json.dumps({k + "_order": orders([r[k] for r in recs], dxs)
                          for k in ("L1", "L2", "LInf")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
