"""Continuous-batching request loop over the batch C-ABI (mcv_batch_*).

Requests (one seeded cubed-sphere solve each) arrive over time.  The loop:
  * prepares each arrived request's mesh (cached per order and size) and
    initial state on a host worker pool, overlapping GPU work;
  * admits prepared requests while batch slots and the HBM budget allow
    (each solve's device footprint from mcv_hip_footprint), so solves join
    while others are mid-run and every solve's state stays on the device;
  * advances the whole batch in chunks of RK stages (hipGraph replays of one
    fused stage launch + one fused boundary launch per stage), each solve at
    its own stage of its own scheme;
  * retires finished solves (read-back, FP64 mass/energy) and frees their
    slot and memory for the next admission; a solve whose state is no longer
    finite is retired as failed ("non-finite") and reported.
Latency is arrival to completed read-back.  The reference has no serving
mode (one solve per process, MCV_SW.F90:48-77); the per-solve step is the
reference's, bit for bit in FP64.
"""
import collections
import concurrent.futures as cf
import dataclasses
import math
import time

import numpy as np

from . import capi, hostlib

GRAVITY = hostlib.GRAVITY
RADIUS = hostlib.RADIUS


@dataclasses.dataclass
class Request:
    rid: int
    t_arrival: float   # s after the start of the run
    dof: int
    n: int
    hours: float       # simulated length
    precision: int     # capi.FP64 or capi.FP32C
    seed: int
    u0: float = 20.0
    h0: float = 5960.0
    lmax: int = 8
    rms_h: float = 30.0
    scheme: str = "RK4"

    @property
    def dt(self):
        return cfl_dt(self.dof, self.n, self.u0, self.h0)

    @property
    def nsteps(self):
        return max(1, int(math.ceil(self.hours * 3600.0 / self.dt)))

    @property
    def cells(self):
        return 6 * self.n * self.n


def cfl_dt(dof, n, u0, h0, courant=0.5):
    """Time step at the given Courant number on the fastest gravity wave,
    sqrt(g h0) + u0, over the smallest point spacing of the gnomonic grid
    (the panel corners, where a point spacing of dx/H shrinks by 1/sqrt(2)...
    1/2 in the two directions; the factor 0.5 is used)."""
    h = dof - 1
    spacing = RADIUS * (math.pi / 2 / n) / h * 0.5
    return courant * spacing / (math.sqrt(GRAVITY * h0) + u0)


def draw_requests(rate, duration, seed=0, hours=(6.0, 48.0), sizes=(32, 64, 128, 256, 384),
                  weights=(0.3, 0.3, 0.2, 0.15, 0.05), p_mcv4=0.5, p_fp32c=0.5):
    """Poisson arrivals of the serving workload: order MCV3/MCV4, panel size
    from `sizes` with `weights`, length uniform in `hours`, precision FP64 or
    FP32C; balanced jets with seeded perturbations."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        k = len(out)
        out.append(Request(
            rid=k, t_arrival=t, dof=4 if rng.random() < p_mcv4 else 3,
            n=int(rng.choice(sizes, p=np.asarray(weights) / np.sum(weights))),
            hours=float(rng.uniform(*hours)),
            precision=capi.FP32C if rng.random() < p_fp32c else capi.FP64,
            seed=1000 + k, u0=float(rng.uniform(10.0, 40.0)),
            rms_h=float(rng.uniform(10.0, 50.0))))


class Server:
    """One GPU, one batch, a host worker pool for request preparation."""

    def __init__(self, capacity=64, hbm_budget=200 << 30, stages_per_poll=12, workers=8,
                 device=0, lib=None):
        self.capacity = capacity
        self.hbm_budget = hbm_budget
        self.stages_per_poll = stages_per_poll
        self.workers = workers
        self.device = device
        self.lib = lib
        self._meshes = {}

    def _mesh(self, dof, n):
        key = (dof, n)
        if key not in self._meshes:
            self._meshes[key] = hostlib.mesh(dof, n)
        return self._meshes[key]

    def _prepare(self, r):
        m = dict(self._mesh(r.dof, r.n))  # jet() sets phi_s on its own copy
        st = hostlib.jet(m, u0=r.u0, h0=r.h0, lmax=r.lmax, seed=r.seed, rms_h=r.rms_h)
        return m, st

    def run(self, requests, keep=("phi", "u", "v"), log=None):
        """Serve `requests` (sorted or not by arrival); returns a report dict and
        {rid: result} with the kept fields, mass and energy of every request."""
        reqs = sorted(requests, key=lambda r: r.t_arrival)
        pending = collections.deque(reqs)
        preparing = {}          # future -> request
        ready = collections.deque()
        active = {}             # slot -> (request, bytes, t_admit)
        results, lat, rows = {}, [], []
        used = 0
        footprint = {}
        # meshes of every order and size in the run are built up front (not
        # per request); initial states are per request
        for key in sorted({(r.dof, r.n) for r in reqs}):
            self._mesh(*key)
        pool = cf.ThreadPoolExecutor(max_workers=self.workers)
        b = capi.Batch(self.capacity, device=self.device, lib=self.lib)
        t0 = time.perf_counter()
        busy = 0.0
        try:
            while pending or preparing or ready or active:
                now = time.perf_counter() - t0
                while pending and pending[0].t_arrival <= now:
                    r = pending.popleft()
                    preparing[pool.submit(self._prepare, r)] = r
                for f in [f for f in preparing if f.done()]:
                    ready.append((preparing.pop(f), f.result()))
                while ready and len(active) < self.capacity:
                    r, (m, st) = ready[0]
                    key = (r.dof, r.n, r.precision)
                    if key not in footprint:
                        footprint[key] = capi.footprint(r.dof, r.n, precision=r.precision,
                                                        lib=self.lib)
                    need = footprint[key]
                    if active and used + need > self.hbm_budget:
                        break  # wait for memory (a lone request is always admitted)
                    ready.popleft()
                    slot = b.add(r.dof, r.n, r.dt, m, st, r.nsteps, scheme=r.scheme,
                                 precision=r.precision)
                    active[slot] = (r, need, time.perf_counter() - t0)
                    used += need
                if not active:
                    time.sleep(0.0005)
                    continue
                tb = time.perf_counter()
                b.run(self.stages_per_poll)
                left, _ = b.poll()
                busy += time.perf_counter() - tb
                for slot in [s for s in active if left[s] == 0]:
                    r, need, t_admit = active.pop(slot)
                    status = "ok"
                    try:
                        out = b.sync(slot, fields=keep)
                    except capi.MCVError as e:
                        if e.code != capi.MCV_ENONFINITE:
                            raise
                        status, out = "non-finite", None  # the solve blew up
                    b.remove(slot)
                    used -= need
                    t_done = time.perf_counter() - t0
                    lat.append(t_done - r.t_arrival)
                    if out is not None:
                        results[r.rid] = out
                    rows.append({"rid": r.rid, "dof": r.dof, "n": r.n, "precision": r.precision,
                                 "hours": r.hours, "u0": r.u0, "rms_h": r.rms_h,
                                 "steps": r.nsteps, "arrival": r.t_arrival, "admit": t_admit,
                                 "done": t_done, "latency": t_done - r.t_arrival,
                                 "status": status})
                    if log:
                        log(f"request {r.rid} {status}: MCV{r.dof} n={r.n} "
                            f"{'FP32C' if r.precision else 'FP64'} {r.nsteps} steps, "
                            f"latency {t_done - r.t_arrival:.2f} s")
        finally:
            pool.shutdown(wait=True)
            b.close()
        span = time.perf_counter() - t0
        lat = np.asarray(lat)
        cell_steps = float(sum(r.cells * r.nsteps for r in reqs))
        report = {
            "requests": len(reqs),
            "completed": int(len(lat)),
            "non_finite": [x["rid"] for x in rows if x["status"] != "ok"],
            "makespan_s": span,
            "solves_per_s": len(lat) / span if span > 0 else 0.0,
            "cell_steps_per_s": cell_steps / span if span > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "latency_max_s": float(lat.max()) if len(lat) else None,
            "gpu_loop_busy_frac": busy / span if span > 0 else 0.0,
            "per_request": rows,
        }
        return report, results
