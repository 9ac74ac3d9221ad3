"""ctypes binding of the MCV step's C-ABI (include/mcv_hip.h).

This is the Python counterpart of the reference's time-integration seam
(`call RK4(stat(new), stat(old))`, MCV_SW.F90:58-59): `Solver(cfg, mesh,
state)` uploads the mesh and initial state (mcv_hip_init), `step(n)` advances
n time steps on the GPU (mcv_hip_step), `sync()` reads the state and the FP64
mass/energy integrals back (mcv_hip_sync).

Arrays are numpy float64 in the reference layout: Fortran (i, j, panel) with
the halo, i.e. C-contiguous shape (6, nh, nh) indexed [panel, j, i]; 2x2
matrices have shape (6, nh, nh, 2, 2) indexed [panel, j, i, col, row].

The library is loaded from the package directory (built in-tree by
`python -c "import __graft_entry__; __graft_entry__.build()"`); there is no
fallback: a missing library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcv_hip.so")

ABI_VERSION = 1
MCV_OK, MCV_EINVAL, MCV_ENOMEM, MCV_EDEVICE, MCV_ENONFINITE, MCV_ESTATE = 0, -1, -2, -3, -4, -5
RK3_TVD, RK4 = 3, 4
FP64, FP32C = 0, 1
SCHEMES = {"RK3_TVD": RK3_TVD, "RK4": RK4}
TILE_VARIANTS = 4

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


class Config(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32), ("dof", ctypes.c_int32), ("n", ctypes.c_int32),
                ("scheme", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("device", ctypes.c_int32), ("dt", ctypes.c_double), ("dx", ctypes.c_double)]


class Mesh(ctypes.Structure):
    _fields_ = [(k, _D) for k in ("sqrtG", "matrixG", "matrixIG", "matrixA", "matrixIA", "f",
                                  "phi_s", "ghost_coef")] + [("ghost_iref", _I)]


class State(ctypes.Structure):
    _fields_ = [(k, _D) for k in ("phi", "phiG", "u", "v", "contraU", "contraV")]


STATE_FIELDS = ("phi", "phiG", "u", "v", "contraU", "contraV")

# one history record (history_write_stat, output_mod.F90:200-212)
HISTORY_POINT_FIELDS = ("u", "v", "contraU", "contraV", "phi", "phi_t", "zonal", "meridional")
HISTORY_CELL_FIELDS = ("uC", "vC", "phiC")


class HistoryInfo(ctypes.Structure):
    _fields_ = [("dof", ctypes.c_int32), ("n", ctypes.c_int32), ("xhalo", ctypes.c_int32),
                ("yhalo", ctypes.c_int32), ("case_num", ctypes.c_int32), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("dt", ctypes.c_double)]


class HistoryFields(ctypes.Structure):
    _fields_ = [(k, _D) for k in HISTORY_POINT_FIELDS + HISTORY_CELL_FIELDS]


class MCVError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"MCV error {code}: {msg}")
        self.code = code


_libs = {}


def load(path=None):
    """Load the C-ABI library (default: the in-tree HIP build); an already
    loaded library is returned as is."""
    if isinstance(path, ctypes.CDLL):
        return path
    path = os.path.abspath(path or LIB_PATH)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
    L = ctypes.CDLL(path)
    H = ctypes.c_void_p
    L.mcv_hip_init.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Mesh),
                               ctypes.POINTER(State), ctypes.POINTER(H)]
    L.mcv_hip_step.argtypes = [H, ctypes.c_int32]
    L.mcv_hip_sync.argtypes = [H, ctypes.POINTER(State), _D, _D]
    L.mcv_hip_sphere_winds.argtypes = [H, _D, _D]
    L.mcv_hip_tendency.argtypes = [H, ctypes.POINTER(State), _D, _D, _D]
    L.mcv_hip_tendency_noalloc.argtypes = [H, ctypes.POINTER(State), _D, _D, _D]
    L.mcv_hip_steps_done.argtypes = [H, ctypes.POINTER(ctypes.c_int64)]
    L.mcv_hip_step_timed.argtypes = [H, ctypes.c_int32, _D, _D]
    L.mcv_hip_copy_probe.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _D]
    L.mcv_hip_copy_probe.restype = ctypes.c_int
    L.mcv_hip_stream_probe.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _D,
                                       ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32)]
    L.mcv_hip_stream_probe.restype = ctypes.c_int
    L.mcv_hip_autotune.argtypes = [H, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _D]
    L.mcv_hip_autotune.restype = ctypes.c_int
    L.mcv_hip_set_tile.argtypes = [H, ctypes.c_int32]
    L.mcv_hip_set_tile.restype = ctypes.c_int
    L.mcv_hip_free.argtypes = [H]
    L.mcv_hip_last_error.argtypes = []
    L.mcv_hip_last_error.restype = ctypes.c_char_p
    L.mcv_hip_footprint.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_int64)]
    L.mcv_hip_device_bytes.argtypes = [H, ctypes.POINTER(ctypes.c_int64)]
    B = ctypes.c_void_p
    _L64 = ctypes.POINTER(ctypes.c_int64)
    L.mcv_batch_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(B)]
    L.mcv_batch_add.argtypes = [B, ctypes.POINTER(Config), ctypes.POINTER(Mesh),
                                ctypes.POINTER(State), ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_int32)]
    L.mcv_batch_remove.argtypes = [B, ctypes.c_int32]
    L.mcv_batch_run.argtypes = [B, ctypes.c_int32]
    L.mcv_batch_poll.argtypes = [B, _L64, _L64]
    L.mcv_batch_sync.argtypes = [B, ctypes.c_int32, ctypes.POINTER(State), _D, _D]
    L.mcv_batch_handle.argtypes = [B, ctypes.c_int32, ctypes.POINTER(H)]
    L.mcv_batch_autotune.argtypes = [B, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _D]
    L.mcv_batch_set_tile.argtypes = [B, ctypes.c_int32]
    L.mcv_batch_set_graphs.argtypes = [B, ctypes.c_int32]
    L.mcv_batch_run_timed.argtypes = [B, ctypes.c_int32, _D, _D]
    L.mcv_batch_free.argtypes = [B]
    L.mcv_history_init.argtypes = [ctypes.c_char_p, ctypes.POINTER(HistoryInfo)] + [_D] * 6
    L.mcv_history_write.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.POINTER(HistoryFields)]
    L.mcv_hip_cell_averages.argtypes = [H, _D, _D, _D]
    L.mcv_hip_history_fields.argtypes = [H, ctypes.POINTER(HistoryFields)]
    L.mcv_hip_history_write.argtypes = [H, ctypes.c_char_p, ctypes.c_int32]
    for f in ("mcv_hip_init", "mcv_hip_step", "mcv_hip_sync", "mcv_hip_sphere_winds",
              "mcv_hip_tendency", "mcv_hip_tendency_noalloc", "mcv_hip_steps_done", "mcv_hip_step_timed", "mcv_hip_free",
              "mcv_hip_footprint", "mcv_hip_device_bytes", "mcv_batch_create", "mcv_batch_add",
              "mcv_batch_remove", "mcv_batch_run", "mcv_batch_poll", "mcv_batch_sync",
              "mcv_batch_handle", "mcv_batch_autotune", "mcv_batch_set_tile",
              "mcv_batch_set_graphs", "mcv_batch_run_timed", "mcv_batch_free",
              "mcv_history_init", "mcv_history_write", "mcv_hip_cell_averages",
              "mcv_hip_history_fields", "mcv_hip_history_write"):
        getattr(L, f).restype = ctypes.c_int
    _libs[path] = L
    return L


def copy_probe(nbytes, reps=20, device=0, lib=None):
    """Milliseconds per plain device copy of nbytes (read nbytes + write nbytes)."""
    L = load(lib)
    ms = ctypes.c_double()
    rc = L.mcv_hip_copy_probe(device, int(nbytes), int(reps), ctypes.byref(ms))
    if rc != MCV_OK:
        raise MCVError(rc, "mcv_hip_copy_probe failed")
    return ms.value


def stream_probe(nbytes, reps=10, device=0, lib=None):
    """Fastest streaming copy of nbytes (several loads in flight per thread,
    non-temporal accesses): (ms per copy, loads-in-flight variant, workgroups
    per CU)."""
    L = load(lib)
    ms = ctypes.c_double()
    v, p = ctypes.c_int32(), ctypes.c_int32()
    rc = L.mcv_hip_stream_probe(device, int(nbytes), int(reps), ctypes.byref(ms), ctypes.byref(v),
                                ctypes.byref(p))
    if rc:
        raise MCVError(rc, "mcv_hip_stream_probe failed")
    return ms.value, v.value, p.value


def dims(dof, n):
    h = dof - 1
    npv = n * h + 1
    return h, npv, npv + 2 * h


def _ptr(a, dtype=np.float64):
    assert a.dtype == dtype and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_I if dtype == np.int32 else _D)


def _inputs(dof, n, dt, mesh, state, scheme, precision, device):
    """ctypes structs of one solve's configuration, mesh and initial state; the
    returned list keeps the contiguous host arrays alive for the call."""
    keep = []

    def arr(a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return _ptr(a, dtype)

    cfg = Config(ABI_VERSION, dof, n, SCHEMES.get(scheme, scheme), precision, device,
                 float(dt), float(mesh["dx"]))
    m = Mesh(arr(mesh["sqrtG"]), arr(mesh["G"]), arr(mesh["IG"]), arr(mesh["A"]),
             arr(mesh["IA"]), arr(mesh["f"]), arr(mesh["phi_s"]), arr(mesh["coef"]),
             arr(mesh["iref"], np.int32))
    s = State(*[arr(state[k]) for k in STATE_FIELDS])
    return cfg, m, s, keep


def _read_state(call, nh, fields, integrals):
    """Runs a sync-like entry point `call(state*, mass*, energy*)`."""
    out = {k: np.empty((6, nh, nh)) for k in fields}
    st = State(*[_ptr(out[k]) if k in out else None for k in STATE_FIELDS])
    mass, energy = ctypes.c_double(), ctypes.c_double()
    rc = call(ctypes.byref(st), ctypes.byref(mass) if integrals else None,
              ctypes.byref(energy) if integrals else None)
    if integrals:
        out["mass"], out["energy"] = mass.value, energy.value
    return rc, out


def _history_struct(rec, nh, nc):
    """HistoryFields over the arrays of `rec` (a dict; missing fields NULL);
    returns the struct and the arrays it points into."""
    keep = {}
    for k in HISTORY_POINT_FIELDS + HISTORY_CELL_FIELDS:
        if k in rec and rec[k] is not None:
            a = np.ascontiguousarray(rec[k], dtype=np.float64)
            want = (6, nh, nh) if k in HISTORY_POINT_FIELDS else (6, nc, nc)
            if a.shape != want:
                raise ValueError(f"{k}: shape {a.shape}, expected {want}")
            keep[k] = a
    f = HistoryFields(*[_ptr(keep[k]) if k in keep else None
                        for k in HISTORY_POINT_FIELDS + HISTORY_CELL_FIELDS])
    return f, keep


def history_init(path, dof, n, dt, cells, lonP, latP, phi_s, case_num=0, dx=None, lib=None):
    """history_init (output_mod.F90:13-166): create the history file with the
    static variables.  `cells` is hostlib.cells(dof, n); lonP, latP in radians."""
    L = load(lib)
    dxr = (np.pi / 2) / n if dx is None else dx
    info = HistoryInfo(dof, n, 1, 1, case_num, dxr, dxr, float(dt))
    arrs = [np.ascontiguousarray(a, dtype=np.float64)
            for a in (cells["lonC"], cells["latC"], lonP, latP, cells["areaCell"], phi_s)]
    rc = L.mcv_history_init(os.fsencode(path), ctypes.byref(info), *[_ptr(a) for a in arrs])
    if rc != MCV_OK:
        raise MCVError(rc, L.mcv_hip_last_error().decode())


def history_write(path, time_slot, rec, dof, n, lib=None):
    """history_write_stat (output_mod.F90:168-223) of host arrays."""
    L = load(lib)
    f, keep = _history_struct(rec, dims(dof, n)[2], n + 2)
    rc = L.mcv_history_write(os.fsencode(path), int(time_slot), int(dof), int(n),
                             ctypes.byref(f))
    if rc != MCV_OK:
        raise MCVError(rc, L.mcv_hip_last_error().decode())


def footprint(dof, n, precision=FP64, scheme="RK4", device=0, lib=None):
    """Device bytes one solve of this configuration allocates at most."""
    L = load(lib)
    cfg = Config(ABI_VERSION, dof, n, SCHEMES.get(scheme, scheme), precision, device, 1.0,
                 (np.pi / 2) / n)
    b = ctypes.c_int64()
    rc = L.mcv_hip_footprint(ctypes.byref(cfg), ctypes.byref(b))
    if rc != MCV_OK:
        raise MCVError(rc, L.mcv_hip_last_error().decode())
    return b.value


class Solver:
    """One cubed-sphere solve held on the GPU (an mcv_handle)."""

    def __init__(self, dof, n, dt, mesh, state, scheme="RK4", precision=FP64, device=0,
                 lib=None):
        self.lib = load(lib)
        self.dof, self.n = dof, n
        self.h, self.npv, self.nh = dims(dof, n)
        cfg, m, s, keep = _inputs(dof, n, dt, mesh, state, scheme, precision, device)
        self._h = ctypes.c_void_p()
        self._check(self.lib.mcv_hip_init(ctypes.byref(cfg), ctypes.byref(m), ctypes.byref(s),
                                          ctypes.byref(self._h)))
        del keep

    def _check(self, rc):
        if rc != MCV_OK:
            raise MCVError(rc, self.lib.mcv_hip_last_error().decode())

    def step(self, nsteps=1):
        self._check(self.lib.mcv_hip_step(self._h, int(nsteps)))

    def sync(self, fields=STATE_FIELDS, integrals=True):
        """Blocking read-back: dict of fields (6, nh, nh) plus 'mass', 'energy'."""
        rc, out = _read_state(lambda st, m, e: self.lib.mcv_hip_sync(self._h, st, m, e),
                              self.nh, fields, integrals)
        self._check(rc)
        return out

    @property
    def device_bytes(self):
        b = ctypes.c_int64()
        self._check(self.lib.mcv_hip_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def sphere_winds(self):
        z = np.empty((6, self.nh, self.nh))
        m = np.empty((6, self.nh, self.nh))
        self._check(self.lib.mcv_hip_sphere_winds(self._h, _ptr(z), _ptr(m)))
        return z, m

    def cell_averages(self):
        """calc_VIA of u, v, phi on the device: (uC, vC, phiC), each (6, n+2, n+2)."""
        out = [np.empty((6, self.n + 2, self.n + 2)) for _ in range(3)]
        self._check(self.lib.mcv_hip_cell_averages(self._h, *[_ptr(o) for o in out]))
        return tuple(out)

    def history_fields(self):
        """The record fields of the output block after a step (dict)."""
        rec = {k: np.empty((6, self.nh, self.nh)) for k in HISTORY_POINT_FIELDS}
        rec.update({k: np.empty((6, self.n + 2, self.n + 2)) for k in HISTORY_CELL_FIELDS})
        f, _ = _history_struct(rec, self.nh, self.n + 2)
        self._check(self.lib.mcv_hip_history_fields(self._h, ctypes.byref(f)))
        return rec

    def history_write(self, path, time_slot):
        """The output block of MCV_SW.F90:62-70: record `time_slot` of `path`."""
        self._check(self.lib.mcv_hip_history_write(self._h, os.fsencode(path), int(time_slot)))

    def tendency(self, state, noalloc=False):
        """The spatial operator alone on `state`; noalloc: through
        mcv_hip_tendency_noalloc (library scratch, no allocation in the call)."""
        keep = [np.ascontiguousarray(state[k], dtype=np.float64) for k in STATE_FIELDS]
        st = State(*[_ptr(a) for a in keep])
        out = [np.empty((6, self.npv, self.npv)) for _ in range(3)]
        fn = self.lib.mcv_hip_tendency_noalloc if noalloc else self.lib.mcv_hip_tendency
        self._check(fn(self._h, ctypes.byref(st), *[_ptr(o) for o in out]))
        return dict(zip(("phiG", "u", "v"), out))

    def autotune(self, reps=5):
        """Time the stage-kernel tile shapes on this solve, keep the fastest;
        returns (chosen index, [ms per stage launch of each shape])."""
        t = ctypes.c_int32()
        ms = (ctypes.c_double * TILE_VARIANTS)()
        self._check(self.lib.mcv_hip_autotune(self._h, int(reps), ctypes.byref(t), ms))
        return t.value, list(ms)

    def set_tile(self, tile):
        self._check(self.lib.mcv_hip_set_tile(self._h, int(tile)))

    def step_timed(self, nsteps):
        """Advance nsteps with a HIP event after every kernel; returns the
        summed (stage kernel ms, boundary kernel ms)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mcv_hip_step_timed(self._h, int(nsteps), ctypes.byref(a),
                                                ctypes.byref(b)))
        return a.value, b.value

    @property
    def steps_done(self):
        s = ctypes.c_int64()
        self._check(self.lib.mcv_hip_steps_done(self._h, ctypes.byref(s)))
        return s.value

    def close(self):
        if self._h:
            self.lib.mcv_hip_free(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Solves of mixed order, size, precision and scheme advanced together on
    one GPU (an mcv_batch): one fused stage launch and one fused boundary
    launch per RK stage for all of them, replayed from a captured hipGraph."""

    GRAPH_STAGES = 12

    def __init__(self, capacity, device=0, lib=None):
        self.lib = load(lib)
        self.capacity = capacity
        self._b = ctypes.c_void_p()
        self.info = {}
        self._check(self.lib.mcv_batch_create(int(device), int(capacity), ctypes.byref(self._b)))
        self.device = device

    def _check(self, rc):
        if rc != MCV_OK:
            raise MCVError(rc, self.lib.mcv_hip_last_error().decode())

    def add(self, dof, n, dt, mesh, state, nsteps, scheme="RK4", precision=FP64):
        """Admit a solve that will run nsteps time steps; returns its slot."""
        cfg, m, s, keep = _inputs(dof, n, dt, mesh, state, scheme, precision, self.device)
        slot = ctypes.c_int32()
        self._check(self.lib.mcv_batch_add(self._b, ctypes.byref(cfg), ctypes.byref(m),
                                           ctypes.byref(s), int(nsteps), ctypes.byref(slot)))
        del keep
        self.info[slot.value] = dict(dof=dof, n=n, dt=dt, scheme=scheme, precision=precision,
                                     nsteps=nsteps)
        return slot.value

    def remove(self, slot):
        self._check(self.lib.mcv_batch_remove(self._b, int(slot)))
        self.info.pop(slot, None)

    def run(self, nstages):
        """Enqueue nstages stage launches (each advances every solve one stage)."""
        self._check(self.lib.mcv_batch_run(self._b, int(nstages)))

    def poll(self):
        """(steps_left, steps_done) per slot, -1 for free slots (blocks)."""
        left = np.empty(self.capacity, dtype=np.int64)
        done = np.empty(self.capacity, dtype=np.int64)
        self._check(self.lib.mcv_batch_poll(
            self._b, left.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            done.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return left, done

    def sync(self, slot, fields=STATE_FIELDS, integrals=True):
        nh = dims(self.info[slot]["dof"], self.info[slot]["n"])[2]
        rc, out = _read_state(lambda st, m, e: self.lib.mcv_batch_sync(self._b, int(slot), st, m, e),
                              nh, fields, integrals)
        self._check(rc)
        return out

    def autotune(self, reps=3):
        t = ctypes.c_int32()
        ms = (ctypes.c_double * TILE_VARIANTS)()
        self._check(self.lib.mcv_batch_autotune(self._b, int(reps), ctypes.byref(t), ms))
        return t.value, list(ms)

    def set_tile(self, tile):
        self._check(self.lib.mcv_batch_set_tile(self._b, int(tile)))

    def set_graphs(self, on):
        self._check(self.lib.mcv_batch_set_graphs(self._b, 1 if on else 0))

    def run_timed(self, nstages):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.mcv_batch_run_timed(self._b, int(nstages), ctypes.byref(a),
                                                 ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if self._b:
            self.lib.mcv_batch_free(self._b)
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
