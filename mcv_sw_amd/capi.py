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


class MCVError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"MCV error {code}: {msg}")
        self.code = code


_libs = {}


def load(path=None):
    """Load the C-ABI library (default: the in-tree HIP build)."""
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
    L.mcv_hip_steps_done.argtypes = [H, ctypes.POINTER(ctypes.c_int64)]
    L.mcv_hip_step_timed.argtypes = [H, ctypes.c_int32, _D, _D]
    L.mcv_hip_copy_probe.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _D]
    L.mcv_hip_copy_probe.restype = ctypes.c_int
    L.mcv_hip_autotune.argtypes = [H, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _D]
    L.mcv_hip_autotune.restype = ctypes.c_int
    L.mcv_hip_set_tile.argtypes = [H, ctypes.c_int32]
    L.mcv_hip_set_tile.restype = ctypes.c_int
    L.mcv_hip_free.argtypes = [H]
    L.mcv_hip_last_error.argtypes = []
    L.mcv_hip_last_error.restype = ctypes.c_char_p
    for f in ("mcv_hip_init", "mcv_hip_step", "mcv_hip_sync", "mcv_hip_sphere_winds",
              "mcv_hip_tendency", "mcv_hip_steps_done", "mcv_hip_step_timed", "mcv_hip_free"):
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


def dims(dof, n):
    h = dof - 1
    npv = n * h + 1
    return h, npv, npv + 2 * h


def _ptr(a, dtype=np.float64):
    assert a.dtype == dtype and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_I if dtype == np.int32 else _D)


class Solver:
    """One cubed-sphere solve held on the GPU (an mcv_handle)."""

    def __init__(self, dof, n, dt, mesh, state, scheme="RK4", precision=FP64, device=0,
                 lib=None):
        self.lib = load(lib)
        self.dof, self.n = dof, n
        self.h, self.npv, self.nh = dims(dof, n)
        cfg = Config(ABI_VERSION, dof, n, SCHEMES.get(scheme, scheme), precision, device,
                     float(dt), float(mesh["dx"]))
        self._keep = []

        def arr(a, dtype=np.float64):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return _ptr(a, dtype)

        m = Mesh(arr(mesh["sqrtG"]), arr(mesh["G"]), arr(mesh["IG"]), arr(mesh["A"]),
                 arr(mesh["IA"]), arr(mesh["f"]), arr(mesh["phi_s"]), arr(mesh["coef"]),
                 arr(mesh["iref"], np.int32))
        s = State(*[arr(state[k]) for k in STATE_FIELDS])
        self._h = ctypes.c_void_p()
        self._check(self.lib.mcv_hip_init(ctypes.byref(cfg), ctypes.byref(m), ctypes.byref(s),
                                          ctypes.byref(self._h)))
        self._keep = []

    def _check(self, rc):
        if rc != MCV_OK:
            raise MCVError(rc, self.lib.mcv_hip_last_error().decode())

    def step(self, nsteps=1):
        self._check(self.lib.mcv_hip_step(self._h, int(nsteps)))

    def sync(self, fields=STATE_FIELDS, integrals=True):
        """Blocking read-back: dict of fields (6, nh, nh) plus 'mass', 'energy'."""
        out = {k: np.empty((6, self.nh, self.nh)) for k in fields}
        st = State(*[_ptr(out[k]) if k in out else None for k in STATE_FIELDS])
        mass, energy = ctypes.c_double(), ctypes.c_double()
        rc = self.lib.mcv_hip_sync(self._h, ctypes.byref(st),
                                   ctypes.byref(mass) if integrals else None,
                                   ctypes.byref(energy) if integrals else None)
        if integrals:
            out["mass"], out["energy"] = mass.value, energy.value
        self._check(rc)
        return out

    def sphere_winds(self):
        z = np.empty((6, self.nh, self.nh))
        m = np.empty((6, self.nh, self.nh))
        self._check(self.lib.mcv_hip_sphere_winds(self._h, _ptr(z), _ptr(m)))
        return z, m

    def tendency(self, state):
        keep = [np.ascontiguousarray(state[k], dtype=np.float64) for k in STATE_FIELDS]
        st = State(*[_ptr(a) for a in keep])
        out = [np.empty((6, self.npv, self.npv)) for _ in range(3)]
        self._check(self.lib.mcv_hip_tendency(self._h, ctypes.byref(st), *[_ptr(o) for o in out]))
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
