"""Python face of the C restatement (oracle/mcv_oracle.c).

TEST INFRASTRUCTURE ONLY.  The test suite, the smoke check and the benchmark's
CPU-baseline leg use it as the checker; the product never imports it.

All arrays are numpy float64 in the reference layout: Fortran order
(i, j, panel) with the halo, i.e. C-contiguous numpy arrays of shape
(6, nh, nh) indexed [panel, j, i]; 2x2 matrices have shape (6, nh, nh, 2, 2)
indexed [panel, j, i, col, row] (Fortran (2,2,i,j,p) read back in C order).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmcv_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


def build():
    """Compile the restatement with gcc (make -C oracle lib)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "lib"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_mesh_build.argtypes = [ctypes.c_int, ctypes.c_int] + [_D] * 10
        L.orc_ghost_build.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D, _I]
        L.orc_run.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [_D] * 8 +
                              [_I] + [_D] * 8 + [ctypes.c_int])
        L.orc_tendency.argtypes = ([ctypes.c_int, ctypes.c_int, ctypes.c_double] + [_D] * 13)
        L.orc_boundary.argtypes = ([ctypes.c_int, ctypes.c_int, ctypes.c_double] + [_D] * 5 +
                                   [_I] + [_D] * 7)
        L.orc_diag.argtypes = [ctypes.c_int, ctypes.c_int] + [_D] * 10
        for f in ("orc_mesh_build", "orc_ghost_build", "orc_run", "orc_tendency",
                  "orc_boundary", "orc_diag"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.int32:
        return a.ctypes.data_as(_I)
    assert a.dtype == np.float64
    return a.ctypes.data_as(_D)


def dims(dof, n):
    h = dof - 1
    npv = n * h + 1
    return h, npv, npv + 2 * h


def mesh(dof, n):
    """Mesh of mesh_mod.F90:132-163 as restated; returns a dict of arrays."""
    h, npv, nh = dims(dof, n)
    s = (6, nh, nh)
    m = {k: np.zeros(s) for k in ("xP", "yP", "lonP", "latP", "sqrtG", "f")}
    for k in ("G", "IG", "A", "IA"):
        m[k] = np.zeros(s + (2, 2))
    rc = lib().orc_mesh_build(dof, n, *[_p(m[k]) for k in ("xP", "yP", "lonP", "latP", "sqrtG",
                                                           "G", "IG", "A", "IA", "f")])
    if rc:
        raise RuntimeError(f"orc_mesh_build failed: {rc}")
    coef = np.zeros((h, npv))
    iref = np.zeros((h, npv), dtype=np.int32)
    rc = lib().orc_ghost_build(dof, n, _p(m["xP"]), _p(m["yP"]), _p(coef), _p(iref))
    if rc:
        raise RuntimeError(f"orc_ghost_build failed: {rc}")
    m["coef"], m["iref"] = coef, iref
    m["dx"] = (90.0 / n) * (2.0 * np.arcsin(1.0) / 180.0)
    return m


FIELDS = ("phi", "phiG", "u", "v", "contraU", "contraV", "zonal", "merid")


def run(dof, n, scheme, dt, mesh_arrays, state, nsteps):
    """Advance `state` (dict of the 8 FIELDS) by nsteps; returns a new dict."""
    m = mesh_arrays
    st = {k: np.ascontiguousarray(state[k], dtype=np.float64).copy() for k in FIELDS}
    rc = lib().orc_run(dof, n, scheme, dt, m["dx"], _p(m["sqrtG"]), _p(m["G"]), _p(m["IG"]),
                       _p(m["A"]), _p(m["IA"]), _p(m["f"]), _p(m["phi_s"]), _p(m["coef"]),
                       _p(m["iref"]), *[_p(st[k]) for k in FIELDS], nsteps)
    if rc:
        raise RuntimeError(f"orc_run failed: {rc}")
    return st


def tendency(dof, n, mesh_arrays, state):
    m = mesh_arrays
    h, npv, nh = dims(dof, n)
    out = [np.zeros((6, npv, npv)) for _ in range(3)]
    rc = lib().orc_tendency(dof, n, m["dx"], _p(m["sqrtG"]), _p(m["IG"]), _p(m["f"]),
                            _p(m["phi_s"]), *[_p(np.ascontiguousarray(state[k]))
                                              for k in FIELDS[:6]], *[_p(o) for o in out])
    if rc:
        raise RuntimeError(f"orc_tendency failed: {rc}")
    return dict(zip(("phiG", "u", "v"), out))


def boundary(dof, n, mesh_arrays, state):
    m = mesh_arrays
    st = {k: np.ascontiguousarray(state[k], dtype=np.float64).copy() for k in FIELDS}
    rc = lib().orc_boundary(dof, n, m["dx"], _p(m["G"]), _p(m["IG"]), _p(m["A"]), _p(m["IA"]),
                            _p(m["coef"]), _p(m["iref"]), _p(st["phi"]), _p(st["u"]),
                            _p(st["v"]), _p(st["contraU"]), _p(st["contraV"]), _p(st["zonal"]),
                            _p(st["merid"]))
    if rc:
        raise RuntimeError(f"orc_boundary failed: {rc}")
    return st


def diag(dof, n, mesh_arrays, state):
    m = mesh_arrays
    mass = ctypes.c_double()
    energy = ctypes.c_double()
    rc = lib().orc_diag(dof, n, _p(m["sqrtG"]), _p(m["phi_s"]),
                        *[_p(np.ascontiguousarray(state[k])) for k in FIELDS[:6]],
                        ctypes.byref(mass), ctypes.byref(energy))
    if rc:
        raise RuntimeError(f"orc_diag failed: {rc}")
    return mass.value, energy.value
