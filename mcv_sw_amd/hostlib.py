"""ctypes face of the Fortran host setup library (mcv_sw_amd/host/mcv_host_lib.F90).

Builds what the reference's host modules build before the time loop:
geometry (mesh_mod/projection_mod), ghost tables (ghost_mod initGhost) and
initial states (test_case_mod), in the reference layout used by capi.Solver.
"""
import ctypes
import os

import numpy as np

from .capi import dims

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcv_host.so")

# constants_mod.F90:12-18
GRAVITY = 9.80616
RADIUS = 6371220.0
OMEGA = 7.292e-5
PI = 2.0 * np.arcsin(1.0)

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        ci, cd = ctypes.c_int, ctypes.c_double
        L.mcv_host_mesh.argtypes = [ci, ci, cd, cd] + [_D] * 10
        L.mcv_host_ghost.argtypes = [ci, ci, _D, _I]
        L.mcv_host_case.argtypes = [ci, ci, cd, cd, cd] + [_D] * 6
        L.mcv_host_jet.argtypes = [ci, cd, cd, cd, cd, cd, ci, ctypes.c_int64, cd] + [_D] * 6
        L.mcv_host_handoff.argtypes = [ci] + [_D] * 11
        L.mcv_host_cells.argtypes = [ci, ci, cd, _D, _D, _D]
        for f in ("mcv_host_mesh", "mcv_host_ghost", "mcv_host_case", "mcv_host_jet",
                  "mcv_host_handoff", "mcv_host_cells"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I if a.dtype == np.int32 else _D)


def mesh(dof, n, radius=RADIUS, omega=OMEGA):
    """Geometry + ghost tables of the 6 x n x n mesh (dict of arrays, see capi)."""
    h, npv, nh = dims(dof, n)
    s = (6, nh, nh)
    m = {k: np.empty(s) for k in ("xP", "yP", "lonP", "latP", "sqrtG", "f")}
    for k in ("G", "IG", "A", "IA"):
        m[k] = np.empty(s + (2, 2))
    rc = load().mcv_host_mesh(dof, n, radius, omega,
                              *[_p(m[k]) for k in ("xP", "yP", "lonP", "latP", "sqrtG",
                                                   "G", "IG", "A", "IA", "f")])
    if rc:
        raise ValueError(f"mcv_host_mesh(dof={dof}, n={n}) failed: {rc}")
    m["coef"] = np.empty((h, npv))
    m["iref"] = np.empty((h, npv), dtype=np.int32)
    rc = load().mcv_host_ghost(dof, n, _p(m["coef"]), _p(m["iref"]))
    if rc:
        raise ValueError(f"mcv_host_ghost(dof={dof}, n={n}) failed: {rc}")
    m["phi_s"] = np.zeros(s)
    m["dx"] = (90.0 / n) * (PI / 180.0)
    m["dof"], m["n"] = dof, n
    return m


def cells(dof, n, radius=RADIUS):
    """Cell geometry of the history file (mesh_mod.F90:165-214): lonC, latC
    (radians) and areaCell (m^2), each (6, n+2, n+2), the halo ring of
    areaCell zero."""
    s = (6, n + 2, n + 2)
    c = {k: np.empty(s) for k in ("lonC", "latC", "areaCell")}
    rc = load().mcv_host_cells(dof, n, radius, _p(c["lonC"]), _p(c["latC"]), _p(c["areaCell"]))
    if rc:
        raise ValueError(f"mcv_host_cells(dof={dof}, n={n}) failed: {rc}")
    return c


def handoff(m, phi, zonal, merid):
    """test_case_mod.F90:30-39: the full state the step starts from."""
    np_ = phi.size
    st = {"phi": np.ascontiguousarray(phi, dtype=np.float64)}
    for k in ("phiG", "u", "v", "contraU", "contraV"):
        st[k] = np.empty_like(st["phi"])
    z = np.ascontiguousarray(zonal, dtype=np.float64)
    w = np.ascontiguousarray(merid, dtype=np.float64)
    rc = load().mcv_host_handoff(np_, _p(m["sqrtG"]), _p(m["G"]), _p(m["IA"]), _p(st["phi"]),
                                 _p(z), _p(w), _p(st["phiG"]), _p(st["u"]), _p(st["v"]),
                                 _p(st["contraU"]), _p(st["contraV"]))
    if rc:
        raise RuntimeError(f"mcv_host_handoff failed: {rc}")
    st["zonal"], st["merid"] = z, w
    return st


def williamson(m, case_num, radius=RADIUS, omega=OMEGA, gravity=GRAVITY):
    """Williamson test case 2, 5 or 6 (test_case_mod.F90:52-191); sets m['phi_s']."""
    s = m["lonP"].shape
    phi, z, w, hs = (np.empty(s) for _ in range(4))
    rc = load().mcv_host_case(case_num, m["lonP"].size, radius, omega, gravity, _p(m["lonP"]),
                              _p(m["latP"]), _p(phi), _p(z), _p(w), _p(hs))
    if rc:
        raise ValueError(f"unknown test case {case_num}")
    m["phi_s"] = hs
    return handoff(m, phi, z, w)


def jet(m, u0=20.0, h0=5960.0, lmax=8, seed=0, rms_h=30.0, radius=RADIUS, omega=OMEGA,
        gravity=GRAVITY):
    """Balanced zonal jet + seeded spherical-harmonic height perturbation with its
    matching (regularised geostrophic) wind; sets m['phi_s'] = 0."""
    s = m["lonP"].shape
    phi, z, w, hs = (np.empty(s) for _ in range(4))
    rc = load().mcv_host_jet(m["lonP"].size, radius, omega, gravity, u0, h0, lmax, seed, rms_h,
                             _p(m["lonP"]), _p(m["latP"]), _p(phi), _p(z), _p(w), _p(hs))
    if rc:
        raise ValueError(f"mcv_host_jet failed: {rc}")
    m["phi_s"] = hs
    return handoff(m, phi, z, w)
