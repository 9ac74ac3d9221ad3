"""Inputs for the parity tests: mesh from the oracle's restated geometry and
analytic initial states (Williamson cases of test_case_mod.F90, a balanced
jet with a bump), handed over as test_case_mod.F90:30-39 does.

Test infrastructure: these build inputs only.  The GPU path and the oracle
receive identical arrays, so the comparison does not depend on how the
inputs were made.
"""
import numpy as np

from oracle import oracle

GRAVITY = 9.80616
RADIUS = 6371220.0
OMEGA = 7.292e-5
PI = 2.0 * np.arcsin(1.0)


def mesh(dof, n):
    m = oracle.mesh(dof, n)
    m["phi_s"] = np.zeros_like(m["sqrtG"])
    return m


def handoff(m, phi, zonal, merid):
    """test_case_mod.F90:30-39: phiG, covariant and contravariant winds."""
    G, IA = m["G"], m["IA"]
    # (G*IA)(r, c) = G(r,1) IA(1,c) + G(r,2) IA(2,c); arrays are [..., col, row]
    M = np.empty_like(G)
    for c in range(2):
        for r in range(2):
            M[..., c, r] = G[..., 0, r] * IA[..., c, 0] + G[..., 1, r] * IA[..., c, 1]
    st = {
        "phi": phi.copy(),
        "phiG": phi * m["sqrtG"],
        "u": M[..., 0, 0] * zonal + M[..., 1, 0] * merid,
        "v": M[..., 0, 1] * zonal + M[..., 1, 1] * merid,
        "contraU": IA[..., 0, 0] * zonal + IA[..., 1, 0] * merid,
        "contraV": IA[..., 0, 1] * zonal + IA[..., 1, 1] * merid,
        "zonal": zonal.copy(),
        "merid": merid.copy(),
    }
    return st


def case2(m):
    """Williamson case 2 (test_case_mod.F90:52-81)."""
    lat = m["latP"]
    u0 = 2.0 * PI * RADIUS / (12.0 * 86400.0)
    phi = 29400.0 - (RADIUS * OMEGA * u0 + u0 ** 2 / 2.0) * np.sin(lat) ** 2
    m["phi_s"] = np.zeros_like(phi)
    return handoff(m, phi, u0 * np.cos(lat), np.zeros_like(phi))


def case5(m):
    """Williamson case 5, isolated mountain (test_case_mod.F90:84-130)."""
    lon = np.where(m["lonP"] < 0, 2.0 * PI + m["lonP"], m["lonP"])
    lat = m["latP"]
    rr, lc, tc, u0, hs0 = PI / 9.0, 1.5 * PI, PI / 6.0, 20.0, 2000.0
    r = np.sqrt(np.minimum(rr ** 2, (lon - lc) ** 2 + (lat - tc) ** 2))
    hs = GRAVITY * hs0 * (1.0 - r / rr)
    phi = 5960.0 * GRAVITY - (RADIUS * OMEGA * u0 + u0 ** 2 / 2.0) * np.sin(lat) ** 2 - hs
    m["phi_s"] = hs
    return handoff(m, phi, u0 * np.cos(lat), np.zeros_like(phi))


def bumped_jet(m, u0=20.0, h0=5960.0, amp=30.0):
    """Balanced zonal jet plus an unbalanced Gaussian height bump."""
    lon, lat = m["lonP"], m["latP"]
    phi = GRAVITY * h0 - (RADIUS * OMEGA * u0 + u0 ** 2 / 2.0) * np.sin(lat) ** 2
    d2 = (lon - 0.7) ** 2 + (lat - 0.4) ** 2
    phi = phi + GRAVITY * amp * np.exp(-d2 / 0.05)
    m["phi_s"] = np.zeros_like(phi)
    return handoff(m, phi, u0 * np.cos(lat), np.zeros_like(phi))


def rotated_jet(m, alpha, u0=20.0, h0=5960.0, amp=60.0):
    """Williamson case 2 with its axis tilted by `alpha` (test_case_mod.F90:52-81
    at alpha = 0), plus the unbalanced Gaussian height bump of bumped_jet."""
    lon, lat = m["lonP"], m["latP"]
    zonal = u0 * (np.cos(lat) * np.cos(alpha) + np.cos(lon) * np.sin(lat) * np.sin(alpha))
    merid = -u0 * np.sin(lon) * np.sin(alpha)
    s = -np.cos(lon) * np.cos(lat) * np.sin(alpha) + np.sin(lat) * np.cos(alpha)
    phi = GRAVITY * h0 - (RADIUS * OMEGA * u0 + u0 ** 2 / 2.0) * s ** 2
    d2 = (lon - 0.7) ** 2 + (lat - 0.4) ** 2
    phi = phi + GRAVITY * amp * np.exp(-d2 / 0.05)
    m["phi_s"] = np.zeros_like(phi)
    return handoff(m, phi, zonal, merid)


def masks(dof, n):
    """Boolean (nh, nh) masks: interior, edge lines, halo strips (corners excluded)."""
    h = dof - 1
    npv = n * h + 1
    nh = npv + 2 * h
    inside = np.zeros((nh, nh), bool)
    inside[h:h + npv, h:h + npv] = True
    corner = np.zeros((nh, nh), bool)
    for a in (slice(0, h), slice(h + npv, nh)):
        for b in (slice(0, h), slice(h + npv, nh)):
            corner[a, b] = True
    strips = ~inside & ~corner
    edges = inside.copy()
    edges[h + 1:h + npv - 1, h + 1:h + npv - 1] = False
    return {"interior": inside, "edges": edges, "strips": strips, "defined": inside | strips}


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
