"""Closed-form and convergence pins of the oracle, independent of the reference.

Each check fails on a plausible restatement mistake on its own:
  - rest state: any asymmetric stencil weight or sign error gives a nonzero
    tendency (exactly zero is required);
  - steady geostrophic flow (Williamson case 2): the tendency must vanish at
    the design order DOF as the mesh is refined (a dropped term or a wrong
    polyfit/interior coefficient drops the order);
  - ghost interpolation of a smooth field: halo values converge to the
    analytic ones at order DOF (a wrong orientation or index is O(1));
  - discrete mass: the w-weighted sum of the phiG tendency telescopes to zero
    up to round-off (diag_mod.F90:10-36 weights);
  - RK4 / SSP-RK3: the local error of one step falls as dt^5 / dt^4 in the
    panel interiors (a wrong stage weight drops the order; the planted
    mistake of the negative control does).
The same pins run on the emulated product code.
"""
import numpy as np
import pytest

import cases
from mcv_sw_amd import capi, hostlib
from oracle import oracle

FIELDS = ("phi", "phiG", "u", "v", "contraU", "contraV")


def _tend(dof, n, m, st, emu=None):
    if emu is None:
        return oracle.tendency(dof, n, m, st)
    with capi.Solver(dof, n, 60.0, m, st, lib=emu) as s:
        return s.tendency(st)


@pytest.mark.parametrize("dof", [3, 4])
@pytest.mark.parametrize("impl", ["oracle", "emulated"])
def test_rest_state_has_zero_momentum_tendency(dof, impl, emu_lib):
    """At rest with constant phi the energy E is constant, so every one-sided
    fit, Riemann flux and Coriolis term of the momentum equations is exactly
    zero.  (The mass tendency is not: the Rusanov term acts on the jumps of
    the one-sided derivatives of phiG = phi*sqrtG, which is not polynomial;
    it vanishes with the mesh spacing, see the design-order test below.)"""
    n = 6
    m = hostlib.mesh(dof, n)
    phi = np.full_like(m["sqrtG"], 5.0e4)  # small integer: every polyfit product is exact
    zero = np.zeros_like(phi)
    st = hostlib.handoff(m, phi, zero, zero)
    t = _tend(dof, n, m, st, None if impl == "oracle" else emu_lib)
    for k in ("u", "v"):
        assert np.all(t[k] == 0.0), k
    assert np.all(np.isfinite(t["phiG"]))


def _case2_residual(dof, n):
    m = hostlib.mesh(dof, n)
    st = hostlib.williamson(m, 2)
    t = oracle.tendency(dof, n, m, st)
    h, npv, nh = capi.dims(dof, n)
    sl = (slice(None), slice(h, h + npv), slice(h, h + npv))
    # relative to the balanced terms (pressure gradient ~ Coriolis ~ f u)
    scale_u = np.abs(st["u"][sl]).max() * 1.4e-4
    return max(np.abs(t["u"]).max(), np.abs(t["v"]).max()) / scale_u


@pytest.mark.parametrize("dof", [3, 4])
def test_steady_flow_tendency_converges_at_design_order(dof):
    """Point values of a degree DOF-1 reconstruction have derivatives accurate
    to O(dx^(DOF-1)) (measured 2.0 for MCV3 and 3.0 for MCV4); a dropped term
    or wrong fit coefficient leaves an O(1) or growing residual."""
    e1, e2 = _case2_residual(dof, 12), _case2_residual(dof, 24)
    order = np.log2(e1 / e2)
    assert order > dof - 1.4, f"observed order {order:.2f} for MCV{dof}"


def _ghost_error(dof, n):
    """Max halo error of phi = a smooth analytic field after the boundary pass."""
    m = hostlib.mesh(dof, n)
    lon, lat = m["lonP"], m["latP"]

    def field(lo, la):
        return 5.0e4 + 3.0e3 * np.cos(la) ** 2 * np.cos(2 * lo) + 1.0e3 * np.sin(la)

    phi = field(lon, lat)
    zon = 10.0 * np.cos(lat) * (1.0 + 0.3 * np.sin(lon))
    mer = 5.0 * np.sin(2 * lon) * np.cos(lat)
    st = hostlib.handoff(m, phi, zon, mer)
    exact = {"phi": st["phi"].copy(), "u": st["u"].copy(), "v": st["v"].copy()}
    out = oracle.boundary(dof, n, m, st)
    strips = cases.masks(dof, n)["strips"]
    return {k: np.abs(out[k][:, strips] - exact[k][:, strips]).max() / np.abs(exact[k]).max()
            for k in ("phi", "u", "v")}


@pytest.mark.parametrize("dof", [3, 4])
def test_ghost_interpolation_converges_at_design_order(dof):
    e1, e2 = _ghost_error(dof, 10), _ghost_error(dof, 20)
    for k in ("phi", "u", "v"):
        order = np.log2(e1[k] / e2[k])
        assert order > dof - 0.5, f"{k}: ghost order {order:.2f} for MCV{dof}"
        assert e2[k] < 1e-3


@pytest.mark.parametrize("dof", [3, 4])
def test_tendency_conserves_mass_discretely(dof):
    n = 8
    m = hostlib.mesh(dof, n)
    st = hostlib.jet(m, u0=30.0, h0=6000.0, lmax=6, seed=3, rms_h=80.0)
    t = oracle.tendency(dof, n, m, st)["phiG"]
    w1 = np.array([1.0, 4.0, 1.0]) / 6.0 if dof == 3 else np.array([1.0, 3.0, 3.0, 1.0]) / 8.0
    h = dof - 1
    tot, mag = 0.0, 0.0
    for p in range(6):
        for cj in range(n):
            for ci in range(n):
                blk = t[p, cj * h:cj * h + dof, ci * h:ci * h + dof]
                wb = np.outer(w1, w1)
                tot += float((wb * blk).sum())
                mag += float((wb * np.abs(blk)).sum())
    assert abs(tot) <= 1e-12 * mag


def _compose_step(dof, n, m, y, scheme, dt, planted=False):
    """time_scheme_mod.F90:13-123 composed from the operator and the boundary
    pass, in numpy (elementwise IEEE arithmetic, no contraction).  With
    `planted`, a plausible weight mistake for the order test's negative
    control: RK4 final weights (2, 1, 1, 2)/6, SSP-RK3 second stage with
    (1/2) dt k2 in place of (1/4) dt k2."""
    h, npv, nh = capi.dims(dof, n)
    I = (slice(None), slice(h, h + npv), slice(h, h + npv))
    sg = m["sqrtG"]

    def finish(s):
        s["phi"] = s["phiG"] / sg
        s = oracle.boundary(dof, n, m, s)
        s["phiG"] = s["phi"] * sg
        return s

    def update(base, coef_fn):
        s = {k: v.copy() for k, v in base.items()}
        for k in ("phiG", "u", "v"):
            s[k][I] = coef_fn(k)
        return finish(s)

    k1 = oracle.tendency(dof, n, m, y)
    if scheme == 4:
        y2 = update(y, lambda f: y[f][I] + (0.5 * dt) * k1[f])
        k2 = oracle.tendency(dof, n, m, y2)
        y3 = update(y, lambda f: y[f][I] + (0.5 * dt) * k2[f])
        k3 = oracle.tendency(dof, n, m, y3)
        y4 = update(y, lambda f: y[f][I] + dt * k3[f])
        k4 = oracle.tendency(dof, n, m, y4)
        w = (2.0, 1.0, 1.0, 2.0) if planted else (1.0, 2.0, 2.0, 1.0)
        t = {f: (w[0] * k1[f] + w[1] * k2[f] + w[2] * k3[f] + w[3] * k4[f]) / 6.0
             for f in ("phiG", "u", "v")}
        return update(y, lambda f: y[f][I] + dt * t[f])
    y1 = update(y, lambda f: y[f][I] + dt * k1[f])
    k2 = oracle.tendency(dof, n, m, y1)
    c2 = 0.5 if planted else 0.25
    y2 = update(y, lambda f: 0.75 * y[f][I] + 0.25 * y1[f][I] + (c2 * dt) * k2[f])
    k3 = oracle.tendency(dof, n, m, y2)
    return update(y, lambda f: y[f][I] / 3.0 + (2.0 / 3.0) * y2[f][I] + ((2.0 / 3.0) * dt) * k3[f])


@pytest.mark.parametrize("scheme", [4, 3])
@pytest.mark.parametrize("dof", [3, 4])
def test_rk_step_is_the_reference_composition(dof, scheme):
    """One oracle step equals RK4 / SSP-RK3 built from the separately pinned
    operator and boundary pass.  (A global self-convergence test in dt cannot
    pin the RK weights on this scheme: the boundary pass makes the step's
    error at the panel edges O(dt^2), and the upwind switch of the cross-wind
    derivative is discontinuous where the contravariant wind changes sign, so
    the whole-sphere error converges at first order; the weights are pinned
    in the panel interiors by test_rk_weights_pinned_by_local_error_order.)"""
    n = 5
    m = hostlib.mesh(dof, n)
    st = hostlib.jet(m, u0=25.0, h0=6000.0, lmax=5, seed=2, rms_h=40.0)
    y0 = oracle.run(dof, n, scheme, 300.0, m, st, 1)
    ref = oracle.run(dof, n, scheme, 300.0, m, y0, 1)
    got = _compose_step(dof, n, m, y0, scheme, 300.0)
    msk = cases.masks(dof, n)["defined"]
    for k in ("phi", "phiG", "u", "v", "contraU", "contraV"):
        assert np.array_equal(got[k][:, msk], ref[k][:, msk]), k


def _interior_local_error(step, dof, n, st, dt, dmin=6):
    """max over h, u, v of |one step of dt - two steps of dt/2|, relative to
    the field's maximum, over the points at least `dmin` points away from
    every panel edge."""
    h, npv, nh = capi.dims(dof, n)
    jj, ii = np.meshgrid(np.arange(npv), np.arange(npv), indexing="ij")
    far = np.minimum(np.minimum(ii, npv - 1 - ii), np.minimum(jj, npv - 1 - jj)) >= dmin
    a, b = step(st, dt, 1), step(st, 0.5 * dt, 2)
    return max(np.abs(a[f] - b[f])[:, h:h + npv, h:h + npv][:, far].max() / np.abs(st[f]).max()
               for f in ("phi", "u", "v"))


@pytest.mark.parametrize("scheme", [4, 3])
@pytest.mark.parametrize("dof", [3, 4])
def test_rk_weights_pinned_by_local_error_order(dof, scheme):
    """The RK weights against the mathematics, not a restatement: the local
    error of one step (one step of dt against two of dt/2) of RK4 / SSP-RK3
    falls as dt^5 / dt^4, so halving dt divides it by 32 / 16.  Measured in
    the panel interiors (>= 6 points from every edge) of a rotated jet with a
    height bump: at the panel edges the reference's boundary pass fills the
    ghosts from the edge lines before it averages them, so the pass is not a
    projection and the step's error there is O(dt^2) whatever the weights
    (time_scheme_mod.F90:136-140).  The flow (rotation 0.7 rad, 60 m bump,
    6x8x8 mesh, dt 80/40/20 s) keeps the upwind switch of the cross-wind
    derivative away from its sign change over these steps.  The negative
    control composes the same step with a plausible weight mistake and must
    lose the order."""
    n = 8
    m = cases.mesh(dof, n)
    st = oracle.run(dof, n, 4, 1.0, m, cases.rotated_jet(m, 0.7, amp=60.0), 1)
    want = 2.0 ** (scheme + 1)

    def ref_step(s, dt, k):
        return oracle.run(dof, n, scheme, dt, m, s, k)

    def planted_step(s, dt, k):
        for _ in range(k):
            s = _compose_step(dof, n, m, s, scheme, dt, planted=True)
        return s

    e = [_interior_local_error(ref_step, dof, n, st, dt) for dt in (80.0, 40.0, 20.0)]
    r = [e[0] / e[1], e[1] / e[2]]
    print(f"[rk order] dof={dof} RK{scheme}: local errors {e}, ratios {r}", flush=True)
    assert all(0.85 * want <= x <= 1.15 * want for x in r), (e, r)
    ep = [_interior_local_error(planted_step, dof, n, st, dt) for dt in (40.0, 20.0)]
    assert ep[0] / ep[1] < 0.6 * want, ep


def test_exact_division_helper_is_ieee_division(tmp_path):
    """The kernels' division by a known divisor (q = a*y, r = fma(-q, b, a),
    copysign(fma(r, y, q), q) with y = RN(1/b)) equals IEEE a / b bit for
    bit, signed zeros included, on the scheme's divisors and random ones."""
    import os
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")
    exe = str(tmp_path / "div_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math",
                    f"-I{here}", "-o", exe, os.path.join(here, "div_check.cpp")], check=True)
    p = subprocess.run([exe], capture_output=True, text=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert " 0 mismatches" in p.stdout
