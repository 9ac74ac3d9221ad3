"""Pins against the reference build's golden vectors (tests/golden).

golden.json holds SHA-256 digests of the reference's own arrays (mesh, step-0
state, final state on every defined point) written by make_golden.py from
oracle/_ref.  Three implementations are checked against them bit for bit:
  - the oracle's C restatement (mesh, ghost tables, trajectories),
  - the product's Fortran host library (mesh, ghost tables, initial states),
  - the product's HIP sources compiled for the CPU by tests/emu (trajectories).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import cases
from mcv_sw_amd import hostlib
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))
FIELDS = ("phi", "phiG", "u", "v", "contraU", "contraV")
MESH_KEYS = ("xP", "yP", "lonP", "latP", "sqrtG", "G", "IG", "A", "IA", "f", "coef", "iref")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def state_digests(st, dof, n):
    msk = cases.masks(dof, n)["defined"]
    return {k: digest(st[k][:, msk]) for k in FIELDS}


@pytest.mark.parametrize("key", sorted(GOLD["meshes"]))
def test_oracle_mesh_matches_reference(key):
    dof, n = map(int, key.split("_"))
    m = oracle.mesh(dof, n)
    got = {k: digest(m[k]) for k in MESH_KEYS}
    bad = [k for k in MESH_KEYS if got[k] != GOLD["meshes"][key][k]]
    assert not bad, f"oracle mesh differs from the reference in {bad}"


@pytest.mark.parametrize("key", sorted(GOLD["meshes"]))
def test_hostlib_mesh_matches_reference(key):
    dof, n = map(int, key.split("_"))
    m = hostlib.mesh(dof, n)
    got = {k: digest(m[k]) for k in MESH_KEYS}
    bad = [k for k in MESH_KEYS if got[k] != GOLD["meshes"][key][k]]
    assert not bad, f"host-library mesh differs from the reference in {bad}"


def _initial(rec):
    """Initial state of a golden case, built by the product's host library."""
    dof, n = rec["dof"], rec["n"]
    m = hostlib.mesh(dof, n)
    kind, arg = rec["init"]
    st = hostlib.williamson(m, arg) if kind == "case" else hostlib.jet(m, **arg)
    return m, st


# full-size cases are too large for the oracle on the CPU; they pin the GPU path
FULL = [k for k in GOLD["cases"] if k.endswith("_full")]
BITWISE_IC = [k for k, r in GOLD["cases"].items()
              if not (r["init"][0] == "case" and r["init"][1] == 6) and k not in FULL]


@pytest.mark.parametrize("name", sorted(BITWISE_IC) + FULL)
def test_hostlib_initial_state_matches_reference(name):
    rec = GOLD["cases"][name]
    m, st = _initial(rec)
    assert state_digests(st, rec["dof"], rec["n"]) == rec["step0"]
    if "phi_s" in rec:
        assert digest(m["phi_s"]) == rec["phi_s"]


@pytest.mark.parametrize("name", FULL)
def test_hostlib_ghost_tables_full_size(name):
    """At 6x384x384 the host library's ghost offsets equal the reference's
    except where the neighbour-panel abscissa atan(sin/tan) differs in its
    last bit between the two math-library code paths (coef = X - x_left
    inherits that one ulp of X)."""
    rec = GOLD["cases"][name]
    m = hostlib.mesh(rec["dof"], rec["n"])
    z = np.load(os.path.join(HERE, "golden", name + "_ghost.npz"))
    assert np.array_equal(m["iref"], z["iref"])
    d = np.abs(m["coef"] - z["coef"])
    assert np.count_nonzero(d) <= 4
    assert d.max() <= np.spacing(np.pi / 4)


@pytest.mark.gpu
@pytest.mark.parametrize("name", FULL)
def test_full_size_solve_matches_reference(name, hip_lib):
    """The single large solve at its full size (MCV4 6x384x384, RK4): the
    handed-over state and the first steps on the GPU equal the reference
    build's arrays bit for bit."""
    from mcv_sw_amd import capi

    print(f"[golden] {name}", flush=True)
    rec = GOLD["cases"][name]
    m, st = _initial(rec)
    # the reference host's own ghost tables, as it would hand them over
    z = np.load(os.path.join(HERE, "golden", name + "_ghost.npz"))
    m["coef"], m["iref"] = z["coef"], z["iref"]
    with capi.Solver(rec["dof"], rec["n"], rec["dt"], m, st, scheme=rec["scheme"],
                     lib=hip_lib) as s:
        assert state_digests(s.sync(integrals=False), rec["dof"], rec["n"]) == rec["step0"]
        s.step(rec["steps"])
        out = s.sync()
    assert state_digests(out, rec["dof"], rec["n"]) == rec["final"]
    assert abs(out["mass"] - rec["mass"]) <= 1e-12 * abs(rec["mass"])


def test_hostlib_case6_matches_reference_to_rounding():
    """Case 6 (Rossby-Haurwitz wave) is written from Williamson et al. (1992)
    in its own evaluation order; it equals the reference's initial state to
    a few rounding errors of each field's magnitude."""
    z = np.load(os.path.join(HERE, "golden", "tiny_mcv4_case6_rk3.npz"))
    m = hostlib.mesh(4, 4)
    st = hostlib.williamson(m, 6)
    msk = cases.masks(4, 4)["defined"]
    for k in FIELDS:
        a, b = st[k][:, msk], z["s0_" + k][:, msk]
        assert np.abs(a - b).max() <= 4 * np.finfo(float).eps * np.abs(b).max(), k


def _ref_initial(name, rec):
    """Step-0 state exactly as the reference had it."""
    if name.startswith("tiny_"):
        z = np.load(os.path.join(HERE, "golden", name + ".npz"))
        m = hostlib.mesh(rec["dof"], rec["n"])
        m["phi_s"] = z["phi_s"]
        st = {k: z["s0_" + k] for k in FIELDS}
        st["zonal"], st["merid"] = z["s0_zonal"], z["s0_merid"]
        return m, st
    assert name in BITWISE_IC
    return _initial(rec)


TRAJ = [k for k in GOLD["cases"] if k.startswith("tiny_") or k in BITWISE_IC]


@pytest.mark.parametrize("name", sorted(TRAJ))
def test_oracle_trajectory_matches_reference(name):
    rec = GOLD["cases"][name]
    m, st = _ref_initial(name, rec)
    out = oracle.run(rec["dof"], rec["n"], 4 if rec["scheme"] == "RK4" else 3, rec["dt"], m, st,
                     rec["steps"])
    assert state_digests(out, rec["dof"], rec["n"]) == rec["final"]
    mass, energy = oracle.diag(rec["dof"], rec["n"], m, out)
    assert abs(mass - rec["mass"]) <= 1e-14 * abs(rec["mass"])
    assert abs(energy - rec["energy"]) <= 1e-14 * abs(rec["energy"])


def test_history_output_leaves_trajectory_unchanged():
    """MCV_SW.F90:62-74 rewrites halo corners and sphere winds at output
    times; the reference trajectory is the same with output every step."""
    a = GOLD["cases"]["mcv3_n8_case5_rk4_history"]["final"]
    b = GOLD["cases"]["mcv3_n8_case5_rk4"]["final"]
    assert a == b


@pytest.mark.parametrize("name", ["tiny_mcv3_case5_rk4", "tiny_mcv4_case6_rk3", "mcv3_n8_case5_rk4",
                                  "mcv4_n12_case2_rk4", "mcv4_n16_jet_rk4"])
def test_emulated_product_matches_reference(name, emu_lib):
    """The product's HIP sources, compiled for the CPU, reproduce the reference."""
    from mcv_sw_amd import capi

    rec = GOLD["cases"][name]
    m, st = _ref_initial(name, rec)
    with capi.Solver(rec["dof"], rec["n"], rec["dt"], m, st, scheme=rec["scheme"],
                     lib=emu_lib) as s:
        s.step(rec["steps"])
        out = s.sync()
    assert state_digests(out, rec["dof"], rec["n"]) == rec["final"]
    assert abs(out["mass"] - rec["mass"]) <= 1e-14 * abs(rec["mass"])


@pytest.mark.parametrize("tile", range(4))
@pytest.mark.parametrize("name", ["tiny_mcv3_case5_rk4", "mcv4_n12_case2_rk4"])
def test_emulated_tile_shapes_match_reference(name, tile, emu_lib):
    """Every stage-kernel tile shape the autotuner may pick is bit-exact."""
    from mcv_sw_amd import capi

    rec = GOLD["cases"][name]
    m, st = _ref_initial(name, rec)
    with capi.Solver(rec["dof"], rec["n"], rec["dt"], m, st, scheme=rec["scheme"],
                     lib=emu_lib) as s:
        s.set_tile(tile)
        s.step(rec["steps"])
        out = s.sync()
    assert state_digests(out, rec["dof"], rec["n"]) == rec["final"]


DAY = sorted(k for k in TRAJ if k.endswith("_day"))


@pytest.mark.gpu
@pytest.mark.parametrize("tile", range(4))
@pytest.mark.parametrize("name", sorted(TRAJ))
def test_gpu_trajectory_matches_reference(name, tile, hip_lib):
    """Every golden trajectory of the reference build (both orders, RK4 and
    RK3_TVD, Williamson cases and seeded jets, up to one simulated day) on
    the MI355X through the C-ABI, for every stage-kernel tile shape: equal to
    the reference's arrays bit for bit, mass within 1e-14 of the reference's
    integral and energy within 1e-13 (time_scheme_mod.F90:13-142,
    diag_mod.F90:10-67)."""
    from mcv_sw_amd import capi

    rec = GOLD["cases"][name]
    print(f"[golden gpu] {name} tile {tile}", flush=True)
    m, st = _ref_initial(name, rec)
    with capi.Solver(rec["dof"], rec["n"], rec["dt"], m, st, scheme=rec["scheme"],
                     lib=hip_lib) as s:
        s.set_tile(tile)
        s.step(rec["steps"])
        out = s.sync()
    assert state_digests(out, rec["dof"], rec["n"]) == rec["final"]
    assert abs(out["mass"] - rec["mass"]) <= 1e-14 * abs(rec["mass"])
    assert abs(out["energy"] - rec["energy"]) <= 1e-13 * abs(rec["energy"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", DAY)
def test_gpu_fp32c_one_day_against_reference(name, hip_lib):
    """The FP32-compensated form over the same simulated day: h within 1e-6
    relative L2 of the reference's FP64 trajectory (the FP64 drop-in run
    beside it equals the reference's digests), u and v within 1e-5, and its
    FP64 mass integral within 1e-9 of the initial mass (the FP32 flux
    arithmetic conserves mass to its own rounding: 3e-11 to 3.5e-10 measured
    over these days)."""
    from mcv_sw_amd import capi

    rec = GOLD["cases"][name]
    dof, n = rec["dof"], rec["n"]
    print(f"[golden gpu fp32c] {name}", flush=True)
    m, st = _ref_initial(name, rec)
    res = {}
    for prec in (capi.FP64, capi.FP32C):
        with capi.Solver(dof, n, rec["dt"], m, st, scheme=rec["scheme"], precision=prec,
                         lib=hip_lib) as s:
            mass0 = s.sync(fields=())["mass"]
            s.step(rec["steps"])
            res[prec] = s.sync()
    assert state_digests(res[capi.FP64], dof, n) == rec["final"]
    msk = cases.masks(dof, n)["interior"]
    err = {k: cases.rel_l2(res[capi.FP32C][k][:, msk], res[capi.FP64][k][:, msk])
           for k in ("phi", "u", "v")}
    drift = abs(res[capi.FP32C]["mass"] - mass0) / abs(mass0)
    print(f"[golden gpu fp32c] {name}: rel L2 h {err['phi']:.3e} u {err['u']:.3e} "
          f"v {err['v']:.3e}, mass drift {drift:.3e}", flush=True)
    assert err["phi"] <= 1e-6 and err["u"] <= 1e-5 and err["v"] <= 1e-5, err
    assert drift <= 1e-9
