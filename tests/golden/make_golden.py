"""Write the golden vectors of tests/golden from the reference build.

Runs oracle/_ref/mcv{3,4}/ref_driver (the reference's own Fortran modules,
compiled unmodified; see oracle/Makefile) in this container and records:
  golden.json   per case: SHA-256 of the reference's arrays on every defined
                point (interior and halo strips; the halo corners hold fill
                or stale values in the reference and are excluded), the
                integrals of diag_mod.F90 at the first and last step, and
                the reference's mesh digests
  tiny_*.npz    full step-0 and final arrays of the 4x4-cell cases
  *_full_ghost.npz  the reference's ghost tables (coef, iref) of the
                full-size case
  history_*.npz the history records of small runs (the arguments of
                history_write_stat after the output block of MCV_SW.F90:62-70:
                addFillValue, convert_wind_P2SP, calc_VIA) with their initial
                state and cell geometry; golden.json "cells" holds digests of
                the reference's lonC, latC, areaCell per mesh
Run: python tests/golden/make_golden.py [case ...]   (needs `make -C oracle ref`)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref_runner  # noqa: E402

FIELDS = ("phi", "phiG", "u", "v", "contraU", "contraV")
MESH_KEYS = ("xP", "yP", "lonP", "latP", "sqrtG", "G", "IG", "A", "IA", "f", "coef", "iref")


def defined_mask(dof, n):
    h = dof - 1
    npv = n * h + 1
    nh = npv + 2 * h
    m = np.ones((nh, nh), bool)
    for a in (slice(0, h), slice(h + npv, nh)):
        for b in (slice(0, h), slice(h + npv, nh)):
            m[a, b] = False
    return m


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def state_digests(st, dof, n):
    msk = defined_mask(dof, n)
    return {k: digest(st[k][:, msk]) for k in FIELDS}


# (name, dof, n, initial state, scheme, dt, steps, extra)
CASES = [
    ("tiny_mcv3_case5_rk4", 3, 4, ("case", 5), "RK4", 1200.0, 8, {}),
    ("tiny_mcv4_case6_rk3", 4, 4, ("case", 6), "RK3_TVD", 900.0, 8, {}),
    ("mcv3_n8_case6_rk3", 3, 8, ("case", 6), "RK3_TVD", 600.0, 20, {}),
    ("mcv4_n12_case2_rk4", 4, 12, ("case", 2), "RK4", 300.0, 48, {}),
    ("mcv3_n8_case5_rk4_history", 3, 8, ("case", 5), "RK4", 600.0, 10, {"history_every": 1}),
    ("mcv3_n8_case5_rk4", 3, 8, ("case", 5), "RK4", 600.0, 10, {}),
    ("mcv3_n32_case5_rk4_day", 3, 32, ("case", 5), "RK4", 240.0, 360, {}),
    ("mcv3_n32_jet_rk4_day", 3, 32, ("jet", dict(u0=20.0, h0=5960.0, lmax=8, seed=0, rms_h=30.0)),
     "RK4", 240.0, 360, {}),
    ("mcv4_n16_jet_rk4", 4, 16, ("jet", dict(u0=40.0, h0=8000.0, lmax=16, seed=0, rms_h=60.0)),
     "RK4", 300.0, 24, {}),
    # one simulated day of each order on the GPU's parity tests
    ("mcv4_n24_jet_rk4_day", 4, 24, ("jet", dict(u0=40.0, h0=8000.0, lmax=16, seed=0, rms_h=60.0)),
     "RK4", 120.0, 720, {}),
    ("mcv3_n32_jet_rk3_day", 3, 32, ("jet", dict(u0=20.0, h0=5960.0, lmax=12, seed=1, rms_h=20.0)),
     "RK3_TVD", 120.0, 720, {}),
    # the single large solve at its full size (a few steps); its ghost tables
    # are stored too (the reference's host passes its own tables)
    ("mcv4_n384_jet_rk4_full", 4, 384,
     ("jet", dict(u0=40.0, h0=8000.0, lmax=16, seed=0, rms_h=60.0)), "RK4", 10.0, 3, {}),
]

MESHES = [(3, 4), (4, 4), (3, 8), (3, 32), (4, 12), (4, 16), (4, 48), (3, 90)]
CELL_KEYS = ("lonC", "latC", "areaCell")

# (name, dof, n, case_num, scheme, dt, steps, history_every)
HISTORY = [
    ("history_mcv3_n4_case5_rk4", 3, 4, 5, "RK4", 1200.0, 6, 3),
    ("history_mcv4_n4_case6_rk3", 4, 4, 6, "RK3_TVD", 900.0, 6, 3),
]


def write_history(name, dof, n, case_num, scheme, dt, steps, every):
    r = ref_runner.run(dof, n, dt, steps, scheme=scheme, case_num=case_num, history_every=every,
                       dump_steps=tuple(range(every, steps + 1, every)))
    arrs = {"phi_s": r["mesh"]["phi_s"]}
    for k in CELL_KEYS:
        arrs[k] = r["mesh"][k]
    s0 = r["states"][0]
    for k in FIELDS + ("zonal", "merid"):
        arrs["s0_" + k] = s0[k]
    for step in range(every, steps + 1, every):
        st = r["states"][step]
        for k in ("u", "v", "contraU", "contraV", "phi", "zonal", "merid", "uC", "vC", "phiC"):
            arrs[f"r{step}_{k}"] = st[k]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)
    return {"dof": dof, "n": n, "case_num": case_num, "scheme": scheme, "dt": dt,
            "steps": steps, "history_every": every}


def main(only=None):
    """Regenerate every case, or with `only` (case names) just those, keeping
    the other records of golden.json."""
    path = os.path.join(HERE, "golden.json")
    if only:
        with open(path) as fh:
            out = json.load(fh)
        run_cases = [c for c in CASES if c[0] in only]
        meshes = []
    else:
        out = {"cases": {}, "meshes": {}}
        run_cases, meshes = CASES, MESHES
    for dof, n in meshes:
        r = ref_runner.run(dof, n, dt=60.0, nsteps=0, case_num=2)
        out["meshes"][f"{dof}_{n}"] = {k: digest(r["mesh"][k]) for k in MESH_KEYS}
        print("mesh", dof, n, flush=True)
    if not only or "cells" in only:
        out["cells"] = {}
        for dof, n in MESHES:
            r = ref_runner.run(dof, n, dt=60.0, nsteps=0, case_num=2)
            out["cells"][f"{dof}_{n}"] = {k: digest(r["mesh"][k]) for k in CELL_KEYS}
            print("cells", dof, n, flush=True)
    out.setdefault("history", {})
    for h in HISTORY:
        if not only or h[0] in only:
            out["history"][h[0]] = write_history(*h)
            print("history", h[0], flush=True)
    for name, dof, n, init, scheme, dt, steps, extra in run_cases:
        ic = None
        case_num = 2
        if init[0] == "case":
            case_num = init[1]
        else:
            from mcv_sw_amd import hostlib  # the IC generator is product code

            m = hostlib.mesh(dof, n)
            st = hostlib.jet(m, **init[1])
            ic = {"phi": st["phi"], "zonal": st["zonal"], "merid": st["merid"],
                  "phi_s": m["phi_s"]}
        r = ref_runner.run(dof, n, dt, steps, scheme=scheme, case_num=case_num, ic=ic,
                           threads=8, **extra)
        s0, s1 = r["states"][0], r["states"][steps]
        rec = {"dof": dof, "n": n, "init": [init[0], init[1]], "scheme": scheme, "dt": dt,
               "steps": steps, "extra": extra,
               "step0": state_digests(s0, dof, n), "final": state_digests(s1, dof, n),
               "mass0": s0["mass"], "energy0": s0["energy"], "mass": s1["mass"],
               "energy": s1["energy"]}
        if "mesh" in r:
            rec["phi_s"] = digest(r["mesh"]["phi_s"])
        if name.endswith("_full"):
            np.savez_compressed(os.path.join(HERE, name + "_ghost.npz"),
                                coef=r["mesh"]["coef"], iref=r["mesh"]["iref"])
        out["cases"][name] = rec
        if name.startswith("tiny_"):
            arrs = {}
            for k in FIELDS:
                arrs["s0_" + k] = s0[k]
                arrs["s1_" + k] = s1[k]
            arrs["s0_zonal"], arrs["s0_merid"] = s0["zonal"], s0["merid"]
            arrs["phi_s"] = r["mesh"]["phi_s"]
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)
        print("case", name, flush=True)
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(only=sys.argv[1:] or None)
