"""Run the reference build (oracle/_ref/mcv{3,4}/ref_driver) and read its output.

TEST INFRASTRUCTURE ONLY: used to write golden vectors (tests/golden) and to
time the reference CPU path.  The binary is compiled from the reference's
own sources by `make -C oracle ref`; this module never reads the reference
sources themselves.
"""
import json
import os
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def binary(dof):
    return os.path.join(_HERE, "_ref", f"mcv{dof}", "ref_driver")


def available(dof=3):
    return os.path.exists(binary(dof))


def dims(dof, n):
    h = dof - 1
    npv = n * h + 1
    return h, npv, npv + 2 * h


def _fmt_deg(n):
    d = 90.0 / n
    s = repr(d)
    assert float(s) * n == 90.0, f"90/{n} is not exact in decimal: {s}"
    return s


def _unlimited_stack():
    # the reference keeps whole-field automatic arrays on the stack
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run(dof, n, dt, nsteps, scheme="RK4", case_num=2, ic=None, dump_steps=(),
        history_every=0, threads=1, workdir=None, dump_mesh=True, dump_states=True,
        timeout=None):
    """Run the reference for nsteps; returns dict(mesh, ghost, states, seconds).

    ic: optional dict with 'phi', 'zonal', 'merid', 'phi_s' (each (6, nh, nh))
    that replaces the reference's test case (the reference's test_case_mod
    tail then derives phiG, u, v, contraU, contraV from it).
    """
    h, npv, nh = dims(dof, n)
    own = workdir is None
    wd = tempfile.mkdtemp(prefix="mcvref_") if own else workdir
    try:
        with open(os.path.join(wd, "namelist.input"), "w") as fh:
            fh.write("&time_settings\nrun_days=0\nrun_hours=0\nrun_minutes=0\n"
                     f"run_seconds={int(round(max(nsteps, 1) * dt))}\ndt={dt!r}\n"
                     "history_interval=86400\n"
                     f"integral_scheme='{scheme}'\n/\n"
                     f"&case_select\ncase_num={case_num}\n/\n"
                     f"&domain\ndx={_fmt_deg(n)}\ndy={_fmt_deg(n)}\nxhalo=1\nyhalo=1\n/\n")
        icf = ""
        if ic is not None:
            icf = os.path.join(wd, "ic.bin")
            with open(icf, "wb") as fh:
                for k in ("phi", "zonal", "merid", "phi_s"):
                    np.ascontiguousarray(ic[k], dtype="<f8").tofile(fh)
        ds = list(dump_steps) + [-1] * (64 - len(dump_steps))
        with open(os.path.join(wd, "driver.nml"), "w") as fh:
            fh.write(f"&driver\nnsteps_drv={nsteps}\nic_file='{icf}'\nout_dir='{wd}'\n"
                     f"dump_steps={','.join(str(x) for x in ds)}\n"
                     f"history_every={history_every}\n"
                     f"dump_mesh={'.true.' if dump_mesh else '.false.'}\n"
                     f"dump_states={'.true.' if dump_states else '.false.'}\n/\n")
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        p = subprocess.run([binary(dof)], cwd=wd, env=env, capture_output=True, text=True,
                           timeout=timeout, preexec_fn=_unlimited_stack)
        if p.returncode != 0:
            raise RuntimeError(f"ref_driver failed ({p.returncode}): {p.stderr[-2000:]}")
        timing = None
        for line in p.stdout.splitlines():
            if line.strip().startswith("{"):
                timing = json.loads(line.replace("E", "e"))
        out = {"seconds": timing["ref_seconds"] if timing else None, "states": {}}
        s = (6, nh, nh)
        if dump_mesh:
            a = np.fromfile(os.path.join(wd, "mesh.bin"), dtype="<f8")
            np_ = 6 * nh * nh
            names = ("xP", "yP", "lonP", "latP", "sqrtG")
            m = {}
            off = 0
            for k in names:
                m[k] = a[off:off + np_].reshape(s); off += np_
            for k in ("G", "IG", "A", "IA"):
                m[k] = a[off:off + 4 * np_].reshape(s + (2, 2)); off += 4 * np_
            for k in ("f", "phi_s"):
                m[k] = a[off:off + np_].reshape(s); off += np_
            g = os.path.join(wd, "ghost.bin")
            m["coef"] = np.fromfile(g, dtype="<f8", count=npv * h).reshape(h, npv)
            m["iref"] = np.fromfile(g, dtype="<i4", offset=npv * h * 8).reshape(h, npv)
            c = np.fromfile(os.path.join(wd, "cells.bin"), dtype="<f8").reshape(
                (3, 6, n + 2, n + 2))
            m["lonC"], m["latC"], m["areaCell"] = c[0], c[1], c[2]
            out["mesh"] = m
        for fn in sorted(os.listdir(wd)):
            if fn.startswith("state_") and fn.endswith(".bin"):
                step = int(fn[6:13])
                a = np.fromfile(os.path.join(wd, fn), dtype="<f8")
                flds = a[:-2].reshape((8,) + s)
                st = dict(zip(("phi", "phiG", "u", "v", "contraU", "contraV", "zonal", "merid"),
                              flds))
                st["mass"], st["energy"] = float(a[-2]), float(a[-1])
                vf = os.path.join(wd, f"via_{step:07d}.bin")
                if os.path.exists(vf):
                    c = np.fromfile(vf, dtype="<f8").reshape((3, 6, n + 2, n + 2))
                    st["uC"], st["vC"], st["phiC"] = c[0], c[1], c[2]
                out["states"][step] = st
        return out
    finally:
        if own:
            import shutil
            shutil.rmtree(wd, ignore_errors=True)
