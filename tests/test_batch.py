"""Batched solves: every solve in a batch gives the same bits as when run
alone, and matches the oracle (FP64 bit for bit, FP32C within tolerance).

The batch kernels run the same per-tile and per-boundary-item code as the
single-solve kernels, on device tables instead of kernel arguments, so the
bar is bit-identity (FP64 and FP32C alike) for every mix of order,
resolution, precision, scheme and admission time, with graph replay on or
off and for every tile shape.
"""
import numpy as np
import pytest

import cases
from mcv_sw_amd import capi
from oracle import oracle

FIELDS = ("phi", "phiG", "u", "v", "contraU", "contraV")

# (dof, n, dt, scheme, precision, nsteps)
MIX = [(3, 5, 300.0, "RK4", capi.FP64, 3), (4, 4, 200.0, "RK3_TVD", capi.FP64, 2),
       (3, 6, 250.0, "RK3_TVD", capi.FP32C, 4), (4, 5, 150.0, "RK4", capi.FP32C, 2)]


def _inputs(spec, amp=50.0):
    dof, n = spec[0], spec[1]
    m = cases.mesh(dof, n)
    return m, cases.bumped_jet(m, u0=30.0, h0=6000.0, amp=amp)


def _alone(lib, spec, inp):
    dof, n, dt, sch, prec, ns = spec
    with capi.Solver(dof, n, dt, *inp, scheme=sch, precision=prec, lib=lib) as s:
        s.step(ns)
        return s.sync()


def _same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in FIELDS) and a["mass"] == b["mass"]


def _matches_oracle(spec, inp, got):
    """The batched result against the oracle's FP64 step on the same inputs
    (time_scheme_mod.F90:13-142): FP64 bit for bit on every defined point,
    FP32C within relative L2 1e-6 in h and 1e-5 in u, v."""
    dof, n, dt, sch, prec, ns = spec
    ref = oracle.run(dof, n, 4 if sch == "RK4" else 3, dt, *inp, ns)
    if prec == capi.FP64:
        msk = cases.masks(dof, n)["defined"]
        return all(np.array_equal(got[k][:, msk], ref[k][:, msk]) for k in FIELDS)
    msk = cases.masks(dof, n)["interior"]
    e = {k: cases.rel_l2(got[k][:, msk], ref[k][:, msk]) for k in ("phi", "u", "v")}
    return e["phi"] <= 1e-6 and e["u"] <= 1e-5 and e["v"] <= 1e-5


def _drain(b, slots, chunk=12, limit=100):
    for _ in range(limit):
        left, _ = b.poll()
        if all(left[s] == 0 for s in slots):
            return
        b.run(chunk)
    raise AssertionError("solves did not finish")


def _mixed_batch(lib, graphs, tile):
    inputs = [_inputs(s) for s in MIX]
    alone = [_alone(lib, s, i) for s, i in zip(MIX, inputs)]
    with capi.Batch(6, lib=lib) as b:
        b.set_graphs(graphs)
        b.set_tile(tile)
        slots = [b.add(s[0], s[1], s[2], *i, s[5], scheme=s[3], precision=s[4])
                 for s, i in zip(MIX, inputs)]
        _drain(b, slots)
        _, done = b.poll()
        for s, a, spec, inp in zip(slots, alone, MIX, inputs):
            assert done[s] == spec[5]
            got = b.sync(s)
            assert _same(got, a), f"slot {s} {spec}"
            assert _matches_oracle(spec, inp, got), f"slot {s} {spec}"


def _continuous(lib):
    """Admission while others are mid-step, slot reuse, and idling solves."""
    A, B, C = MIX[0], MIX[2], MIX[3]
    ia, ib, ic = _inputs(A), _inputs(B, 70.0), _inputs(C, 20.0)
    with capi.Batch(2, lib=lib) as b:
        sa = b.add(A[0], A[1], A[2], *ia, A[5], scheme=A[3], precision=A[4])
        b.run(2)  # A is now inside its first step
        sb = b.add(B[0], B[1], B[2], *ib, B[5], scheme=B[3], precision=B[4])
        with pytest.raises(capi.MCVError) as e:  # full
            b.add(C[0], C[1], C[2], *ic, C[5], scheme=C[3], precision=C[4])
        assert e.value.code == capi.MCV_ESTATE
        with pytest.raises(capi.MCVError) as e:  # A is between stages
            b.sync(sa)
        assert e.value.code == capi.MCV_ESTATE
        _drain(b, [sa], chunk=1)
        assert _same(b.sync(sa), _alone(lib, A, ia))
        b.remove(sa)
        sc = b.add(C[0], C[1], C[2], *ic, C[5], scheme=C[3], precision=C[4])
        assert sc == sa  # the freed slot is reused
        _drain(b, [sb, sc], chunk=5)
        b.run(7)  # finished solves idle: further stages change nothing
        assert _same(b.sync(sb), _alone(lib, B, ib))
        assert _same(b.sync(sc), _alone(lib, C, ic))


def _kind_change(lib):
    """The stage kernel is compiled per set of solve kinds in the batch: a
    graph captured for one kind is replaced when another kind joins, and
    again when it leaves."""
    B, D = MIX[2], MIX[1]  # FP32C MCV3 RK3, FP64 MCV4 RK3
    ib, idd = _inputs(B, 40.0), _inputs(D, 60.0)
    nb, nd = 16, 4
    with capi.Batch(2, lib=lib) as b:
        b.set_graphs(True)
        sb = b.add(B[0], B[1], B[2], *ib, nb, scheme=B[3], precision=B[4])
        b.run(12)  # one graph replay with the FP32C MCV3 body alone
        sd = b.add(D[0], D[1], D[2], *idd, nd, scheme=D[3], precision=D[4])
        b.run(12)  # re-captured with both bodies
        _drain(b, [sd], chunk=12)
        assert _same(b.sync(sd), _alone(lib, (*D[:5], nd), idd))
        b.remove(sd)
        _drain(b, [sb], chunk=12)  # back to one body
        assert _same(b.sync(sb), _alone(lib, (*B[:5], nb), ib))


def _grid_change(lib):
    """The stage and boundary kernels take one workgroup per work item; a
    captured graph keeps its grids while the work list fits them and fills at
    least half.  A larger solve joining a replayed batch outgrows the grids
    (re-capture), and its leaving shrinks the list below half (re-capture)."""
    small = (3, 4, 300.0, "RK3_TVD", capi.FP32C, 12)
    big = (3, 16, 100.0, "RK3_TVD", capi.FP32C, 4)
    isml, ibig = _inputs(small, 40.0), _inputs(big, 60.0)
    with capi.Batch(2, lib=lib) as b:
        b.set_graphs(True)
        b.set_tile(0)
        s1 = b.add(small[0], small[1], small[2], *isml, small[5], scheme=small[3],
                   precision=small[4])
        b.run(12)  # the small solve's grids
        s2 = b.add(big[0], big[1], big[2], *ibig, big[5], scheme=big[3], precision=big[4])
        b.run(12)  # grown
        _drain(b, [s2], chunk=12)
        assert _same(b.sync(s2), _alone(lib, big, ibig))
        b.remove(s2)
        _drain(b, [s1], chunk=12)  # shrunk
        _, done = b.poll()
        assert done[s1] == small[5]
        assert _same(b.sync(s1), _alone(lib, small, isml))


def _one_kind(lib, spec, tile):
    """A batch of one kind of solve (one body, one RK scheme) runs the stage
    kernel compiled for that kind alone (for FP32C at the 8x7/6x4 and 10x5
    shapes, its 8-wave build); two such solves give the bits of each run
    alone."""
    inputs = [_inputs(spec, 40.0), _inputs(spec, 70.0)]
    with capi.Batch(2, lib=lib) as b:
        b.set_graphs(True)
        b.set_tile(tile)
        slots = [b.add(spec[0], spec[1], spec[2], *i, spec[5], scheme=spec[3],
                       precision=spec[4]) for i in inputs]
        _drain(b, slots)
        for s, i in zip(slots, inputs):
            assert _same(b.sync(s), _alone(lib, spec, i)), f"slot {s} {spec}"


# ---------------------------------------------------------------------------
# CPU: product sources on the emulation shim
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("graphs", [False, True])
def test_emulated_batch_matches_alone(emu_lib, graphs):
    _mixed_batch(emu_lib, graphs, tile=0)


@pytest.mark.parametrize("tile", [1, 2, 3])
def test_emulated_batch_tile_shapes(emu_lib, tile):
    _mixed_batch(emu_lib, True, tile)


def test_emulated_continuous_batching(emu_lib):
    _continuous(emu_lib)


def test_emulated_batch_kind_change(emu_lib):
    _kind_change(emu_lib)


@pytest.mark.parametrize("tile", [0, 2])
@pytest.mark.parametrize("k", range(len(MIX)))
def test_emulated_batch_one_kind(emu_lib, k, tile):
    _one_kind(emu_lib, MIX[k], tile)


def test_emulated_batch_grid_change(emu_lib):
    _grid_change(emu_lib)


def test_emulated_batch_autotune_keeps_results(emu_lib):
    spec = MIX[1]
    inp = _inputs(spec)
    with capi.Batch(2, lib=emu_lib) as b:
        s = b.add(spec[0], spec[1], spec[2], *inp, spec[5], scheme=spec[3], precision=spec[4])
        tile, ms = b.autotune(reps=1)
        assert 0 <= tile < capi.TILE_VARIANTS and len(ms) == capi.TILE_VARIANTS
        b.run(1)
        with pytest.raises(capi.MCVError) as e:  # inside a step
            b.autotune(reps=1)
        assert e.value.code == capi.MCV_ESTATE
        _drain(b, [s], chunk=1)
        assert _same(b.sync(s), _alone(emu_lib, spec, inp))


def test_emulated_batch_rejects_bad_slots(emu_lib):
    with capi.Batch(1, lib=emu_lib) as b:
        with pytest.raises(capi.MCVError) as e:
            b.remove(0)
        assert e.value.code == capi.MCV_EINVAL
        b.run(3)  # an empty batch runs
        left, done = b.poll()
        assert left[0] == -1 and done[0] == -1


@pytest.mark.parametrize("dof,n,prec", [(3, 6, capi.FP64), (4, 5, capi.FP32C)])
def test_emulated_footprint_bounds_device_bytes(emu_lib, dof, n, prec):
    m, st = _inputs((dof, n))
    with capi.Solver(dof, n, 100.0, m, st, precision=prec, lib=emu_lib) as s:
        used = s.device_bytes
    fp = capi.footprint(dof, n, precision=prec, lib=emu_lib)
    # the bound counts the panel-invariant metric terms on all six panels
    assert used <= fp <= 1.3 * used, (used, fp)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [False, True])
def test_batch_matches_alone(hip_lib, graphs):
    print(f"[batch] graphs={graphs}", flush=True)
    _mixed_batch(hip_lib, graphs, tile=0)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [1, 2, 3])
def test_batch_tile_shapes(hip_lib, tile):
    print(f"[batch] tile={tile}", flush=True)
    _mixed_batch(hip_lib, True, tile)


@pytest.mark.gpu
def test_batch_kind_change(hip_lib):
    print("[batch] kind change with graph replay", flush=True)
    _kind_change(hip_lib)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [0, 2])
@pytest.mark.parametrize("k", range(len(MIX)))
def test_batch_one_kind(hip_lib, k, tile):
    _one_kind(hip_lib, MIX[k], tile)


@pytest.mark.gpu
def test_batch_grid_change(hip_lib):
    _grid_change(hip_lib)


@pytest.mark.gpu
def test_continuous_batching(hip_lib):
    print("[batch] continuous", flush=True)
    _continuous(hip_lib)


@pytest.mark.gpu
def test_batch_ensemble_matches_alone(hip_lib):
    """Twelve MCV3 members on 6x24x24 (RK3, FP32C) plus two FP64 MCV4
    solves, after autotuning."""
    print("[batch] ensemble", flush=True)
    specs = [(3, 24, 300.0, "RK3_TVD", capi.FP32C, 6)] * 12 + \
        [(4, 20, 150.0, "RK4", capi.FP64, 5), (4, 33, 100.0, "RK4", capi.FP64, 4)]
    inputs = [_inputs(s, amp=10.0 + 5.0 * k) for k, s in enumerate(specs)]
    with capi.Batch(len(specs), lib=hip_lib) as b:
        slots = [b.add(s[0], s[1], s[2], *i, s[5], scheme=s[3], precision=s[4])
                 for s, i in zip(specs, inputs)]
        b.autotune(reps=2)
        _drain(b, slots)
        for s, spec, inp in zip(slots, specs, inputs):
            got = b.sync(s)
            assert _same(got, _alone(hip_lib, spec, inp)), f"slot {s}"
            assert _matches_oracle(spec, inp, got), f"slot {s} {spec}"
