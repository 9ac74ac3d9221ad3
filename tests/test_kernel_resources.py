"""Register and scratch budgets of the HIP kernels, from the compiler's own
resource report (hipcc -Rpass-analysis=kernel-resource-usage, gfx950, the
product's flags).  Compiles on the CPU; no GPU needed.

  - no kernel spills registers to scratch memory;
  - the 8-wave builds of the batched stage kernel reach 8 waves per SIMD
    (their extra SGPRs spill to VGPR lanes only);
  - the FP64 MCV4 single-solve stage kernels at the 6x4 and 5x5 cell tiles
    (the shapes the autotuner takes at 6x384^2, whose LDS, 27-28 KB, allows
    5 workgroups per CU) stay within 80 VGPRs and run at least 5 waves per
    SIMD.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from mcv_sw_amd import build

SOURCES = ("mcv_hip.hip", "mcv_batch.hip")


def _field(block, key):
    m = re.search(re.escape(key) + r": (\d+)", block)
    return int(m.group(1)) if m else None


@pytest.fixture(scope="module")
def kernels():
    hipcc = build.HIPCC if os.path.exists(build.HIPCC) else shutil.which("hipcc")
    if hipcc is None:
        pytest.skip("hipcc is not installed")
    out = {}
    with tempfile.TemporaryDirectory(prefix="mcv_res_") as tmp:
        procs = []
        for src in SOURCES:
            cmd = ([hipcc] + [f for f in build.HIP_FLAGS if f not in ("-shared", "-Wall")] +
                   ["-c", "--offload-device-only", "-Rpass-analysis=kernel-resource-usage",
                    "-I", os.path.join(build.ROOT, "include"),
                    "-o", os.path.join(tmp, src + ".o"), os.path.join(build.CSRC, src)])
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.STDOUT, text=True)))
        for src, p in procs:
            text, _ = p.communicate()
            assert p.returncode == 0, f"{src}:\n{text[-3000:]}"
            for block in re.split(r"remark: [^\n]*Function Name: ", text)[1:]:
                name = block.split()[0]
                out[name] = {"vgpr": _field(block, "VGPRs"),
                             "scratch": _field(block, "ScratchSize [bytes/lane]"),
                             "occupancy": _field(block, "Occupancy [waves/SIMD]")}
    assert out, "no resource report parsed"
    return out


def test_no_scratch_spills(kernels):
    bad = {k: v["scratch"] for k, v in kernels.items() if v["scratch"]}
    assert not bad, bad


def test_eight_wave_batched_kernels(kernels):
    eight = {k: v for k, v in kernels.items() if "batch_stage_kernel_8w" in k}
    assert len(eight) == 8
    assert all(v["occupancy"] == 8 for v in eight.values()), eight


def test_fp64_stage_kernels_within_80_vgprs(kernels):
    fp64 = {k: v for k, v in kernels.items()
            if re.match(r"_ZN3mcv12stage_kernelILi4ELi\dELi(6ELi4|5ELi5)E", k)}
    assert len(fp64) == 8 * 2  # stage kinds (tendency + 7) x tile shapes
    assert all(v["vgpr"] <= 80 and v["occupancy"] >= 5 for v in fp64.values()), fp64


def test_div_small_exact_on_its_domain():
    """div_small<D>(t) (mcv_device.hpp) forms t / D as (t * ceil(2^21 / D)) >> 21
    with one 24-bit multiply: exhaustively, for every divisor 1..1024 that is
    not a power of two and every t < 2048, the quotient is exact and the
    product stays below 2^32."""
    import numpy as np

    t = np.arange(2048, dtype=np.uint64)
    for d in range(3, 1025):
        if d & (d - 1) == 0:
            continue
        m = ((1 << 21) + d - 1) // d
        assert m < (1 << 24)
        prod = t * np.uint64(m)
        assert int(prod.max()) < (1 << 32), d
        assert np.array_equal(prod >> np.uint64(21), t // np.uint64(d)), d
