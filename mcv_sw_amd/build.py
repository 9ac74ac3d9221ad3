"""In-tree build of the product libraries.

  libmcv_hip.so   HIP kernels + C-ABI for gfx950 (hipcc)

Every HIP source is compiled with -ffp-contract=off (the FP64 path must round
exactly where the reference does) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIP_LIB = os.path.join(HERE, "libmcv_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


HIP_SOURCES = ["mcv_hip.hip", "mcv_batch.hip", "mcv_probe.hip", "mcv_netcdf.cpp"]


def build_hip(verbose=False):
    """Compiles every source to an object in parallel, then links."""
    inc = ["-I", os.path.join(ROOT, "include")]
    with tempfile.TemporaryDirectory(prefix="mcv_build_") as tmpdir:
        objs, procs = [], []
        for f in HIP_SOURCES:
            obj = os.path.join(tmpdir, os.path.splitext(f)[0] + ".o")
            objs.append(obj)
            cmd = [HIPCC] + HIP_FLAGS + ["-c"] + inc + ["-o", obj, os.path.join(CSRC, f)]
            procs.append((f, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                              stderr=subprocess.STDOUT, text=True)))
        failed = []
        for f, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                failed.append(f"{f}:\n{out[-6000:]}")
            elif verbose and out.strip():
                print(out)
        if failed:
            raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
        tmp = HIP_LIB + ".tmp"
        cmd = [HIPCC] + HIP_FLAGS + ["-o", tmp] + objs
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("hipcc link failed:\n" + p.stderr[-6000:])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


FC = os.environ.get("MCV_FC", "/opt/rocm/bin/amdflang")
HOST_DIR = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "libmcv_host.so")
# The reference's own kind semantics (REAL = REAL(8), DOUBLE PRECISION =
# REAL(16)) so the host setup reproduces the reference geometry bit for bit.
FFLAGS = ["-cpp", "-fdefault-real-8", "-O3", "-fPIC"]


def build_host(verbose=False):
    mod_dir = os.path.join(HOST_DIR, "_build")
    os.makedirs(mod_dir, exist_ok=True)
    tmp = HOST_LIB + ".tmp"
    cmd = [FC] + FFLAGS + ["-shared", "-J", mod_dir, "-o", tmp,
                           os.path.join(HOST_DIR, "mcv_host_lib.F90")]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + p.stderr[-6000:])
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


DRIVER = os.path.join(HOST_DIR, "mcv_sw_gpu")


def build_driver(hip_lib=HIP_LIB, out=DRIVER, verbose=False):
    """The Fortran host program (host/mcv_sw_driver.F90) linked against a
    library exporting the C-ABI (the HIP build by default)."""
    lib_dir = os.path.dirname(os.path.abspath(hip_lib))
    # module files and the linked program go to files of this process only, so
    # that concurrent builders (parallel test workers) do not collide
    tmp = f"{out}.tmp{os.getpid()}"
    with tempfile.TemporaryDirectory(prefix="mcv_fmod_") as mod_dir:
        cmd = [FC] + FFLAGS + ["-J", mod_dir, "-o", tmp,
                               os.path.join(HOST_DIR, "mcv_host_lib.F90"),
                               os.path.join(HOST_DIR, "mcv_sw_driver.F90"),
                               os.path.abspath(hip_lib), f"-Wl,-rpath,{lib_dir}"]
        p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + p.stderr[-6000:])
    if verbose and p.stderr.strip():
        print(p.stderr)
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    return [build_hip(verbose), build_host(verbose), build_driver(verbose=verbose)]


if __name__ == "__main__":
    for path in build_all(verbose=True):
        print("built", path)
