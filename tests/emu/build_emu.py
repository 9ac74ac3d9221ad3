"""Build the CPU-emulated product library (test infrastructure).

Compiles mcv_sw_amd/csrc/mcv_hip.hip with g++ against tests/emu/hip/ (a
test-only shim, see its header) so the CPU suite can run the product's
kernels and host orchestration without a GPU.  Two variants:
  libmcv_hip_emu.so   -O2, loaded through ctypes by the parity tests
  asan_driver         -O1 with AddressSanitizer/UBSan: runs every stage kind
                      on a small problem and fails on any out-of-bounds access
  tsan_driver         -O1, each block run by one host thread per GPU thread
                      with real barriers, under ThreadSanitizer: fails on any
                      data race between the threads of a block
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRCS = [os.path.join(ROOT, "mcv_sw_amd", "csrc", f) for f in ("mcv_hip.hip", "mcv_batch.hip", "mcv_probe.hip",
                                                                 "mcv_netcdf.cpp")]
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libmcv_hip_emu.so")
ASAN = os.path.join(OUT, "asan_driver")

COMMON = ["g++", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", f"-I{HERE}",
          "-w"]


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _deps():
    d = os.path.join(ROOT, "mcv_sw_amd", "csrc")
    out = [os.path.join(d, f) for f in os.listdir(d) if f.endswith((".hip", ".hpp", ".cpp"))]
    return out + [os.path.join(HERE, "hip", "hip_runtime.h"), os.path.join(HERE, "asan_driver.cpp"),
                  os.path.join(ROOT, "include", "mcv_hip.h")]


def build_lib(force=False):
    os.makedirs(OUT, exist_ok=True)
    if force or _stale(LIB, _deps()):
        subprocess.run(COMMON + ["-O2", "-fPIC", "-shared", "-x", "c++"] + SRCS + ["-o", LIB],
                       check=True)
    return LIB


TSAN = os.path.join(OUT, "tsan_driver")


def build_tsan(force=False):
    """Threaded blocks (one host thread per GPU thread, real barriers) under
    ThreadSanitizer: intra-block data races in the kernels are reported."""
    os.makedirs(OUT, exist_ok=True)
    if force or _stale(TSAN, _deps()):
        cmd = [c if c != "-std=c++17" else "-std=c++20" for c in COMMON]
        subprocess.run(cmd + ["-O1", "-g", "-DMCV_EMU_THREADS", "-fsanitize=thread", "-pthread",
                              "-x", "c++"] + SRCS + ["-x", "c++",
                              os.path.join(HERE, "asan_driver.cpp"), "-o", TSAN], check=True)
    return TSAN


def build_asan(force=False):
    os.makedirs(OUT, exist_ok=True)
    if force or _stale(ASAN, _deps()):
        subprocess.run(COMMON + ["-O1", "-g", "-fsanitize=address,undefined",
                                 "-fno-sanitize-recover=all", "-x", "c++"] + SRCS + ["-x", "c++",
                                 os.path.join(HERE, "asan_driver.cpp"), "-o", ASAN], check=True)
    return ASAN


if __name__ == "__main__":
    build_lib(force=True)
    if len(sys.argv) > 1 and sys.argv[1] == "asan":
        build_asan(force=True)
