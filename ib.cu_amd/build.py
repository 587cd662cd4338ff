"""Builds libibm.so (HIP, gfx950) in-tree.

Every source under csrc/ is compiled with hipcc for gfx950 and linked into
ib.cu_amd/libibm.so next to this file.  Objects go to ib.cu_amd/_build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libibm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++20", "-fPIC",
    f"--offload-arch={ARCH}",
    # Packed FP32 instructions stay off: hipcc must not emit v_pk_{add,mul,fma}_f32.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
    "-Wno-unused-command-line-argument",
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _headers_mtime():
    latest = 0.0
    for d in (CSRC, os.path.join(ROOT, "include")):
        for name in os.listdir(d):
            if name.endswith((".hpp", ".h")):
                latest = max(latest, os.path.getmtime(os.path.join(d, name)))
    return latest


def _compile(src, incremental=False, headers=0.0):
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), headers):
        return obj
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip", *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def device_asm(src, out):
    """gfx950 assembly of one source, with the flags of the library build."""
    cmd = [HIPCC, *COMMON, "--cuda-device-only", "-S", src, "-o", out]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip", *COMMON, "--cuda-device-only", "-S", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc -S failed on {src}:\n{r.stderr}")
    return out


def build(verbose=True, incremental=False):
    """Compiles every source (incremental: only those older than their
    object or any header) and links libibm.so."""
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    headers = _headers_mtime()
    with ThreadPoolExecutor(max_workers=max(jobs, 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental, headers), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_api_check(verbose)
    return LIB


API_SRC = os.path.join(ROOT, "tests", "cpp", "api_check.cpp")
API_BIN = os.path.join(ROOT, "tests", "cpp", "api_check")


def build_api_check(verbose=True):
    """The C++ API program the GPU tests run (links libibm.so)."""
    cmd = [HIPCC, "-O2", "-std=c++20", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), API_SRC, "-o", API_BIN + ".tmp",
           "-L", HERE, "-l:libibm.so", "-Wl,-rpath,$ORIGIN/../../ib.cu_amd", "-Wno-unused-command-line-argument"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {API_SRC}:\n{r.stderr}")
    os.replace(API_BIN + ".tmp", API_BIN)
    if verbose:
        print(f"built {API_BIN}")


if __name__ == "__main__":
    try:
        build(incremental="--incremental" in sys.argv)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
