"""Build libassist.so in-tree: HIP kernels for gfx950 plus the C/C++ host library.

    python assist_amd/build.py            # build (always recompiles)
    python assist_amd/build.py --check    # build only if sources are newer

(It imports nothing from the package, whose import needs the library this builds.)

The kernels are compiled with hipcc for gfx950 only, with packed FP32
instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops); the
host sources are compiled with the system C++ compiler against the ROCm headers.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libassist.so")
OBJ = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("ASSIST_AMD_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "mlp.hip"]
CPP_SOURCES = ["spk.cpp", "ascii.cpp", "assist_api.cpp", "batch.cpp"]
HEADERS = ["device_types.h", "force_model.h", "ias15_tables.h", "kernels.h", "internal.h"]

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CPP_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
             "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def sources():
    return [os.path.join(CSRC, f) for f in HIP_SOURCES + CPP_SOURCES + HEADERS] + [
        os.path.join(ROOT, "include", "assist.h"), os.path.join(ROOT, "include", "ascii_ephem.h")]


def up_to_date():
    if not os.path.exists(LIB):
        return False
    t = os.path.getmtime(LIB)
    return all(os.path.getmtime(s) <= t for s in sources())


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for f in HIP_SOURCES:
        o = os.path.join(OBJ, f + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    for f in CPP_SOURCES:
        o = os.path.join(OBJ, f + ".o")
        jobs.append(["g++"] + CPP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    with ThreadPoolExecutor(max_workers=4) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(OBJ, f + ".o") for f in HIP_SOURCES + CPP_SOURCES]
    tmp = LIB + ".tmp%d" % os.getpid()
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    if "--check" in sys.argv and up_to_date():
        print(LIB, "(up to date)")
    else:
        print(build(verbose=True))
