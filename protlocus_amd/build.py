"""Build libnasw_hip.so (gfx950 HIP kernels + the ns_* C-ABI) in-tree.

    python -m protlocus_amd.build        # or protlocus_amd.build.build()

Each source is compiled on its own (in parallel) into protlocus_amd/_build/, then linked
into protlocus_amd/libnasw_hip.so. hipcc cross-compiles gfx950 code objects without a GPU.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnasw_hip.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-ffp-contract=off",
          "-I" + os.path.join(os.path.dirname(HERE), "include")]
# device code: gfx950 only; packed-FP32 VOP3P forms are kept out of all kernels
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src, verbose=False):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip"] + DEVICE + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (src, msgs))
    if msgs and verbose:
        print(msgs, flush=True)
    return obj


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    # C drivers of the reference's calling patterns over this library (the benchmark's drop-in path)
    drv = os.path.join(HERE, "libnsh_drivers.so")
    cmd = ["gcc", "-std=c99", "-O2", "-fPIC", "-shared", "-o", drv + ".tmp", os.path.join(HERE, "drivers", "caller_patterns.c"),
           "-L" + HERE, "-l:libnasw_hip.so", "-Wl,-rpath,$ORIGIN", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stdout + r.stderr)
    os.replace(drv + ".tmp", drv)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
