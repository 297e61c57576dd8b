"""Build libhpcg_amd.so, the C++ surface and the driver, in-tree.

* Each .hip file under hpcg_amd/csrc/ is compiled with hipcc for gfx950 to
  an object in hpcg_amd/_build/ (in parallel), then linked into
  hpcg_amd/libhpcg_amd.so.
* hpcg_amd/host/*.cpp (the reference's optimisable C++ surface over the C
  ABI) is compiled with g++ into hpcg_amd/libhpcg_host.a.
* hpcg_amd/host/driver/*.cpp is linked with both into hpcg_amd/xhpcg_amd,
  the benchmark driver.
Objects are rebuilt when their source or any header is newer.

    python -m hpcg_amd.build [--force]
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhpcg_amd.so")
HOST = os.path.join(PKG, "host")
HOST_LIB = os.path.join(PKG, "libhpcg_host.a")
DRIVER = os.path.join(PKG, "xhpcg_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ROCM = "/opt/rocm"
ARCH = "gfx950"

# -ffp-contract=off: no fused multiply-add, so row sums round like the
#   reference built for x86-64 (separate multiply and add).
# -packed-fp32-ops: keep the compiler from emitting packed FP32 arithmetic.
# -mcode-object-version=5: loadable by torch's bundled HIP runtime too.
FLAGS = [
    "--offload-arch=" + ARCH, "-mcode-object-version=5", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


# host C++: plain g++ against the HIP runtime headers.  HPCG_CONTIGUOUS_ARRAYS:
# the driver's generator keeps each matrix array in one block, and
# DeleteMatrix frees it that way.
HOST_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
    "-D__HIP_PLATFORM_AMD__", "-DHPCG_CONTIGUOUS_ARRAYS", "-I" + os.path.join(ROCM, "include"), "-I" + os.path.join(ROOT, "include"),
]


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, obj):
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the GPU-only target feature: harmless
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, err))
    return err.strip()


def build(force=False, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    jobs = []
    objs = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s)[:-4] + ".o")
        objs.append(o)
        if force or _stale(o, [s] + headers):
            jobs.append((s, o))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    if jobs:
        with cf.ThreadPoolExecutor(workers) as ex:
            futs = {ex.submit(_compile, s, o): s for s, o in jobs}
            for f in cf.as_completed(futs):
                msg = f.result()
                if verbose:
                    print("compiled", os.path.relpath(futs[f], ROOT), flush=True)
                    if msg:
                        print(msg, flush=True)
    if force or jobs or _stale(LIB, objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-mcode-object-version=5", "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr)
        os.replace(tmp, LIB)
        if verbose:
            print("linked", os.path.relpath(LIB, ROOT), flush=True)
    _build_host(force, verbose)
    return LIB


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s" % (what, r.stderr))
    return r.stderr.strip()


def _build_host(force, verbose):
    """libhpcg_host.a from host/*.cpp, then xhpcg_amd from host/driver/*.cpp."""
    hobj = os.path.join(BUILD, "host")
    os.makedirs(os.path.join(hobj, "driver"), exist_ok=True)
    headers = (glob.glob(os.path.join(HOST, "*.hpp")) + glob.glob(os.path.join(HOST, "driver", "*.hpp"))
               + glob.glob(os.path.join(ROOT, "include", "*.h")))
    groups = {}
    jobs = []
    for sub in ("", "driver"):
        objs = []
        for src in sorted(glob.glob(os.path.join(HOST, sub, "*.cpp"))):
            obj = os.path.join(hobj, sub, os.path.basename(src)[:-4] + ".o")
            objs.append(obj)
            if force or _stale(obj, [src] + headers):
                jobs.append((src, obj))
        groups[sub] = objs
    if jobs:
        workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
        with cf.ThreadPoolExecutor(workers) as ex:
            futs = {ex.submit(_run, [CXX] + HOST_FLAGS + ["-c", s, "-o", o], "g++ " + s): s for s, o in jobs}
            for f in cf.as_completed(futs):
                msg = f.result()
                if verbose:
                    print("compiled", os.path.relpath(futs[f], ROOT), flush=True)
                    if msg:
                        print(msg, flush=True)
    if force or jobs or _stale(HOST_LIB, groups[""]):
        tmp = HOST_LIB + ".tmp"
        if os.path.exists(tmp):
            os.remove(tmp)
        _run(["ar", "rcs", tmp] + groups[""], "ar")
        os.replace(tmp, HOST_LIB)
        if verbose:
            print("archived", os.path.relpath(HOST_LIB, ROOT), flush=True)
    if force or jobs or _stale(DRIVER, groups["driver"] + [HOST_LIB, LIB]):
        tmp = DRIVER + ".tmp"
        _run([CXX, "-fopenmp", "-o", tmp] + groups["driver"] +
             [HOST_LIB, "-L" + PKG, "-lhpcg_amd", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
              "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")], "link xhpcg_amd")
        os.replace(tmp, DRIVER)
        if verbose:
            print("linked", os.path.relpath(DRIVER, ROOT), flush=True)


if __name__ == "__main__":
    build(force="--force" in sys.argv)
