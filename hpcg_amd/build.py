"""Build libhpcg_amd.so in-tree with hipcc for gfx950.

Each .hip file under hpcg_amd/csrc/ is compiled to an object in
hpcg_amd/_build/ (in parallel), then linked into hpcg_amd/libhpcg_amd.so.
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
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
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
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv)
