"""The gfx950 code this repository builds makes no device function calls:
every row helper is inlined into its kernel.  When a helper was left as a
function of its own, the block PCG's colour step made one call per row and
lost a third of its speed with the same results, which no parity test can
see.  This test disassembles the build's own code objects (hpcg_amd/_build,
written by `python -m hpcg_amd.build`) and looks for s_swappc, the call
instruction.  CPU only; skipped where the objects or the ROCm LLVM tools are
absent."""
import glob
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "hpcg_amd", "_build")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def device_disassembly(obj, tmp):
    """AMD assembly of the gfx950 code object embedded in `obj`, or None
    where the object holds no device code."""
    fatbin = os.path.join(tmp, "fatbin")
    code = os.path.join(tmp, "code")
    r = subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section=.hip_fatbin=" + fatbin,
                        obj, os.path.join(tmp, "copy.o")], capture_output=True, text=True)
    if r.returncode != 0 or not os.path.exists(fatbin):
        return None
    subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o",
                    "--input=" + fatbin, "--targets=" + TARGET, "--output=" + code], check=True)
    return subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", "--mcpu=gfx950", code],
                          capture_output=True, text=True, check=True).stdout


def test_no_device_function_calls():
    objs = sorted(glob.glob(os.path.join(BUILD, "*.o")))
    tools = all(os.path.exists(os.path.join(LLVM, t))
                for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-objdump"))
    if not objs or not tools:
        pytest.skip("no build objects or no ROCm LLVM tools")
    checked = 0
    for obj in objs:
        with tempfile.TemporaryDirectory() as tmp:
            dis = device_disassembly(obj, tmp)
        if dis is None:
            continue
        checked += 1
        calls = [l for l in dis.splitlines() if "s_swappc" in l]
        assert not calls, "%s: %d device function calls" % (os.path.basename(obj), len(calls))
    assert checked >= 5
