"""FillRandomVector's private generator (hpcg_amd::reference_rand in
hpcg_amd/host/Vector.hpp) against the C library's own rand() from its default
seed, on the CPU: the same sequence, and a fresh sequence each process."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 100000

PROGRAM = r"""
#include <cstdio>
#include "Vector.hpp"
int main() {
  for (int i = 0; i < %d; ++i) std::printf("%%d\n", hpcg_amd::reference_rand());
  Vector v;
  InitializeVector(v, 3);
  FillRandomVector(v);
  std::printf("%%.17g %%.17g %%.17g\n", v.values[0], v.values[1], v.values[2]);
}
""" % N


@pytest.fixture(scope="module")
def program_output(tmp_path_factory):
    d = tmp_path_factory.mktemp("rand")
    src = d / "rr.cpp"
    src.write_text(PROGRAM)
    exe = d / "rr"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "hpcg_amd", "host"),
                    "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")


def test_reference_rand_is_glibc_default_sequence(program_output):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    want = [libc.rand() for _ in range(N + 3)]
    got = [int(v) for v in program_output[:N]]
    assert got == want[:N]
    # FillRandomVector continues the same sequence: rand()/RAND_MAX + 1
    vals = [float(v) for v in program_output[N].split()]
    assert vals == [w / 2147483647.0 + 1.0 for w in want[N:N + 3]]
