"""Writes tests/golden/ref_report_16x16x16.txt: the result file the
reference's own driver (compiled from its sources by `make -C oracle ref`
into oracle/_ref/xhpcg) writes for a 16^3 quick-path run on one rank.  The
driver test compares the key list and the deterministic values of
xhpcg_amd's result file with it.

    python tests/golden/make_report_fixture.py
"""
import glob
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "xhpcg")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "ref"], check=True)
    d = tempfile.mkdtemp()
    try:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([exe, "--nx=16", "--ny=16", "--nz=16", "--rt=0"], cwd=d, env=env,
                       check=True, capture_output=True, timeout=600)
        (out,) = glob.glob(os.path.join(d, "HPCG-Benchmark_3.1_*.txt"))
        shutil.copy(out, os.path.join(HERE, "ref_report_16x16x16.txt"))
    finally:
        shutil.rmtree(d)


if __name__ == "__main__":
    main()
