"""Regenerate the golden fixtures in this folder from the compiled reference.

Runs ``oracle/_ref/ref_golden`` (the HPCG 3.1 reference sources compiled
single-threaded by ``make -C oracle ref``) for each grid below and packs its
outputs into ``ref_<nx>x<ny>x<nz>.npz``: inputs, per-kernel outputs, CG
residual trajectories, TestCG / TestSymmetry / TestNorms results and the
flop/byte model parsed from the reference's own report file.

This script only runs where the reference tree exists; the fixtures it writes
are data and are committed.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_golden")
GRIDS = [(16, 16, 16), (8, 16, 24), (32, 32, 32)]
# at 32^3 only scalars and trajectories are kept (the vectors are large)
LITE = {(32, 32, 32)}
LITE_KEEP = ("dims", "cg_", "testcg", "testsymmetry", "testnorms", "dot", "model_")
DT = {"f64": np.float64, "i32": np.int32, "u8": np.uint8}

# keys of the reference report (ReportResults.cpp:331-341) kept as the model
REPORT_KEYS = {
    "Floating Point Operations Summary::Raw DDOT": "model_ddot",
    "Floating Point Operations Summary::Raw WAXPBY": "model_waxpby",
    "Floating Point Operations Summary::Raw SpMV": "model_spmv",
    "Floating Point Operations Summary::Raw MG": "model_mg",
    "Floating Point Operations Summary::Total": "model_total",
    "Floating Point Operations Summary::Total with convergence overhead": "model_total_ref",
    "GB/s Summary::Raw Read B/W": "model_read_bw",
    "GB/s Summary::Raw Write B/W": "model_write_bw",
}


def run(grid):
    nx, ny, nz = grid
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([HARNESS, "golden", str(nx), str(ny), str(nz), d], check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        man = json.load(open(os.path.join(d, "manifest.json")))
        out = {"dims": np.array(man.pop("dims"), dtype=np.int32)}
        for name, (dt, count) in man.items():
            a = np.fromfile(os.path.join(d, name + ".bin"), dtype=DT[dt])
            assert a.size == count, name
            if name.endswith("_cols"):
                # store column offsets from the row index: they compress well
                rows = np.repeat(np.arange(count // 27, dtype=np.int64), 27)
                name = name[:-5] + "_coldelta"
                a = (a.astype(np.int64) - rows).astype(np.int32)
            out[name] = a
        rep = [f for f in os.listdir(d) if f.startswith("HPCG-Benchmark")]
        assert len(rep) == 1, rep
        for line in open(os.path.join(d, rep[0])):
            key, _, val = line.rstrip("\n").partition("=")
            if key in REPORT_KEYS:
                out[REPORT_KEYS[key]] = np.array([float(val)])
    if grid in LITE:
        out = {k: v for k, v in out.items() if k.startswith(LITE_KEEP)}
    path = os.path.join(HERE, "ref_%dx%dx%d.npz" % grid)
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    for g in GRIDS:
        run(g)
