"""CheckProblem on the device (hpcgx_check_problem; the reference's
CheckProblem.cpp:79-148, which main.cpp:150-159 runs on every level).

Generated problems and fleets pass (hpcgx_generate_problem and
hpcgx_generate_fleet already ran the matrix checks), and so do the
reference's own matrices loaded through OptimizeProblem's host path (the
golden 16^3 hierarchy written by the compiled reference).  One altered
value, column, f2c entry or b entry in that hierarchy is found and named by
its natural row; a level without grid dimensions is refused.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hp():
    import torch
    assert torch.cuda.is_available()
    import hpcg_amd
    return hpcg_amd


def golden_levels(g, dims=(16, 16, 16)):
    levels = []
    for l in range(4):
        d = dict(nnz_row=g["L%d_nnz_row" % l].copy(), cols=g["L%d_cols" % l].copy(),
                 vals=g["L%d_vals" % l].copy(), diag_slot=g["L%d_diag_slot" % l].copy(),
                 nx=dims[0] >> l, ny=dims[1] >> l, nz=dims[2] >> l)
        if l < 3:
            d["f2c"] = g["L%d_f2c" % l].copy()
        levels.append(d)
    return levels


def swap_rows(levels, p, q):
    """Level 0 relabelled by the transposition p <-> q: row perm[i] holds
    row i's entries with columns perm[c], in ascending order again."""
    m = levels[0]
    n = m["nnz_row"].size
    perm = np.arange(n)
    perm[p], perm[q] = q, p
    cols = m["cols"].reshape(n, 27).astype(np.int64)
    vals = m["vals"].reshape(n, 27)
    valid = np.arange(27)[None, :] < m["nnz_row"].astype(np.int64)[:, None]
    newc = np.where(valid, perm[cols], n + 1)
    order = np.argsort(newc, axis=1, kind="stable")
    newc = np.take_along_axis(newc, order, axis=1)
    newv = np.take_along_axis(vals, order, axis=1)
    newc = np.where(newc == n + 1, 0, newc)
    inv = np.argsort(perm)                  # new row r holds old row inv[r]
    m["cols"] = np.ascontiguousarray(newc[inv].reshape(-1).astype(np.int32))
    m["vals"] = np.ascontiguousarray(newv[inv].reshape(-1))
    m["nnz_row"] = np.ascontiguousarray(m["nnz_row"][inv])
    rows = np.arange(n)[:, None]
    m["diag_slot"] = np.argmax(newc[inv] == rows, axis=1).astype(m["diag_slot"].dtype)
    m["f2c"] = perm[m["f2c"]].astype(np.int32)
    return levels


@pytest.mark.parametrize("dims", [(16, 16, 16), (8, 16, 24), (64, 48, 32)])
def test_generated_problem_passes(hp, dims):
    say("check generated", dims)
    P = hp.Problem.generate(*dims, nlevels=4)
    b, x, xe = P.vectors()
    P.check_problem(b, x, xe)
    bad = b.clone()
    bad[P.nrow(0) // 3] += 1.0
    with pytest.raises(hp.HpcgxError) as e:
        P.check_problem(bad, x, xe)
    assert e.value.code == hp._lib.ECHECK and "b differs" in str(e.value)
    with pytest.raises(hp.HpcgxError) as e:
        P.check_problem(b, xe, xe)
    assert e.value.code == hp._lib.ECHECK and "x is not 0" in str(e.value)
    P.close()


def test_fleet_passes(hp):
    say("check fleet")
    F = hp.Problem.generate_fleet([(16, 16, 16), (32, 16, 24), (8, 8, 8)], nlevels=4)
    b, x, xe = F.vectors()
    F.check_problem(b, x, xe)
    F.close()


def test_reference_matrices_pass_and_alterations_fail(hp, g16):
    say("check host levels")
    P = hp.Problem.from_reference_levels(golden_levels(g16))
    P.check_problem()
    P.close()
    row = 1234                       # natural row of the fine level
    # the row's fourth stored entry: an off-diagonal -1 of the golden matrix
    cases = []
    lv = golden_levels(g16)
    assert lv[0]["vals"][27 * row + 3] == -1.0
    lv[0]["vals"][27 * row + 3] = -2.0
    cases.append(("value", lv, 0, "a value differs"))
    # two interior rows trade places (P A P^T: the pattern stays symmetric
    # and every row keeps 26 / -1 values, but rows 1234 and 2730 and their
    # neighbours no longer hold their own stencil's columns)
    cases.append(("column", swap_rows(golden_levels(g16), 1234, 2730), 0,
                  "a column differs"))
    lv = golden_levels(g16)
    lv[1]["vals"][27 * 100 + int(lv[1]["diag_slot"][100])] = 25.0
    cases.append(("diagonal", lv, 1, "a value differs"))
    lv = golden_levels(g16)
    lv[0]["f2c"][17] += 1            # an odd-x fine point: injected by no coarse row
    cases.append(("f2c", lv, 0, "f2c differs"))
    for name, levels, level, what in cases:
        say("altered", name)
        P = hp.Problem.from_reference_levels(levels)
        with pytest.raises(hp.HpcgxError) as e:
            P.check_problem()
        msg = str(e.value)
        assert e.value.code == hp._lib.ECHECK, (name, msg)
        assert ("level %d" % level) in msg and what in msg, (name, msg)
        if name == "value":
            assert "natural row %d" % row in msg, msg
        P.close()


def test_unknown_grid_refused(hp, g16):
    lv = golden_levels(g16)
    for d in lv:
        d["nx"] = d["ny"] = d["nz"] = 0
    P = hp.Problem.from_reference_levels(lv)
    with pytest.raises(hp.HpcgxError) as e:
        P.check_problem()
    assert e.value.code == 1 and "unknown" in str(e.value)
    P.close()
