"""HPCG flop/byte accounting.

* :func:`hpcg_model` -- the reference's own flop and memory-traffic model
  (ReportResults.cpp:75-136), used for the "HPCG GFLOP/s" figure.  nnz is a
  64-bit count here; the reference's int counter wraps above 2^31 nonzeros
  (GenerateProblem_ref.cpp:126,177).
* :func:`canonical_bytes` -- compulsory HBM traffic per kernel (true nnz, no
  ELL padding), the numerator of every roofline fraction this project
  reports.  With a = bytes per stored entry (value + column index) and
  v = bytes per vector element:
      SpMV                      a*m + 2v*n
      symmetric SYMGS           2a*m + 6v*n   (x read+written and r read, per sweep)
      residual + restriction    a*m_inj + v*n_f + 2v*n_c
      prolongation              3v*n_c
      dot                       2v*n          (v*n when x is y)
      WAXPBY                    3v*n
      CG iteration              SpMV + 14v*n + V-cycle
"""


def hpcg_model(nrow, nnz, nsets, ref_iters, opt_iters, pre=1, post=1):
    """ReportResults.cpp:75-136.  nrow[l], nnz[l] per MG level."""
    L = len(nrow)
    fsets = float(nsets)
    fit = fsets * opt_iters
    n0, m0 = float(nrow[0]), float(nnz[0])
    ddot = (3.0 * fit + fsets) * 2.0 * n0
    waxpby = (3.0 * fit + fsets) * 2.0 * n0
    spmv = (fit + fsets) * 2.0 * m0
    mg = 0.0
    for l in range(L - 1):
        m = float(nnz[l])
        mg += pre * fit * 4.0 * m + fit * 2.0 * m + post * fit * 4.0 * m
    mg += fit * 4.0 * float(nnz[L - 1])
    total = ddot + waxpby + spmv + mg
    D, I = 8.0, 4.0
    reads = (3.0 * fit + fsets) * 2.0 * n0 * D * 2.0  # ddot + waxpby reads
    writes = (3.0 * fit + fsets) * D + (3.0 * fit + fsets) * n0 * D
    reads += (fit + fsets) * (m0 * (D + I) + n0 * D)
    writes += (fit + fsets) * n0 * D
    for l in range(L - 1):
        m, n = float(nnz[l]), float(nrow[l])
        reads += pre * fit * (2.0 * m * (D + I) + n * D)
        writes += pre * fit * n * D
        reads += fit * (m * (D + I) + n * D)
        writes += fit * m * D
        reads += post * fit * (2.0 * m * (D + I) + n * D)
        writes += post * fit * m * D
    m, n = float(nnz[L - 1]), float(nrow[L - 1])
    reads += fit * (2.0 * m * (D + I) + n * D)
    writes += fit * n * D
    scale = float(ref_iters) / float(opt_iters)
    return dict(ddot=ddot, waxpby=waxpby, spmv=spmv, mg=mg, total=total,
                total_ref=total * scale, reads=reads, writes=writes,
                reads_ref=reads * scale, writes_ref=writes * scale)


def grid_levels(nx, ny, nz, nlevels=4):
    """Rows, stored nonzeros and nonzeros of the injected rows of each level
    of the 27-point problem, from closed forms: nnz = prod(3 n_a - 2); the
    injected rows (even coordinates) hold prod(3 n_a / 2 - 1)."""
    out = []
    for l in range(nlevels):
        n = nx * ny * nz
        nnz = (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2)
        m_inj = ((3 * (nx // 2) - 1) * (3 * (ny // 2) - 1) * (3 * (nz // 2) - 1)
                 if l + 1 < nlevels else 0)
        out.append(dict(dims=(nx, ny, nz), n=n, nnz=nnz, m_inj=m_inj))
        nx, ny, nz = nx // 2, ny // 2, nz // 2
    return out


def canonical_bytes(levels, a=12, v=8, a_outer=12, v_outer=8):
    """Compulsory bytes per kernel and per CG iteration.  `a`, `v` describe
    the V-cycle's storage; `a_outer`, `v_outer` the outer PCG's SpMV and
    vectors (FP64 in every mode)."""
    L = len(levels)
    out = {}
    vcycle = 0.0
    for l, lv in enumerate(levels):
        n, m = lv["n"], lv["nnz"]
        out["spmv_L%d" % l] = a * m + 2 * v * n
        out["symgs_L%d" % l] = 2 * a * m + 6 * v * n
        if l + 1 < L:
            nc = levels[l + 1]["n"]
            out["restrict_L%d" % l] = a * lv["m_inj"] + v * n + 2 * v * nc
            out["prolong_L%d" % l] = 3 * v * nc
            vcycle += 2 * out["symgs_L%d" % l] + out["restrict_L%d" % l] + out["prolong_L%d" % l]
        else:
            vcycle += out["symgs_L%d" % l]
    n0, m0 = levels[0]["n"], levels[0]["nnz"]
    out["vcycle"] = vcycle
    out["outer_spmv"] = a_outer * m0 + 2 * v_outer * n0
    out["outer_vectors"] = 14 * v_outer * n0
    out["iteration"] = out["outer_spmv"] + out["outer_vectors"] + vcycle
    return out
