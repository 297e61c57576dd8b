"""hpcgx_dot_noalloc (the dot product whose reduction scratch is a static
device array) against hpcgx_dot, the entry point whose kernels it shares,
through the C ABI: bit-identical results on two calls in a row, the second
with a smaller input, plus the empty and negative lengths."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(n, x, y, out):
    from hpcg_amd import _lib
    from hpcg_amd.problem import _ptr, _stream
    return _lib.lib().hpcgx_dot_noalloc(n, _ptr(x, n), _ptr(y, n), _ptr(out, 1), _stream())


def test_dot_noalloc_matches_dot():
    import torch
    import hpcg_amd
    from hpcg_amd import _lib

    torch.cuda.set_device(0)
    P = hpcg_amd.Problem.generate(16, 16, 16, nlevels=4)
    try:
        rng = np.random.default_rng(11)
        # first call: 2442 first-pass workgroups of 2048 elements, the last
        # one partial; second: a smaller length, 35 workgroups
        for n in (5_000_003, 70_001):
            x = torch.from_numpy(rng.uniform(-1, 1, n)).cuda()
            y = torch.from_numpy(rng.uniform(-1, 1, n)).cuda()
            want = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
            got = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
            P.dot_device(x, y, want, n=n)
            rc = _noalloc(n, x, y, got)
            assert rc == 0, _lib.lib().hpcgx_last_error().decode()
            torch.cuda.synchronize()
            w, g = float(want.item()), float(got.item())
            assert np.float64(w).tobytes() == np.float64(g).tobytes(), (n, w, g)
            # and the sum itself, to the dot's bound against an exact-order sum
            xs, ys = x.cpu().numpy(), y.cpu().numpy()
            exact = float(np.sum(xs * ys, dtype=np.longdouble))
            assert abs(g - exact) <= 1e-14 * float(np.sum(np.abs(xs * ys))), (n, g, exact)
    finally:
        P.close()


def test_dot_noalloc_edges():
    import torch
    from hpcg_amd import _lib

    torch.cuda.set_device(0)
    x = torch.ones(4, dtype=torch.float64, device="cuda")
    out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    assert _noalloc(0, x, x, out) == 0
    torch.cuda.synchronize()
    assert float(out.item()) == 0.0
    from hpcg_amd.problem import _ptr, _stream
    rc = _lib.lib().hpcgx_dot_noalloc(-1, _ptr(x), _ptr(x), _ptr(out), _stream())
    assert rc == 1  # HPCGX_EINVAL
    assert "negative" in _lib.lib().hpcgx_last_error().decode()
    rc = _lib.lib().hpcgx_dot_noalloc(4, _ptr(x), _ptr(x), ctypes.c_void_p(0), _stream())
    assert rc == 1


def test_dot_noalloc_long_vectors():
    """Past 131072 first-pass workgroups (2^28 elements) both entry points
    take the grid-stride reduction: still the same bits, and the sum within
    the dot's bound of a float64 reference."""
    import torch
    import hpcg_amd
    from hpcg_amd import _lib

    torch.cuda.set_device(0)
    P = hpcg_amd.Problem.generate(16, 16, 16, nlevels=4)
    try:
        n = 131072 * 2048 + 1
        g = torch.Generator(device="cuda").manual_seed(5)
        x = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        y = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        want = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
        got = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
        P.dot_device(x, y, want, n=n)
        rc = _noalloc(n, x, y, got)
        assert rc == 0, _lib.lib().hpcgx_last_error().decode()
        torch.cuda.synchronize()
        w, gv = float(want.item()), float(got.item())
        assert np.float64(w).tobytes() == np.float64(gv).tobytes(), (w, gv)
        ref = float(torch.dot(x, y))
        assert abs(gv - ref) <= 2e-14 * float((x * y).abs().sum()), (gv, ref)
        del x, y
    finally:
        P.close()
