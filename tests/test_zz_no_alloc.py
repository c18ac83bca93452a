"""The allocation-free entry point cth_de_vjp_noalloc_dev against
cth_de_vjp_dev, through the C ABI: the same kernels on the library's static
scratch instead of a stream-ordered workspace, so the results are identical
bit for bit.  Two calls in a row, the second smaller (it must not see the
first call's records), and a call beyond the scratch's size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32


def _vjp(lib_fn, ctx, shape, pts, powers, grad_d, torch, with_power):
    n = pts.shape[0]
    d = torch.full((n,), float("nan"), device="cuda")
    gp = torch.full((n, 3), float("nan"), device="cuda")
    gpow = torch.full((n,), float("nan"), device="cuda") if with_power else None
    ptr = lambda t: None if t is None else t.data_ptr()
    rc = lib_fn(ctx.handle, C.byref(shape.desc), ptr(pts), ptr(powers), n, ptr(grad_d), ptr(d), ptr(gp), ptr(gpow),
                torch.cuda.current_stream().cuda_stream)
    return rc, d, gp, gpow


def _same(a, b):
    return np.array_equal(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))


def test_vjp_noalloc_matches_vjp(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd.workloads import shell_points
    torch = torch_cuda
    lib = _lib.lib
    cases = []
    # 1: P8 Mandelbulb, 60,000 points (16.3 MB of the 16 MiB scratch), no grad_d
    p1 = np.random.default_rng(11).uniform(-1.3, 1.3, (60000, 3)).astype(F32)
    cases.append((ca.Mandelbulb.classic(15, 2.0, ctx=ctx), torch.from_numpy(p1).cuda(), None, None, False))
    # 2: real-power trig map, 5,000 shell points with per-point powers, an
    # incoming gradient and the power derivative
    p2, w2 = shell_points(5000, seed=12)
    g2 = torch.from_numpy(np.random.default_rng(13).uniform(-2, 2, 5000).astype(F32)).cuda()
    cases.append((ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx), torch.from_numpy(p2).cuda(), torch.from_numpy(w2).cuda(),
                  g2, True))
    for i, (shape, pts, powers, grad_d, with_power) in enumerate(cases):
        rc0, d0, g0, w0 = _vjp(lib.cth_de_vjp_dev, ctx, shape, pts, powers, grad_d, torch, with_power)
        rc1, d1, g1, w1 = _vjp(lib.cth_de_vjp_noalloc_dev, ctx, shape, pts, powers, grad_d, torch, with_power)
        torch.cuda.synchronize()
        assert rc0 == 0 and rc1 == 0, (i, lib.cth_last_error())
        assert _same(d0, d1) and _same(g0, g1), i
        if with_power:
            assert _same(w0, w1), i
        assert torch.isfinite(g1).all(), i
    # beyond the scratch: refused before any launch (the outputs stay NaN)
    big = torch.zeros((61681, 3), device="cuda")
    rc, d, gp, _ = _vjp(lib.cth_de_vjp_noalloc_dev, ctx, cases[0][0], big, None, None, torch, False)
    torch.cuda.synchronize()
    assert rc == _lib.CTH_E_INVALID and b"16 MiB" in lib.cth_last_error()
    assert torch.isnan(d).all() and torch.isnan(gp).all()
