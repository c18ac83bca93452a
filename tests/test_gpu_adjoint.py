"""Adjoint (reverse-sweep) gradients and normals against the FP64 dual-number
oracle, central differences and the 6-tap FD baseline."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
F32 = np.float32


def gpu(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = (a * b).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
    return np.arccos(np.clip(c, -1, 1))


def test_trig_vjp_matches_dual_oracle(ctx, torch_cuda):
    """Shell-normals inputs: shell 0.9 < |p| < 1.3 (seed 1), power U[6, 10], 15 iterations."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import shell_points
    torch = torch_cuda
    pts, powers = shell_points(50000, seed=1)
    shape = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    d, gp, gpow = dev.de_vjp(shape, gpu(torch, pts), gpu(torch, powers), with_power=True)
    torch.cuda.synchronize()
    d, gp, gpow = d.cpu().numpy(), gp.cpu().numpy(), gpow.cpu().numpy()
    d64, g64, it64 = O.de_grad_f64(O.mandelbulb_trig(15, 2.0), pts, powers)
    d_o, it_o, _, _ = O.de_points(O.mandelbulb_trig(15, 2.0), pts, powers)
    # the adjoint's forward is the parity forward kernel's arithmetic
    d_fwd = dev.de_points(shape, gpu(torch, pts), gpu(torch, powers)).cpu().numpy()
    assert (d.view(np.uint32) == d_fwd.view(np.uint32)).all()
    # forward-mode (tangent) gradient: same quantity, other evaluation order
    dt, gt, gtp = dev.de_grad(shape, gpu(torch, pts), gpu(torch, powers), with_power=True)
    torch.cuda.synchronize()
    dt, gt, gtp = dt.cpu().numpy(), gt.cpu().numpy(), gtp.cpu().numpy()
    assert (dt.view(np.uint32) == d_fwd.view(np.uint32)).all()
    same = it_o == it64
    ang = angle(gp[same], g64[same, :3])
    relp = np.abs(gpow[same] - g64[same, 3]) / np.maximum(np.abs(g64[same, 3]), 1e-12)
    print(f"trig vjp: iters f32==f64 {same.mean():.4f}, angle p50 {np.median(ang):.2e} p99 {np.percentile(ang, 99):.2e} "
          f"max {ang.max():.2e}; dpower rel p50 {np.median(relp):.2e} p99 {np.percentile(relp, 99):.2e}")
    angt = angle(gt[same], g64[same, :3])
    relt = np.abs(gtp[same] - g64[same, 3]) / np.maximum(np.abs(g64[same, 3]), 1e-12)
    print(f"trig tangent: angle p50 {np.median(angt):.2e} p99 {np.percentile(angt, 99):.2e}; "
          f"dpower rel p50 {np.median(relt):.2e}")
    for a, rp in ((ang, relp), (angt, relt)):
        assert np.median(a) < 1e-5
        assert np.mean(a <= 1e-3) > 0.98
        assert np.median(rp) < 1e-4


@pytest.mark.parametrize("P", [8, 5])
def test_integer_power_normals_match_dual_oracle(ctx, torch_cuda, P):
    """Mandelbulb<P> (P8 polynomial for P = 8, trig map otherwise): fused
    normals agree with the FP64 dual gradient direction."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import shell_points
    torch = torch_cuda
    pts, _ = shell_points(50000, seed=2, r_lo=0.6, r_hi=1.25)
    shape = ca.Mandelbulb(15, 2.0, P, ctx=ctx)
    d, nrm = dev.normals(shape, gpu(torch, pts))
    _, gp = dev.de_vjp(shape, gpu(torch, pts))
    torch.cuda.synchronize()
    nrm, gp = nrm.cpu().numpy(), gp.cpu().numpy()
    _, g64, it64 = O.de_grad_f64(O.mandelbulb(15, 2.0, P), pts)
    _, it_o, _, _ = O.de_points(O.mandelbulb(15, 2.0, P), pts)
    same = it_o == it64
    ang = angle(nrm[same], g64[same, :3])
    angv = angle(gp[same], g64[same, :3])
    _, nfd = dev.normals_fd(shape, gpu(torch, pts), 1e-3)
    angf = angle(nfd.cpu().numpy()[same], g64[same, :3])
    print(f"P={P} normals (tangent): angle p50 {np.median(ang):.2e} p99 {np.percentile(ang, 99):.2e}; "
          f"reverse sweep p50 {np.median(angv):.2e} p99 {np.percentile(angv, 99):.2e}; "
          f"6-tap FD h=1e-3 p50 {np.median(angf):.2e} p99 {np.percentile(angf, 99):.2e}")
    assert np.abs(np.linalg.norm(nrm, axis=1) - 1).max() < 1e-5
    for a in (ang, angv):
        assert np.median(a) < 1e-5
        assert np.mean(a <= 1e-3) > 0.98
    # the derivative kernels are more accurate than the difference stencil
    assert np.median(ang) < np.median(angf)


def test_vjp_is_linear_in_grad_d(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    pts = np.random.default_rng(3).uniform(-1.2, 1.2, (4096, 3)).astype(F32)
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    g = torch.full((4096,), 2.0, device="cuda")
    _, g1 = dev.de_vjp(shape, gpu(torch, pts))
    _, g2 = dev.de_vjp(shape, gpu(torch, pts), grad_d=g)
    torch.cuda.synchronize()
    fin = torch.isfinite(g1).all(1)
    assert torch.equal(g2[fin], 2 * g1[fin])


def test_zero_iteration_gradient_closed_form(ctx, torch_cuda):
    """|p| > bailout: d = 0.5 r ln r, grad = 0.5 (ln r + 1) p / r."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    v = np.random.default_rng(4).normal(size=(5000, 3))
    pts = (v / np.linalg.norm(v, axis=1, keepdims=True) * np.random.default_rng(5).uniform(2.1, 4, (5000, 1)))
    pts = pts.astype(F32)
    _, gp = dev.de_vjp(ca.Mandelbulb.classic(15, 2.0, ctx=ctx), gpu(torch, pts))
    torch.cuda.synchronize()
    r = np.linalg.norm(pts.astype(np.float64), axis=1, keepdims=True)
    expect = 0.5 * (np.log(r) + 1) * pts / r
    assert np.allclose(gp.cpu().numpy(), expect, rtol=2e-6, atol=1e-7)


def test_fd_normals_match_oracle_stencil(ctx, torch_cuda):
    """The 6-tap baseline reproduces the reference stencil: differences of
    parity distances at p +- h e_a (src/mesh/buffer.rs:256-266)."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import shell_points
    torch = torch_cuda
    pts, powers = shell_points(20000, seed=6)
    h = 1e-3
    for shape, oshape, pw in ((ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx), O.mandelbulb_trig(15, 2.0), powers),
                              (ca.Mandelbulb.classic(15, 2.0, ctx=ctx), O.mandelbulb(15, 2.0), None)):
        grad = torch.empty((len(pts), 3), device="cuda")
        d, nrm = dev.normals_fd(shape, gpu(torch, pts), h, None if pw is None else gpu(torch, pw), grad=grad)
        torch.cuda.synchronize()
        diff_o = O.fd_diff(oshape, pts, h, pw)
        diff = grad.cpu().numpy() * np.float32(2 * h)
        ang = angle(nrm.cpu().numpy(), diff_o)
        print(f"fd taps: normal angle to the oracle's stencil p50 {np.median(ang):.2e} "
              f"p99 {np.percentile(ang, 99):.2e}")
        if pw is None:  # P8: parity taps, so the stencil reproduces the oracle's
            assert np.percentile(ang, 99) < 1e-6
        assert np.mean(ang < 1e-3) > 0.98


@pytest.mark.parametrize("sweep", ["forward", "reverse"])
def test_autograd_function(ctx, torch_cuda, sweep):
    """The autograd backward is the chosen kernel's vector-Jacobian product,
    bit for bit, and the two sweeps agree with each other and with the FP64
    dual-number gradient on well-conditioned points."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    pts_np = np.random.default_rng(7).uniform(-1.2, 1.2, (256, 3)).astype(F32)
    pts = torch.tensor(pts_np, device="cuda", requires_grad=True)
    powers = torch.full((256,), 7.5, device="cuda", requires_grad=True)
    shape = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    d = dev.distance(shape, pts, powers, sweep=sweep)
    (d * 3.0).sum().backward()
    fn = dev.de_grad if sweep == "forward" else dev.de_vjp
    _, gp, gpow = fn(shape, pts.detach(), powers.detach(), torch.full((256,), 3.0, device="cuda"), with_power=True)
    torch.cuda.synchronize()
    assert torch.equal(pts.grad, gp) and torch.equal(powers.grad, gpow)
    _, g64, it64 = O.de_grad_f64(O.mandelbulb_trig(15, 2.0), pts_np, np.full(256, 7.5, F32))
    it_o = O.de_points(O.mandelbulb_trig(15, 2.0), pts_np, np.full(256, 7.5, F32))[1]
    ok = (it_o == it64) & np.isfinite(g64).all(1)
    ang = angle(gp.cpu().numpy()[ok], 3.0 * g64[ok, :3])
    assert np.median(ang) < 1e-5 and np.mean(ang <= 1e-3) > 0.97
    with pytest.raises(ValueError):
        dev.distance(shape, pts, powers, sweep="sideways")


@pytest.mark.parametrize("kind", ["p8", "trig"])
def test_vjp_split_z0_and_c(ctx, torch_cuda, kind):
    """The VJP with the point's two roles apart (the orbit's start z0 and the
    constant c of every body): the parts sum, in f32, to cth_de_vjp_dev's
    grad_p, and each part agrees with the FP64 dual-number oracle's to the
    bar of the VJP itself.  c does not reach the distance when no body runs
    or when the only body is at the iteration cap: its part is exactly 0."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import shell_points
    torch = torch_cuda
    if kind == "p8":
        pts, _ = shell_points(30000, seed=2, r_lo=0.6, r_hi=1.25)
        powers = None
        shape, orc = ca.Mandelbulb.classic(15, 2.0, ctx=ctx), O.mandelbulb(15, 2.0)
    else:
        pts, powers = shell_points(30000, seed=3)
        shape, orc = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx), O.mandelbulb_trig(15, 2.0)
    gpts = gpu(torch, pts)
    gpw = None if powers is None else gpu(torch, powers)
    gd = gpu(torch, np.random.default_rng(32).uniform(0.5, 2.0, len(pts)).astype(F32))
    d, gz, gc, gpow = dev.de_vjp_split(shape, gpts, gpw, gd, with_power=True)
    d1, gp, gpow1 = dev.de_vjp(shape, gpts, gpw, gd, with_power=True)
    torch.cuda.synchronize()
    assert torch.equal(d.view(torch.int32), d1.view(torch.int32))
    assert torch.equal(gpow.view(torch.int32), gpow1.view(torch.int32))
    s = (gz + gc).cpu().numpy()
    gp = gp.cpu().numpy()
    assert ((s == gp) | (np.isnan(s) & np.isnan(gp))).all()
    gz, gc, g = gz.cpu().numpy(), gc.cpu().numpy(), gd.cpu().numpy().astype(np.float64)
    _, g64, it64 = O.de_grad_split_f64(orc, pts, powers)
    it_o = O.de_points(orc, pts, powers)[1]
    ok = (it_o == it64) & (it_o >= 1) & np.isfinite(g64).all(1)
    for part, ref in ((gz, g64[:, :3]), (gc, g64[:, 4:])):
        # an angle that is not defined (a zero or non-finite vector on either
        # side) counts as a miss
        with np.errstate(invalid="ignore"):
            ang = np.nan_to_num(angle(part[ok], ref[ok]), nan=np.pi)
            rel = np.nan_to_num(np.linalg.norm(part[ok] - g[ok, None] * ref[ok], axis=1)
                                / np.linalg.norm(g[ok, None] * ref[ok], axis=1), nan=np.inf)
        print(f"{kind}: {int((~np.isfinite(part[ok]).all(1)).sum())} of {int(ok.sum())} not finite; angle p50 "
              f"{np.median(ang):.2e} p98 {np.percentile(ang, 98):.2e}; magnitude rel p50 {np.median(rel):.2e}")
        assert np.median(ang) < 1e-5 and np.mean(ang <= 1e-3) > 0.98
        assert np.median(rel) < 1e-4
    # c does not reach the distance: zero bodies, or one body at the cap
    far = np.random.default_rng(33).normal(size=(4000, 3))
    far = (far / np.linalg.norm(far, axis=1, keepdims=True) * 3.0).astype(F32)
    _, gzf, gcf = dev.de_vjp_split(shape, gpu(torch, far), None if powers is None else gpu(torch, powers[:4000]))
    assert (gcf.cpu().numpy() == 0).all() and (gzf.cpu().numpy() != 0).any(axis=1).all()
    one = ca.Mandelbulb.classic(1, 2.0, ctx=ctx) if kind == "p8" else ca.MandelbulbTrig(1, 2.0, 8.0, ctx=ctx)
    _, _, gc1 = dev.de_vjp_split(one, gpts, gpw)
    torch.cuda.synchronize()
    assert (gc1.cpu().numpy() == 0).all()


def test_host_vjp_batch_equals_device_entries(ctx, torch_cuda):
    """cth_de_vjp_batch (host buffers) gives cth_de_grad_dev's results with
    the forward sweep and cth_de_vjp_dev's with the reverse sweep, bit for
    bit, with per-point powers and a grad_d; a bad sweep is refused."""
    import ctypes as C
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import shell_points
    torch = torch_cuda
    pts, powers = shell_points(20000, seed=4)
    g = np.random.default_rng(41).uniform(-2, 2, len(pts)).astype(F32)
    shape = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    for sweep, fn in ((_lib.CTH_SWEEP_FORWARD, dev.de_grad), (_lib.CTH_SWEEP_REVERSE, dev.de_vjp)):
        d, gp, gw = np.empty(len(pts), F32), np.empty((len(pts), 3), F32), np.empty(len(pts), F32)
        _lib.check(_lib.lib.cth_de_vjp_batch(ctx.handle, C.byref(shape.desc), P(pts), P(powers), len(pts), P(g),
                                             P(d), P(gp), P(gw), sweep))
        d1, gp1, gw1 = (t.cpu().numpy() for t in fn(shape, gpu(torch, pts), gpu(torch, powers), gpu(torch, g),
                                                     with_power=True))
        for a, b in ((d, d1), (gp, gp1), (gw, gw1)):
            assert (a.view(np.uint32) == b.view(np.uint32)).all(), sweep
    with pytest.raises(ca.CthError):
        _lib.check(_lib.lib.cth_de_vjp_batch(ctx.handle, C.byref(shape.desc), P(pts), None, len(pts), None, P(d),
                                             None, None, 7))
