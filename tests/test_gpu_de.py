"""GPU parity of the distance-estimation kernels against the oracle.

Bar: the P8 loop state (final r, dr and the iteration count) is bit-identical
to the reference's f32 sequence; the distance differs only through the final
`ln` (ocml vs glibc), at most a few ulp.  NaNs carry the x86 sign bit.
"""
import numpy as np
import pytest

import oracle as O
import refmath as R
from test_oracle import NAN_SHAPES, nan_input_points

pytestmark = pytest.mark.gpu

F32 = np.float32
GRID = dict(start=(-1.5, -1.5, -1.5), end=(1.5, 1.5, 1.5))


def ulp_dist(a, b):
    a = np.asarray(a, F32).view(np.int32).astype(np.int64)
    b = np.asarray(b, F32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, np.int64(-2 ** 31) - a, a)
    b = np.where(b < 0, np.int64(-2 ** 31) - b, b)
    return np.abs(a - b)


def same_bits(a, b):
    return np.asarray(a, F32).view(np.uint32) == np.asarray(b, F32).view(np.uint32)


def assert_de_parity(d_gpu, d_ref, max_ulp=4, min_bitexact=0.99):
    """NaN pattern and signs identical, <= max_ulp everywhere, and (because the
    only non-reference op is a correctly rounded ln against glibc's logf) at
    least min_bitexact of the values bit-identical."""
    nan_g, nan_r = np.isnan(d_gpu), np.isnan(d_ref)
    bit = same_bits(d_gpu, d_ref) | (nan_g & nan_r)
    assert bit.mean() >= min_bitexact, f"only {bit.mean():.5f} bit-identical"
    assert (nan_g == nan_r).all(), "NaN pattern differs"
    # NaNs carry the x86 default-NaN sign bit
    assert same_bits(d_gpu[nan_g], d_ref[nan_r]).all()
    fin = ~nan_g
    u = ulp_dist(d_gpu[fin], d_ref[fin])
    assert u.max() <= max_ulp, f"max ulp {u.max()}"
    inf = np.isinf(d_ref)
    assert same_bits(d_gpu[inf], d_ref[inf]).all()
    fin &= ~inf
    rel = np.abs(d_gpu[fin].astype(np.float64) - d_ref[fin]) / np.maximum(np.abs(d_ref[fin].astype(np.float64)), 1e-30)
    assert rel.max() < 1e-5


def assert_cr_formula(d_gpu, r_o, dr_o):
    """The distances equal the reference's final formula 0.5 * (ln(r) * r) / dr
    (src/shape/mandelbulb.rs:77-78) on the oracle's loop state with the ln
    correctly rounded, as the device computes it: bit for bit, NaNs by
    pattern, wherever the ln is not a hard case for rounding (its exact value
    within 2^-18 ulps of a rounding boundary, about 2^-17 of inputs, where the
    device's f64 logarithm may round either way)."""
    ln_cr, hard = R.cr_logf(r_o)
    with np.errstate(all="ignore"):
        d_cr = (F32(0.5) * (ln_cr * r_o)) / dr_o
    same = same_bits(d_gpu, d_cr) | (np.isnan(d_gpu) & np.isnan(d_cr))
    assert (same | hard).all(), f"{int((~same & ~hard).sum())} distances differ from the formula"


def assert_same_or_both_nan(a, b):
    """Bit-identical, except that the sign of an intermediate NaN is not part
    of the reference's observable behaviour (only the final distance's is)."""
    na, nb = np.isnan(a), np.isnan(b)
    assert (na == nb).all()
    assert same_bits(a[~na], b[~nb]).all()


def ulp_stable(shape, pts, d_ref, powers=None, tol=1e-6):
    """Points whose f32 reference distance is determined to `tol` at the ulp
    level: moving any coordinate by one ulp either way changes it by at most
    tol relative.  Elsewhere the f32 orbit is chaotic and no implementation
    with different (even correctly rounded) transcendentals can agree."""
    ok = np.ones(len(pts), bool)
    for ax in range(3):
        for direction in (np.inf, -np.inf):
            q = pts.copy()
            q[:, ax] = np.nextafter(q[:, ax], F32(direction))
            dq = O.de_points(shape, q, powers)[0]
            ok &= np.abs(dq.astype(np.float64) - d_ref) <= tol * np.abs(d_ref.astype(np.float64))
    return ok


def grid_points(n, divisor, start=GRID["start"], end=GRID["end"]):
    """The reference's corner formula p = s + (e - s) * (i / divisor) in f32."""
    t = np.arange(n, dtype=F32) / F32(divisor)
    ax = [F32(start[a]) + (F32(end[a]) - F32(start[a])) * t for a in range(3)]
    X, Y, Z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(F32)


def gpu(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("divisor", [63.0, 64.0])
def test_grid64_p8_parity(ctx, torch_cuda, divisor):
    """64^3 grid, P8, 15 iterations, bailout 2 (divisor 64 puts the origin and
    64 z-axis points on the grid)."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    d, it = dev.de_grid(shape, GRID["start"], GRID["end"], 64, divisor, with_iters=True)
    torch.cuda.synchronize()
    dref, itref = O.de_grid(O.mandelbulb(15, 2.0), GRID["start"], GRID["end"], 64, divisor)
    assert (it.cpu().numpy().astype(np.uint32) == itref).all()
    assert_de_parity(d.cpu().numpy(), dref)
    # the loop state from explicit points with the same coordinates is bit-identical
    pts = grid_points(64, divisor)
    r, dr, it2 = dev.de_state(shape, gpu(torch, pts))
    _, it_o, r_o, dr_o = O.de_points(O.mandelbulb(15, 2.0), pts)
    assert (it2.cpu().numpy().astype(np.uint32) == it_o).all()
    assert_same_or_both_nan(r.cpu().numpy(), r_o)
    assert_same_or_both_nan(dr.cpu().numpy(), dr_o)
    assert_cr_formula(d.cpu().numpy(), r_o, dr_o)


def edge_points():
    e = [[0, 0, 0], [-0.0, 0, 0], [0, -0.0, 0.5], [-0.0, -0.0, -0.7], [0, 0, 1e-30], [0, 0, 2.0], [2.0, 0, 0],
         [0, 0, -1.1], [1e-20, 0, 0.3], [0, 1e-20, -0.3], [1e-38, 1e-38, 0.5], [1e-45, 0, 0.1], [3.0, 4.0, 0.0],
         [1.0, 1.0, 1.0], [-1.2, 1.2, -1.2], [0.5, 0.5, 0.0], [1e-3, -1e-3, 1e-3], [0.9, 0.0, -0.0],
         [1.4142135, 1.4142135, 0.0], [np.inf, 0, 0], [1e30, 1e30, 1e30]]
    return np.array(e, F32)


@pytest.mark.parametrize("params", [(15, 2.0), (8, 5.0), (6, 2.5), (1, 2.0), (100, 2.0)])
def test_points_p8_state_bitexact(ctx, torch_cuda, params):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from golden_data import BENCH_POINTS
    torch = torch_cuda
    mi, bo = params
    rng = np.random.default_rng(100 + mi)
    pts = np.concatenate([rng.uniform(-1.5, 1.5, (200000, 3)).astype(F32), edge_points(), BENCH_POINTS])
    shape = ca.Mandelbulb.classic(mi, bo, ctx=ctx)
    g = gpu(torch, pts)
    d, it = dev.de_points(shape, g, with_iters=True)
    r, dr, it2 = dev.de_state(shape, g)
    torch.cuda.synchronize()
    d_o, it_o, r_o, dr_o = O.de_points(O.mandelbulb(mi, bo), pts)
    assert (it.cpu().numpy().astype(np.uint32) == it_o).all()
    assert (it2.cpu().numpy().astype(np.uint32) == it_o).all()
    rg, drg = r.cpu().numpy(), dr.cpu().numpy()
    nan = np.isnan(r_o)
    assert (np.isnan(rg) == nan).all() and same_bits(rg[~nan], r_o[~nan]).all()
    nan = np.isnan(dr_o)
    assert (np.isnan(drg) == nan).all() and same_bits(drg[~nan], dr_o[~nan]).all()
    assert_de_parity(d.cpu().numpy(), d_o)
    assert_cr_formula(d.cpu().numpy(), r_o, dr_o)


@pytest.mark.parametrize("P", [2, 3, 5, 9])
def test_trig_integer_power(ctx, torch_cuda, P):
    """Mandelbulb<P != 8>: the trig map with ocml transcendentals.  Iteration
    counts must agree except at a bailout boundary; distances of points whose
    orbit is well conditioned (the FP64 value is within 1e-6 of the f32
    reference) agree to 1e-5."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    rng = np.random.default_rng(200 + P)
    pts = rng.uniform(-1.5, 1.5, (100000, 3)).astype(F32)
    shape = ca.Mandelbulb(15, 2.0, P, ctx=ctx)
    d, it = dev.de_points(shape, gpu(torch, pts), with_iters=True)
    torch.cuda.synchronize()
    d, it = d.cpu().numpy(), it.cpu().numpy().astype(np.uint32)
    d_o, it_o, _, _ = O.de_points(O.mandelbulb(15, 2.0, P), pts)
    d64, _ = O.de_points_f64(O.mandelbulb(15, 2.0, P), pts)
    rel = np.abs(d.astype(np.float64) - d_o) / np.maximum(np.abs(d_o.astype(np.float64)), 1e-30)
    cond = ulp_stable(O.mandelbulb(15, 2.0, P), pts, d_o)
    print(f"P={P}: iters equal {np.mean(it == it_o):.5f}, rel>1e-5 all {np.mean(rel > 1e-5):.5f}, "
          f"ulp-stable {cond.mean():.3f} of which rel>1e-5 {np.mean(rel[cond] > 1e-5):.6f}, "
          f"vs fp64 rel>1e-5 gpu {np.mean(np.abs(d - d64) > 1e-5 * np.abs(d64)):.5f} "
          f"oracle {np.mean(np.abs(d_o - d64) > 1e-5 * np.abs(d64)):.5f}")
    assert np.mean(it == it_o) > 0.995
    assert np.mean(rel[cond] > 1e-5) < 1e-3


def test_trig_real_power_per_point(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import shell_points
    torch = torch_cuda
    pts, powers = shell_points(100000, seed=1)
    shape = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    d, it = dev.de_points(shape, gpu(torch, pts), gpu(torch, powers), with_iters=True)
    torch.cuda.synchronize()
    d, it = d.cpu().numpy(), it.cpu().numpy().astype(np.uint32)
    s = O.mandelbulb_trig(15, 2.0)
    d_o, it_o, _, _ = O.de_points(s, pts, powers)
    d64, _ = O.de_points_f64(s, pts, powers)
    rel = np.abs(d.astype(np.float64) - d_o) / np.maximum(np.abs(d_o.astype(np.float64)), 1e-30)
    cond = ulp_stable(s, pts, d_o, powers)
    print(f"real power: iters equal {np.mean(it == it_o):.5f}, rel>1e-5 all {np.mean(rel > 1e-5):.5f}, "
          f"ulp-stable {cond.mean():.3f} of which rel>1e-5 {np.mean(rel[cond] > 1e-5):.6f}, "
          f"vs fp64 rel>1e-5 gpu {np.mean(np.abs(d - d64) > 1e-5 * np.abs(d64)):.5f} "
          f"oracle {np.mean(np.abs(d_o - d64) > 1e-5 * np.abs(d64)):.5f}")
    assert np.mean(it == it_o) > 0.99
    assert np.mean(rel[cond] > 1e-5) < 1e-3


def test_sphere_exact(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    pts = np.random.default_rng(3).uniform(-2, 2, (10000, 3)).astype(F32)
    shape = ca.Sphere((0.25, -0.5, 0.125), 0.75, ctx=ctx)
    d = dev.de_points(shape, gpu(torch, pts)).cpu().numpy()
    d_o = O.de_points(O.sphere((0.25, -0.5, 0.125), 0.75), pts)[0]
    assert same_bits(d, d_o).all()


def test_shape_trait_host_queries(ctx):
    import cantucci_amd as ca
    from golden_data import BENCH_POINTS
    m = ca.Mandelbulb.classic(8, 5.0, ctx=ctx)
    d = m.batch_min_distance_from(BENCH_POINTS)
    d_o = O.de_points(O.mandelbulb(8, 5.0), BENCH_POINTS)[0]
    assert_de_parity(d, d_o)
    assert m.min_distance_from(BENCH_POINTS[3]) == d[3]
    assert (m.batch_contains(BENCH_POINTS) == (d < 0)).all()
    with pytest.raises(ca.CthError) as e:
        m.batch_max_distance_from(BENCH_POINTS)
    assert e.value.code == 2
    with pytest.raises(ca.CthError):
        m.batch_bounded_distance_from(BENCH_POINTS)
    assert m.max_distance_from(BENCH_POINTS[0]) is None
    s = ca.Sphere((0, 0, 0), 1.0, ctx=ctx)
    pts = np.array([[0.5, 0, 0], [1.0, 0, 0], [2, 0, 0], [0, 0, 0]], F32)
    assert list(s.batch_contains(pts)) == [True, True, False, True]
    b = s.batch_bounded_distance_from(pts)
    assert (b[:, 0] == b[:, 1]).all() and (s.batch_max_distance_from(pts) == b[:, 0]).all()
    assert m.bounding_box() == ((-1.2, -1.2, -1.2), (1.2, 1.2, 1.2)) or np.allclose(m.bounding_box(), [[-1.2] * 3, [1.2] * 3])
    assert len(m.batch_min_distance_from(np.zeros((0, 3), F32))) == 0


def test_empty_and_invalid_device_calls(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    assert dev.de_points(shape, torch.zeros((0, 3), device="cuda")).numel() == 0
    with pytest.raises(ca.CthError):
        dev.sample_spans(shape, torch.zeros((1, 6), device="cuda"), 6)  # res not a power of two
    # batches of 2^31 points or more are refused before any launch (the
    # buffers below hold 16 points, so nothing is read even if one ran)
    import ctypes as C
    from cantucci_amd import _lib
    buf = torch.zeros((16, 3), device="cuda")
    for n in (2 ** 31, 2 ** 32 + 5):
        rc = _lib.lib.cth_normals_fd_dev(ctx.handle, C.byref(shape.desc), buf.data_ptr(), None, n, 1e-3,
                                         buf.data_ptr(), None, None, None)
        assert rc == _lib.CTH_E_INVALID and b"2^31" in _lib.lib.cth_last_error()
    # the torch wrappers check shapes and dtypes before launching
    with pytest.raises(ValueError):
        dev.de_points(shape, torch.zeros((8, 2), device="cuda"))
    with pytest.raises(ValueError):
        dev.de_points(ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx), buf, torch.ones(15, device="cuda"))
    with pytest.raises(TypeError):
        dev.de_points(ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx), buf, torch.ones(16, dtype=torch.float64,
                                                                              device="cuda"))
    with pytest.raises(ValueError):
        dev.de_vjp(shape, buf, grad_d=torch.ones(4, device="cuda"))
    torch.cuda.synchronize()


def octree_leaf_sample(count=400, seed=0):
    from cantucci_amd.workloads import half_diagonals
    leaves, _ = O.refine(O.mandelbulb(15, 2.0), [-1.5, -1.5, -1.5, 1.5, 1.5, 1.5], 2, 6, half_diagonals(3.0, 6))
    idx = np.random.default_rng(seed).choice(len(leaves), min(count, len(leaves)), replace=False)
    return leaves[np.sort(idx)]


@pytest.mark.parametrize("res", [8, 4, 16])
def test_span_sampling_and_sign_bits(ctx, torch_cuda, res):
    """Octree-leaf spans (depth-6 surface leaves of the octree workload's refinement): corner
    distances at parity and crossing bits identical to the oracle."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    spans = octree_leaf_sample(300 if res <= 8 else 60)
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    d, cb = dev.sample_spans(shape, gpu(torch, spans), res)
    torch.cuda.synchronize()
    d_o, cb_o, _ = O.sample_spans(O.mandelbulb(15, 2.0), spans, res)
    assert_de_parity(d.cpu().numpy().ravel(), d_o.ravel())
    assert (cb.cpu().numpy().view(np.uint64) == cb_o).all()


def test_span_sampling_sphere_and_host_entry(ctx):
    import cantucci_amd as ca
    from cantucci_amd import _lib
    import ctypes as C
    spans = np.array([[-1, -1, -1, 1, 1, 1], [0, 0, 0, 0.5, 0.5, 0.5]], F32)
    res = 8
    per = (res + 1) ** 3
    d = np.empty((2, per), F32)
    cb = np.empty((2, 8), np.uint64)
    shape = _lib.make_shape(_lib.CTH_SHAPE_SPHERE, 0, 1, 0, (0.0, 0.0, 0.0), 0.6)
    _lib.check(_lib.lib.cth_sample_spans(ctx.handle, C.byref(shape), spans.ctypes.data, 2, res, d.ctypes.data,
                                         cb.ctypes.data))
    d_o, cb_o, _ = O.sample_spans(O.sphere((0, 0, 0), 0.6), spans, res)
    assert same_bits(d, d_o).all() and (cb == cb_o).all()


def test_fast_mode_error_is_reported(torch_cuda):
    """Perf mode (FMA, hardware rcp/sqrt) is not bit-exact; its error on the
    64^3 grid is measured here, never used for parity."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    fctx = ca.Context(0, "fast")
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=fctx)
    d, it = dev.de_grid(shape, GRID["start"], GRID["end"], 64, 63.0, with_iters=True)
    torch.cuda.synchronize()
    dref, itref = O.de_grid(O.mandelbulb(15, 2.0), GRID["start"], GRID["end"], 64, 63.0)
    d = d.cpu().numpy()
    rel = np.abs(d.astype(np.float64) - dref) / np.maximum(np.abs(dref.astype(np.float64)), 1e-30)
    frac = np.mean(rel > 1e-5)
    print(f"fast mode: iters equal {np.mean(it.cpu().numpy() == itref):.4f}, rel>1e-5 {frac:.4f}, "
          f"median rel {np.median(rel):.2e}")
    assert np.median(rel) < 1e-5 and frac < 0.1
    fctx.close()


@pytest.mark.parametrize("max_depth", [5, 7])
def test_refinement_matches_oracle(ctx, torch_cuda, max_depth):
    """GPU level-synchronous refinement: same per-level node counts and the same
    surface leaves in the same breadth-first create_spans order."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import half_diagonals
    hd = half_diagonals(3.0, max_depth)
    root = [-1.5, -1.5, -1.5, 1.5, 1.5, 1.5]
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    leaves, lc = dev.refine(shape, root, 2, max_depth, hd)
    torch_cuda.cuda.synchronize()
    leaves_o, lc_o = O.refine(O.mandelbulb(15, 2.0), root, 2, max_depth, hd)
    assert list(lc) == list(lc_o)
    assert same_bits(leaves.cpu().numpy(), leaves_o).all()


def test_refinement_sphere_and_errors(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import half_diagonals
    hd = half_diagonals(2.0, 5)
    root = [-1, -1, -1, 1, 1, 1]
    leaves, lc = dev.refine(ca.Sphere((0.1, 0.0, -0.05), 0.6, ctx=ctx), root, 1, 5, hd)
    leaves_o, lc_o = O.refine(O.sphere((0.1, 0.0, -0.05), 0.6), root, 1, 5, hd)
    assert list(lc) == list(lc_o) and same_bits(leaves.cpu().numpy(), leaves_o).all()
    with pytest.raises(ca.CthError):
        dev.refine(ca.Mandelbulb.classic(15, 2.0, ctx=ctx), [1, -1, -1, -1, 1, 1], 0, 3, hd)


@pytest.mark.parametrize("bailout,lo,hi", [(2.0, 2.0, 5e3), (0.9, 0.9, 1.1), (1e-3, 1e-3, 0.9)])
def test_final_ln_is_correctly_rounded(ctx, torch_cuda, bailout, lo, hi):
    """Points beyond the bailout radius run no loop body, so d = 0.5 (ln r r) / 1
    isolates the device logarithm: it must be the correctly rounded f32 ln (the
    value glibc's logf returns for 99.6% of inputs) on every input whose ln is
    not a hard case for rounding."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    rng = np.random.default_rng(7)
    n = 400000
    rad = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    v = rng.normal(size=(n, 3))
    pts = (v / np.linalg.norm(v, axis=1, keepdims=True) * rad[:, None]).astype(F32)
    r = np.sqrt((pts[:, 0] * pts[:, 0] + pts[:, 1] * pts[:, 1]) + pts[:, 2] * pts[:, 2])
    keep = r > F32(bailout)
    pts, r = pts[keep], r[keep]
    shape = ca.Mandelbulb(15, bailout, 8, ctx=ctx)
    d = dev.de_points(shape, gpu(torch, pts))
    torch.cuda.synchronize()
    ln_cr, hard = R.cr_logf(r)
    expect = (F32(0.5) * (ln_cr * r)) / F32(1.0)
    agree = same_bits(d.cpu().numpy(), expect)
    print(f"bailout {bailout}: {len(pts)} points, correctly rounded ln on {agree.mean():.7f} "
          f"({int(hard.sum())} hard cases for rounding, {int((~agree & hard).sum())} of them differ)")
    assert (agree | hard).all()


def test_trig_integral_powers_zero_and_negative(ctx, torch_cuda):
    """Per-point integral powers 0, -0.0, 1, -2, -3 of the real-power kind take
    the powi ladder like the oracle (a negative exponent as the reciprocal of
    the ladder on its magnitude).  At power +-0 the map is (+-0, +-0, 1) and
    dr stays 1 whatever the transcendentals, so the loop state is
    bit-identical there; the other powers go through ocml's acos/atan2/sin/cos
    and are held to the ulp-stable bar of the other trig tests."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    rng = np.random.default_rng(77)
    pts = rng.uniform(-1.5, 1.5, (50000, 3)).astype(F32)
    choices = np.array([0.0, -0.0, 1.0, -2.0, -3.0], F32)
    powers = choices[np.arange(len(pts)) % len(choices)]
    shape = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    gp, gw = gpu(torch, pts), gpu(torch, powers)
    d = dev.de_points(shape, gp, gw).cpu().numpy()
    r, dr, it = (t.cpu().numpy() for t in dev.de_state(shape, gp, gw))
    s = O.mandelbulb_trig(15, 2.0)
    d_o, it_o, r_o, dr_o = O.de_points(s, pts, powers)
    zero = powers == 0.0
    assert (dr[zero] == 1.0).all() and (dr_o[zero] == 1.0).all()
    assert (it[zero].astype(np.uint32) == it_o[zero]).all()
    assert_same_or_both_nan(r[zero], r_o[zero])
    assert_same_or_both_nan(dr[zero], dr_o[zero])
    assert_de_parity(d[zero], d_o[zero])
    assert np.isfinite(d[zero]).all()  # the ladder's r^-1 instead of r^(2^32 - 1)
    rest = ~zero
    assert np.mean(it[rest].astype(np.uint32) == it_o[rest]) > 0.99
    cond = ulp_stable(s, pts[rest], d_o[rest], powers[rest])
    rel = np.abs(d[rest].astype(np.float64) - d_o[rest]) / np.maximum(np.abs(d_o[rest].astype(np.float64)), 1e-30)
    print(f"powers 1,-2,-3: iters equal {np.mean(it[rest].astype(np.uint32) == it_o[rest]):.5f}, "
          f"ulp-stable {cond.mean():.3f} of which rel>1e-5 {np.mean(rel[cond] > 1e-5):.6f}")
    assert np.mean(rel[cond] > 1e-5) < 1e-3
    # a uniform power of 0 on the shape itself (no per-point array)
    d0 = dev.de_points(ca.MandelbulbTrig(15, 2.0, 0.0, ctx=ctx), gp).cpu().numpy()
    d0_o = O.de_points(O.mandelbulb_trig(15, 2.0, 0.0), pts)[0]
    assert_de_parity(d0, d0_o)


def test_every_entry_point_takes_empty_input(ctx, torch_cuda):
    """Zero points, spans, leaves, cameras or rays (the reference's batch
    methods return an empty Vec for an empty slice, src/shape/util.rs:1-18):
    every device entry point returns at once with empty outputs (the mesh's
    offsets a single 0), and so do the host trait batches."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    p8 = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    trig = ca.MandelbulbTrig(15, 2.0, 7.5, ctx=ctx)
    sph = ca.Sphere((0.0, 0.0, 0.0), 1.0, ctx=ctx)
    z = lambda *s, dt=torch.float32: torch.zeros(s, dtype=dt, device="cuda")
    for shape in (p8, trig, sph):
        assert dev.de_points(shape, z(0, 3)).numel() == 0
        assert shape.batch_min_distance_from(np.zeros((0, 3), F32)).size == 0
        assert shape.batch_contains(np.zeros((0, 3), F32)).size == 0
    assert sph.batch_max_distance_from(np.zeros((0, 3), F32)).size == 0
    assert all(t.numel() == 0 for t in dev.de_state(p8, z(0, 3)))
    d, cb = dev.sample_spans(p8, z(0, 6), 8)
    assert d.shape == (0, 729) and cb.shape == (0, 8)
    voff, ioff, v, ix = dev.mesh(p8, z(0, 6), z(0, 729), 8, "fd6")
    assert voff.cpu().tolist() == [0] and ioff.cpu().tolist() == [0] and v.shape == (0, 7) and ix.numel() == 0
    assert dev.mesh_vertices(p8, z(0, 7)).shape == (0, 7)
    depth, steps = dev.sphere_trace(p8, z(1, 3), z(0, 3))
    assert depth.numel() == 0 and steps.numel() == 0
    depth, steps = dev.render_depth(p8, z(0, 14), 16, 9)
    assert depth.shape == (0, 9, 16) and steps.shape == (0, 9, 16)
    f, fd, fs = dev.get_focii(p8, z(0, 10), 5)
    assert f.shape == (0, 25, 3) and fd.shape == (0, 25)
    for fn in (dev.de_vjp, dev.de_vjp_split, dev.de_grad, dev.normals):
        assert all(t.numel() == 0 for t in fn(trig, z(0, 3)))
    assert all(t.numel() == 0 for t in dev.normals_fd(trig, z(0, 3), 1e-3))
    assert dev.focus_split(ctx, z(0, 6), z(0, 3), (0.0, 0.0, 3.0)).numel() == 0
    assert dev.split_leaves(ctx, z(0, 6), z(0, dt=torch.int32)).shape == (0, 6)
    torch.cuda.synchronize()


@pytest.mark.parametrize("name,mi,_s", NAN_SHAPES, ids=[s[0] + f"-{s[1]}" for s in NAN_SHAPES])
def test_nan_inputs_carry_their_sign(ctx, torch_cuda, name, mi, _s):
    """Points with NaN coordinates of either sign, and NaNs with payloads:
    iteration counts equal the oracle's, r and dr are NaN where the oracle's
    are, and every distance is a NaN with the oracle's sign (the first NaN
    coordinate's, in x, y, z order).  The distance is the canonical NaN of
    that sign; the oracle carries an input payload through, so payloads are
    compared by sign only."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    pay = np.array([0x7FC12345, 0xFFD00001, 0x7FFFFFFF], np.uint32).view(F32)
    extra = np.array([[pay[0], 0.3, -0.2], [0.5, pay[1], 0.1], [0.0, 0.0, pay[2]], [pay[1], pay[0], 0.0]], F32)
    pts = np.concatenate([nan_input_points(), extra])
    if name == "sphere":
        shape, orc = ca.Sphere((0.1, -0.2, 0.3), 1.0, ctx=ctx), O.sphere((0.1, -0.2, 0.3), 1.0)
    elif name == "trig7.5":
        shape, orc = ca.MandelbulbTrig(mi, 2.0, 7.5, ctx=ctx), O.mandelbulb_trig(mi, 2.0, 7.5)
    else:
        P = 8 if name == "p8" else int(name[4:])
        shape, orc = ca.Mandelbulb(mi, 2.0, P, ctx=ctx), O.mandelbulb(mi, 2.0, P)
    print(f"{name} max_iters {mi}: {len(pts)} points", flush=True)
    g = gpu(torch, pts)
    d, it = dev.de_points(shape, g, with_iters=True)
    d, it = d.cpu().numpy(), it.cpu().numpy().astype(np.uint32)
    d_o, it_o, r_o, dr_o = O.de_points(orc, pts)
    assert np.isnan(d_o).all() and np.isnan(d).all()
    assert (np.signbit(d) == np.signbit(d_o)).all()
    canon = np.where(np.signbit(d_o), np.uint32(0xFFC00000), np.uint32(0x7FC00000))
    assert (d.view(np.uint32) == canon).all()
    assert (it == it_o).all()
    if name != "sphere":
        r, dr, it2 = (t.cpu().numpy() for t in dev.de_state(shape, g))
        assert (it2.astype(np.uint32) == it_o).all()
        assert (np.isnan(r) == np.isnan(r_o)).all() and (np.isnan(dr) == np.isnan(dr_o)).all()
    # the host trait batch goes through the same kernels
    dh = shape.batch_min_distance_from(pts)
    assert (np.asarray(dh, F32).view(np.uint32) == canon).all()
