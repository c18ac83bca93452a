"""GPU surface nets (naive_surface_nets steps 2-3) against the oracle.

Bar: with the same corner distances, vertex counts, offsets and every index
(topology, emission order and winding) are bit-identical, and so are the
vertex positions (the reference's f32 operation order).  The 6-tap normals and
distance_from_surface are distance evaluations again: bit-identical for the
sphere, and for the Mandelbulb equal except where the final ln differs by an
ulp between glibc's logf and the correctly rounded device log; with the
oracle's loop state and a correctly rounded ln in the reference's formulas
they are bit-identical, except where that ln is a hard case for rounding
(refmath.cr_logf): its exact value within 2^-18 ulps of a rounding boundary,
where the device's f64 logarithm may round either way.
"""
import numpy as np
import pytest

import oracle as O
import refmath as R

pytestmark = pytest.mark.gpu

F32 = np.float32
ROOT = [-1.5, -1.5, -1.5, 1.5, 1.5, 1.5]


def same_bits(a, b):
    return np.asarray(a, F32).view(np.uint32) == np.asarray(b, F32).view(np.uint32)


def gpu(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = (a * b).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
    return np.arccos(np.clip(c, -1, 1))


def octree_leaves(count=2000, seed=0, depth=8):
    """A seeded sample of the depth-8 surface leaves of the octree workload."""
    from cantucci_amd.workloads import half_diagonals
    leaves, _ = O.refine(O.mandelbulb(15, 2.0), ROOT, 2, depth, half_diagonals(3.0, depth))
    idx = np.random.default_rng(seed).choice(len(leaves), min(count, len(leaves)), replace=False)
    return leaves[np.sort(idx)]


def run_gpu_mesh(torch, shape, spans, res, normals="fd6"):
    from cantucci_amd import device as dev
    sp = gpu(torch, spans)
    d, cb = dev.sample_spans(shape, sp, res)
    cb2 = torch.full_like(cb, -1)
    voff, ioff, v, ix = dev.mesh(shape, sp, d, res, normals, cross=cb2)
    torch.cuda.synchronize()
    assert torch.equal(cb, cb2), "crossing words of the count pass differ from span sampling's"
    return (d.cpu().numpy(), voff.cpu().numpy().astype(np.uint64), ioff.cpu().numpy().astype(np.uint64),
            v.cpu().numpy(), ix.cpu().numpy().view(np.uint32))


def assert_topology(voff, ioff, ix, voff_o, ioff_o, ix_o):
    assert np.array_equal(voff, voff_o), "per-leaf vertex counts differ"
    assert np.array_equal(ioff, ioff_o), "per-leaf index counts differ"
    assert np.array_equal(ix, ix_o), "indices differ"


@pytest.mark.parametrize("res", [8, 1, 2, 4, 16, 32, 64])
def test_sphere_mesh_bitexact(ctx, torch_cuda, res):
    """Sphere: exact closed-form distances, so the whole mesh -- positions,
    6-tap normals and distances -- is bit-identical."""
    import cantucci_amd as ca
    torch = torch_cuda
    center, radius = (0.1, 0.0, -0.05), 0.8
    shape = ca.Sphere(center, radius, ctx=ctx)
    from cantucci_amd.workloads import half_diagonals
    leaves, _ = O.refine(O.sphere(center, radius), [-1, -1, -1, 1, 1, 1], 1, 4, half_diagonals(2.0, 4))
    d, voff, ioff, v, ix = run_gpu_mesh(torch, shape, leaves, res)
    voff_o, ioff_o, v_o, ix_o = O.mesh_batch(O.sphere(center, radius), leaves, res, dists=d)
    assert_topology(voff, ioff, ix, voff_o, ioff_o, ix_o)
    assert voff[-1] > 0 and (ioff[-1] > 0 or res == 1)
    assert same_bits(v, v_o).all()
    # end to end: the oracle's own corner sampling gives the same mesh
    voff_e, ioff_e, v_e, ix_e = O.mesh_batch(O.sphere(center, radius), leaves, res)
    assert_topology(voff, ioff, ix, voff_e, ioff_e, ix_e)
    assert same_bits(v, v_e).all()


def crossing_words(d, res):
    """Crossing words of a batch of corner fields: bit c of a leaf's words is
    cell c = (x * res + y) * res + z, set when its 8 corner sign bits differ."""
    n1 = res + 1
    s = np.signbit(d.reshape(-1, n1, n1, n1))
    c = [s[:, x:x + res, y:y + res, z:z + res] for x in (0, 1) for y in (0, 1) for z in (0, 1)]
    allneg = np.logical_and.reduce(c)
    anyneg = np.logical_or.reduce(c)
    bits = (anyneg & ~allneg).reshape(len(d), -1)
    nwords = (res ** 3 + 63) // 64
    pad = np.zeros((len(d), nwords * 64), bool)
    pad[:, :bits.shape[1]] = bits
    w = np.packbits(pad.reshape(len(d), nwords, 64)[:, :, ::-1], axis=2).view(">u8")[..., 0]
    return w.astype(np.uint64)


@pytest.mark.parametrize("res", [1, 2, 4, 8, 16])
def test_mesh_signed_zero_and_nan_corners(ctx, torch_cuda, res):
    """Crafted corner fields holding -0.0, +0.0 and NaNs of both signs: the
    crossing words (sign bit, the reference's !is_sign_positive), counts,
    every index (winding by `< 0.0`, false for -0.0 and NaN) and the vertex
    positions match the oracle's; the words also match a numpy restatement."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    rng = np.random.default_rng(50 + res)
    spans = octree_leaves(96, seed=7)
    n1 = res + 1
    vals = np.array([-1.0, -0.25, -0.0, 0.0, 0.25, 1.0, np.nan, -np.nan], F32)
    assert np.signbit(vals[2]) and np.signbit(vals[7]) and not np.signbit(vals[6])
    d = vals[rng.integers(0, len(vals), (len(spans), n1 ** 3))]
    shape = ca.Sphere((0.0, 0.0, 0.0), 1.0, ctx=ctx)
    nwords = (res ** 3 + 63) // 64
    cb = torch.full((len(spans), nwords), -1, dtype=torch.int64, device="cuda")
    voff, ioff, v, ix = dev.mesh(shape, gpu(torch, spans), gpu(torch, d), res, "fd6", cross=cb)
    torch.cuda.synchronize()
    voff = voff.cpu().numpy().astype(np.uint64)
    ioff = ioff.cpu().numpy().astype(np.uint64)
    v = v.cpu().numpy()
    ix = ix.cpu().numpy().view(np.uint32)
    assert (cb.cpu().numpy().view(np.uint64) == crossing_words(d, res)).all()
    voff_o, ioff_o, v_o, ix_o = O.mesh_batch(O.sphere((0.0, 0.0, 0.0), 1.0), spans, res, dists=d)
    assert_topology(voff, ioff, ix, voff_o, ioff_o, ix_o)
    assert int(voff[-1]) > 0
    pos_same = same_bits(v[:, :3], v_o[:, :3]) | (np.isnan(v[:, :3]) & np.isnan(v_o[:, :3]))
    assert pos_same.all()


def test_octree_leaf_mesh_parity(ctx, torch_cuda):
    """2,000 seeded depth-8 surface leaves of the octree workload (P8, 15 it,
    bailout 2, 8^3 cells per leaf)."""
    import cantucci_amd as ca
    torch = torch_cuda
    spans = octree_leaves(2000)
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    d, voff, ioff, v, ix = run_gpu_mesh(torch, shape, spans, 8)
    voff_o, ioff_o, v_o, ix_o = O.mesh_batch(O.mandelbulb(15, 2.0), spans, 8, dists=d)
    assert_topology(voff, ioff, ix, voff_o, ioff_o, ix_o)
    nv = int(voff[-1])
    print(f"{len(spans)} leaves: {nv} vertices, {int(ioff[-1]) // 6} quads")
    assert nv > 10 * len(spans)
    assert same_bits(v[:, :3], v_o[:, :3]).all(), "vertex positions differ"
    dist_same = same_bits(v[:, 6], v_o[:, 6]) | (np.isnan(v[:, 6]) & np.isnan(v_o[:, 6]))
    nrm_same = same_bits(v[:, 3:6], v_o[:, 3:6]).all(1)
    print(f"distance bit-exact {dist_same.mean():.5f}, normal bit-exact {nrm_same.mean():.5f}")
    # The distance differs from the oracle's only through the final ln: it is
    # the correctly rounded f32 ln of the (bit-identical) loop state, where
    # glibc's logf is off by an ulp on a few percent of these near-surface r.
    d_cr, _, hard = _cr_distances(O.mandelbulb(15, 2.0), v[:, :3])
    cr_same = same_bits(v[:, 6], d_cr) | (np.isnan(v[:, 6]) & np.isnan(d_cr))
    print(f"distance = correctly rounded formula on {cr_same.mean():.6f}")
    assert dist_same.mean() > 0.97 and (cr_same | hard).all()
    fin = np.isfinite(v[:, 6]) & np.isfinite(v_o[:, 6])
    rel = np.abs(v[fin, 6].astype(np.float64) - v_o[fin, 6]) / np.maximum(np.abs(v_o[fin, 6].astype(np.float64)), 1e-30)
    assert rel.max() < 1e-5
    # 6-tap normals: the same up to those ulps of the tap distances, and the
    # reference's formulas exactly once the taps' ln is correctly rounded
    assert nrm_same.mean() > 0.95
    n_cr, n_hard = _cr_fd_normals(O.mandelbulb(15, 2.0), spans, voff, v, 8)
    n_same = (same_bits(v[:, 3:6], n_cr) | (np.isnan(v[:, 3:6]) & np.isnan(n_cr))).all(1)
    print(f"normal = the formulas with a correctly rounded ln on {n_same.mean():.6f}")
    assert (n_same | n_hard).all()
    fin = np.isfinite(v[:, 3:6]).all(1) & np.isfinite(v_o[:, 3:6]).all(1)
    ang = angle(v[fin, 3:6], v_o[fin, 3:6])
    print(f"6-tap normal angle vs oracle: p99 {np.percentile(ang, 99):.2e} p99.9 {np.percentile(ang, 99.9):.2e}")
    assert np.percentile(ang, 99.9) < 1e-3
    # with the oracle's own corner distances the sign field, and so the
    # topology, is the same
    d_o, _, _ = O.sample_spans(O.mandelbulb(15, 2.0), spans, 8)
    assert (np.signbit(d.reshape(d_o.shape)) == np.signbit(d_o)).all()
    voff_e, ioff_e, _, ix_e = O.mesh_batch(O.mandelbulb(15, 2.0), spans, 8)
    assert_topology(voff, ioff, ix, voff_e, ioff_e, ix_e)


def corner_points(spans, res):
    """The corners of each span in GridTable order, formed as the reference
    forms them (src/mesh/buffer.rs:64-83): s' = s + -(e - s) / res,
    e' = e + (e - s) / res, corner = s' + (e' - s') * (i / res) per axis."""
    L = np.asarray(spans, F32).reshape(-1, 6)
    n1 = res + 1
    t = np.arange(n1, dtype=F32) / F32(res)
    of = (L[:, 3:] - L[:, :3]) / F32(res)
    s1 = L[:, :3] + -of
    across = (L[:, 3:] + of) - s1
    ax = s1[:, :, None] + across[:, :, None] * t[None, None, :]
    P = np.empty((len(L), n1, n1, n1, 3), F32)
    P[..., 0] = ax[:, 0, :, None, None]
    P[..., 1] = ax[:, 1, None, :, None]
    P[..., 2] = ax[:, 2, None, None, :]
    return P.reshape(-1, 3)


def _cr_distances(shape_o, pts):
    """The reference's distances at pts from the oracle's loop state, with the
    final ln correctly rounded (as the kernels compute it) instead of glibc's
    logf; also the iteration counts and the mask of points whose ln is a hard
    case for rounding (refmath.cr_logf), where the device may round either way."""
    _, it, r, dr = O.de_points(shape_o, pts)
    ln_cr, hard = R.cr_logf(r)
    with np.errstate(all="ignore"):
        return (F32(0.5) * (ln_cr * r)) / dr, it, hard


def _cr_fd_normals(shape_o, spans, voff, verts, res):
    """6-tap normals of the vertices (src/mesh/buffer.rs:256-272: half-step
    0.7 (e' - s') / res of the adjusted span, taps p +- unit_a * delta_a,
    g = DE(+) - DE(-), normalize as g * (1 / |g|)) from _cr_distances, and
    the mask of vertices with a tap whose ln is a hard case for rounding."""
    sp = np.asarray(spans, F32).reshape(-1, 6)
    counts = np.diff(voff.astype(np.int64))
    s0 = np.repeat(sp[:, :3], counts, axis=0)
    e0 = np.repeat(sp[:, 3:], counts, axis=0)
    fres = F32(res)
    of = (e0 - s0) / fres
    s2, e2 = s0 + -of, e0 + of
    delta = (F32(0.7) * (e2 - s2)) / fres
    p = verts[:, :3]
    g = np.empty_like(p)
    hard = np.zeros(len(p), bool)
    for a in range(3):
        pp, pm = p.copy(), p.copy()
        pp[:, a] = p[:, a] + delta[:, a]
        pm[:, a] = p[:, a] + -delta[:, a]
        d_p, _, h_p = _cr_distances(shape_o, pp)
        d_m, _, h_m = _cr_distances(shape_o, pm)
        g[:, a] = d_p - d_m
        hard |= h_p | h_m
    with np.errstate(all="ignore"):
        mag = np.sqrt((g[:, 0] * g[:, 0] + g[:, 1] * g[:, 1]) + g[:, 2] * g[:, 2])
        return g * (F32(1.0) / mag)[:, None], hard


@pytest.mark.parametrize("mi", [1, 2, 3])
def test_mesh_taps_short_caps(ctx, torch_cuda, mi):
    """The tap kernel at small iteration caps, where most taps end at the cap
    (with max_iters 1 every tap that does not bail at its start does, before
    any full body): vertex positions bit-exact, and each distance and 6-tap
    normal equal to the reference's formulas on the oracle's loop state with
    the final ln correctly rounded."""
    import cantucci_amd as ca
    torch = torch_cuda
    spans = octree_leaves(300, seed=11)
    shape = ca.Mandelbulb.classic(mi, 2.0, ctx=ctx)
    sh = O.mandelbulb(mi, 2.0)
    d, voff, ioff, v, ix = run_gpu_mesh(torch, shape, spans, 8)
    voff_o, ioff_o, v_o, ix_o = O.mesh_batch(sh, spans, 8, dists=d)
    assert_topology(voff, ioff, ix, voff_o, ioff_o, ix_o)
    nv = int(voff[-1])
    assert nv > 100
    assert same_bits(v[:, :3], v_o[:, :3]).all()
    d_cr, it_o, d_hard = _cr_distances(sh, v[:, :3])
    capped = (it_o == mi).mean()
    cr_same = same_bits(v[:, 6], d_cr) | (np.isnan(v[:, 6]) & np.isnan(d_cr))
    n_cr, n_hard = _cr_fd_normals(sh, spans, voff, v, 8)
    n_same = (same_bits(v[:, 3:6], n_cr) | (np.isnan(v[:, 3:6]) & np.isnan(n_cr))).all(1)
    print(f"max_iters {mi}: {nv} vertices, {capped:.3f} at the cap; distance and normal = the formulas with a "
          f"correctly rounded ln on {cr_same.mean():.6f} / {n_same.mean():.6f}")
    assert capped > 0.3
    assert (cr_same | d_hard).all()
    assert (n_same | n_hard).all()


def test_tangent_normals_on_mesh(ctx, torch_cuda):
    """Forward-mode vertex normals: same vertices and distances as the 6-tap
    mesh, normals within 1e-3 rad of the FP64 dual-number gradient."""
    import cantucci_amd as ca
    torch = torch_cuda
    spans = octree_leaves(500, seed=3)
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    _, voff, ioff, v_fd, ix_fd = run_gpu_mesh(torch, shape, spans, 8, "fd6")
    _, voff_t, ioff_t, v_t, ix_t = run_gpu_mesh(torch, shape, spans, 8, "tangent")
    assert_topology(voff_t, ioff_t, ix_t, voff, ioff, ix_fd)
    assert same_bits(v_t[:, :3], v_fd[:, :3]).all()
    dsame = same_bits(v_t[:, 6], v_fd[:, 6]) | (np.isnan(v_t[:, 6]) & np.isnan(v_fd[:, 6]))
    assert dsame.all()
    pts = v_t[:, :3]
    sh = O.mandelbulb(15, 2.0)
    _, g64, _ = O.de_grad_f64(sh, pts.astype(np.float64))
    # Conditioning of each vertex: how far the exact gradient direction turns
    # when the f32 point moves by one ulp along an axis.  Vertices sit on the
    # fractal surface, where many are so ill-conditioned that no f32
    # evaluation can recover the FP64 direction; the bar applies to the rest.
    turn = np.zeros(len(pts))
    for a in range(3):
        for dirn in (np.inf, -np.inf):
            q = pts.copy()
            q[:, a] = np.nextafter(q[:, a], F32(dirn))
            _, gq, _ = O.de_grad_f64(sh, q.astype(np.float64))
            with np.errstate(all="ignore"):
                turn = np.maximum(turn, np.nan_to_num(angle(gq[:, :3], g64[:, :3]), nan=np.pi))
    ok = np.isfinite(g64[:, :3]).all(1) & (np.linalg.norm(g64[:, :3], axis=1) > 0) & np.isfinite(v_t[:, 3:6]).all(1)
    ang_t = angle(v_t[ok, 3:6], g64[ok, :3])
    ang_fd = angle(v_fd[ok, 3:6], g64[ok, :3])
    stable = turn[ok] < 1e-4
    print(f"tangent p50 {np.median(ang_t):.2e} p99 {np.percentile(ang_t, 99):.2e} "
          f"(<=1e-3 on {np.mean(ang_t <= 1e-3):.4f}); well-conditioned {stable.mean():.4f}: "
          f"p50 {np.median(ang_t[stable]):.2e} p99 {np.percentile(ang_t[stable], 99):.2e} "
          f"(<=1e-3 on {np.mean(ang_t[stable] <= 1e-3):.4f}); "
          f"6-tap (a difference at 0.7 cells, not a point gradient) p50 {np.median(ang_fd):.2e}")
    assert stable.mean() > 0.3
    assert np.median(ang_t[stable]) < 5e-5 and np.mean(ang_t[stable] <= 1e-3) > 0.99


def _host_sample(shape, sp, res):
    """Corner distances through the host-buffer span-sampling entry point."""
    import ctypes as C
    from cantucci_amd import _lib
    d = np.empty((res + 1) ** 3, F32)
    _lib.check(_lib.lib.cth_sample_spans(shape.ctx.handle, C.byref(shape.desc), np.ascontiguousarray(sp).ctypes.data,
                                         1, res, d.ctypes.data, None))
    return d


@pytest.mark.parametrize("res,span", [
    (64, ((-0.6, -0.6, 0.7), (-0.3, -0.3, 1.0))),
    (32, ((0.2, 0.4, -0.3), (0.5, 0.7, 0.0))),
])
def test_generate_for_box(ctx, res, span):
    """MeshBuffer::generate_for_box mirror at the reference's own resolution
    (64) with the app's shape, classic(6, 2.5)."""
    import cantucci_amd as ca
    from cantucci_amd.mesh import MeshBuffer
    shape = ca.Mandelbulb.classic(6, 2.5, ctx=ctx)
    buf, tm = MeshBuffer.generate_for_box(span, shape, res)
    sp = np.array(span[0] + span[1], F32)
    d_o, v_o, ix_o = O.surface_nets(O.mandelbulb(6, 2.5), sp, res)
    assert tm.vertices == len(v_o) and tm.faces * 6 == len(ix_o) and tm.vertices > 100
    assert np.array_equal(buf.indices, ix_o)
    pos = buf.vertices["position"]
    # end to end, positions move only through the corner distances' last-ulp
    # differences (the final ln); with the device's corner distances they
    # are bit-identical
    cell = (sp[3] - sp[0]) / res
    assert np.abs(pos - v_o[:, :3]).max() < 1e-4 * cell
    d_dev = _host_sample(shape, sp, res)
    _, _, v_same, ix_same = O.mesh_batch(O.mandelbulb(6, 2.5), sp[None], res, dists=d_dev)
    assert np.array_equal(buf.indices, ix_same)
    assert same_bits(pos, v_same[:, :3]).all()
    dist_same = same_bits(buf.vertices["distance_from_surface"], v_o[:, 6])
    assert dist_same.mean() > 0.97


def test_mesh_edge_cases(ctx, torch_cuda):
    """No crossing cells (a leaf far outside / inside), an empty batch, and
    the reference's asserts as errors."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.mesh import MeshBuffer
    torch = torch_cuda
    shape = ca.Sphere((0.0, 0.0, 0.0), 0.5, ctx=ctx)
    spans = np.array([[2, 2, 2, 3, 3, 3], [-0.1, -0.1, -0.1, 0.1, 0.1, 0.1], [0.3, -0.2, -0.2, 0.7, 0.2, 0.2]], F32)
    d, voff, ioff, v, ix = run_gpu_mesh(torch, shape, spans, 4)
    assert voff[0] == 0 and voff[1] == 0 and voff[2] == 0 and voff[3] > 0
    voff_o, ioff_o, v_o, ix_o = O.mesh_batch(O.sphere((0, 0, 0), 0.5), spans, 4, dists=d)
    assert_topology(voff, ioff, ix, voff_o, ioff_o, ix_o)
    assert same_bits(v, v_o).all()
    empty = torch.empty((0, 6), device="cuda")
    de = torch.empty((0, 125), device="cuda")
    voff, ioff, v, ix = dev.mesh(shape, empty, de, 4)
    assert voff.cpu().tolist() == [0] and v.shape[0] == 0 and ix.shape[0] == 0
    with pytest.raises(ca.CthError):
        MeshBuffer.generate_for_box(((0, 0, 0), (0, 1, 1)), shape, 8)  # start == end on x
    with pytest.raises(ca.CthError):
        MeshBuffer.generate_for_box(((0, 0, 0), (1, 1, 1)), shape, 12)  # not a power of two
    with pytest.raises(ca.CthError):
        MeshBuffer.generate_for_box(((0, 0, 0), (1, 1, 1)), shape, 8, normals="tangent")  # sphere: closed form


def test_shape_mesh_new_and_update(ctx, torch_cuda):
    """ShapeMesh::new and two updates (src/mesh/mod.rs:45-170) against the
    oracle: the 64 leaves of new(), the leaves the focus-point rule splits,
    and every new leaf's mesh at resolution 64: every corner distance equal
    to the reference's formula with the ln correctly rounded (outside hard
    cases), topology and positions bit for bit against the oracle's surface
    nets on the device's corner distances, and every vertex distance and
    6-tap normal equal to the reference's formulas (outside hard cases)."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.mesh import Camera, ShapeMesh
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(6, 2.5, ctx=ctx)
    sh = O.mandelbulb(6, 2.5)
    sm = ShapeMesh(shape)
    want = np.concatenate([O.create_spans(c) for c in O.create_spans([-1.2] * 3 + [1.2] * 3)])
    spans_of = lambda: np.array([[*a, *b] for (a, b), _ in sm.leaves], F32)
    assert same_bits(spans_of(), want).all() and len(sm.leaves) == 64

    def check_meshes(which):
        sp = spans_of()[which]
        d, _ = dev.sample_spans(shape, gpu(torch, sp), 64, with_cross=False)
        dn = d.cpu().numpy()
        d_cr, _, c_hard = _cr_distances(sh, corner_points(sp, 64))
        c_same = same_bits(dn.reshape(-1), d_cr) | (np.isnan(dn.reshape(-1)) & np.isnan(d_cr))
        assert (c_same | c_hard).all()
        voff_o, ioff_o, v_o, ix_o = O.mesh_batch(sh, sp, 64, dists=dn)
        for k, i in enumerate(which):
            buf = sm.leaves[i][1]
            a, b = int(voff_o[k]), int(voff_o[k + 1])
            assert buf is not None and len(buf.vertices) == b - a
            assert (buf.indices == ix_o[int(ioff_o[k]):int(ioff_o[k + 1])]).all()
            assert same_bits(buf.vertices["position"], v_o[a:b, :3]).all()
        V = np.concatenate([sm.leaves[i][1].vertices.view(F32).reshape(-1, 7) for i in which])
        dv, _, dv_hard = _cr_distances(sh, V[:, :3])
        nv, nv_hard = _cr_fd_normals(sh, sp, voff_o, V, 64)
        dv_same = same_bits(V[:, 6], dv) | (np.isnan(V[:, 6]) & np.isnan(dv))
        nv_same = (same_bits(V[:, 3:6], nv) | (np.isnan(V[:, 3:6]) & np.isnan(nv))).all(1)
        assert (dv_same | dv_hard).all() and (nv_same | nv_hard).all()
        print(f"{len(which)} leaves: {dn.size} corners ({int((~c_same).sum())} differ, {int(c_hard.sum())} hard), "
              f"{len(V)} vertices ({int((~dv_same).sum())} distances and {int((~nv_same).sum())} normals differ)")

    cam = Camera((-1.6, 0.2, 0.1), (1.0, -0.1, -0.05), 1.0, 16 / 9, 0.00004, 10.0)
    sm.update(cam)  # nothing is meshed yet, so nothing splits; the 64 leaves are meshed
    assert same_bits(spans_of(), want).all()
    check_meshes(list(range(64)))
    c10 = np.array([*cam.position, *cam.direction, cam.fov, cam.aspect, cam.near_plane, cam.far_plane], F32)
    pos, depth, _, _ = O.get_focii(sh, c10, 5)
    flags = O.focus_split(want, pos, c10[:3], depth)
    want2, children = [], []
    for i, f in enumerate(flags):
        if f:
            children.extend(range(len(want2), len(want2) + 8))
            want2.extend(O.create_spans(want[i]))
        else:
            want2.append(want[i])
    sm.update(cam)
    print(f"{int(flags.sum())} leaves split, {len(sm.leaves)} leaves, {sm.batch_timings.vertices} vertices meshed")
    assert flags.sum() > 0 and same_bits(spans_of(), np.array(want2, F32)).all()
    assert all(buf is not None for _, buf in sm.leaves)
    check_meshes(children)
