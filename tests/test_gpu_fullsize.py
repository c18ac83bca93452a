"""The benchmark workloads at their full sizes, checked through properties
that do not depend on size and, output by output, against the oracle.

* octree workload: level counts and leaves identical to the oracle's
  refinement; crossing words of span sampling and of the mesh count pass
  identical; per-leaf vertex count = popcount of the leaf's crossing words;
  every index below its leaf's vertex count; and every output (all corner
  distances, vertices and indices) against the oracle.
* 32 x 1080p sphere tracing: one seeded row per pose traced by the oracle,
  and every ray against the oracle's march with the correctly rounded ln.
* 2^22 shell points: every point's tangent (and 6-tap FD) normal against the
  FP64 dual gradient, also without the points whose stencil changes the
  iteration count.
* 64 mixed jobs: every point of every job against the oracle and the FP64
  dual gradient.
"""
import numpy as np
import pytest

import oracle as O
import refmath as R

pytestmark = pytest.mark.gpu
F32 = np.float32


def gpu(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def popcount64(words):
    w = words.view(np.uint64)
    c = np.zeros(w.shape, np.uint64)
    for s in range(0, 64, 8):
        c += np.array([bin(i).count("1") for i in range(256)], np.uint64)[((w >> np.uint64(s)) & np.uint64(255))]
    return c


def _octree_workload(ctx):
    """Refinement, span sampling and 6-tap meshing of the octree workload on
    the device, as the benchmark runs them."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    oc = wl.OCTREE
    hd = wl.half_diagonals(3.0, oc["max_depth"])
    shape = ca.Mandelbulb.classic(oc["max_iters"], oc["bailout"], ctx=ctx)
    leaves, lc = dev.refine(shape, oc["root"], oc["min_depth"], oc["max_depth"], hd)
    dists, cross = dev.sample_spans(shape, leaves, oc["res"])
    return shape, hd, leaves, lc, dists, cross


def test_octree_workload_full_size(ctx, torch_cuda):
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    torch = torch_cuda
    oc = wl.OCTREE
    shape, hd, leaves, lc, dists, cross = _octree_workload(ctx)
    torch.cuda.synchronize()
    leaves_o, lc_o = O.refine(O.mandelbulb(15, 2.0), oc["root"], oc["min_depth"], oc["max_depth"], hd)
    assert list(lc) == list(lc_o) and int(lc[-1]) == 1252864
    lv = leaves.cpu().numpy()
    assert lv.shape == (921604, 6) and (lv.view(np.uint32) == leaves_o.view(np.uint32)).all()

    res = oc["res"]
    cross2 = torch.full_like(cross, -1)
    voff, ioff, verts, idx = dev.mesh(shape, leaves, dists, res, "fd6", cross=cross2)
    torch.cuda.synchronize()
    assert torch.equal(cross, cross2)
    nleaf = lv.shape[0]
    vcount = (voff[1:] - voff[:-1]).cpu().numpy()
    pc = popcount64(cross.cpu().numpy()).sum(1)
    assert (vcount.astype(np.uint64) == pc).all()
    icount = (ioff[1:] - ioff[:-1])
    assert bool((icount % 6 == 0).all())
    lim = torch.repeat_interleave(torch.from_numpy(vcount).cuda(), icount)
    assert bool((idx.to(torch.int64) < lim).all()) and bool((idx >= 0).all())
    del lim
    print(f"{nleaf} leaves, {int(voff[-1])} vertices, {int(ioff[-1]) // 6} quads")
    assert finite_normals(verts) > 0.999


def test_octree_workload_every_output(ctx, torch_cuda):
    """Every output of the octree workload against the oracle, in chunks of
    40,000 leaves (671.8 M corner distances, 75.2 M vertices):
      * each corner distance has the sign of the oracle's (glibc ln) and
        equals the reference's final formula on the oracle's loop state with
        the ln correctly rounded, wherever that ln is not a hard case for
        rounding (refmath.cr_logf);
      * each leaf's vertex and index counts, every index and every vertex
        position equal the oracle's surface nets run on the device's corner
        distances;
      * each vertex distance and 6-tap normal equals the reference's formulas
        on the oracle's loop state with the same ln, outside the vertices with
        a hard case among their taps."""
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    from test_gpu_mesh import _cr_distances, _cr_fd_normals, corner_points
    torch = torch_cuda
    res = wl.OCTREE["res"]
    shape, _, leaves, _, dists, _ = _octree_workload(ctx)
    voff, ioff, verts, idx = dev.mesh(shape, leaves, dists, res, "fd6")
    torch.cuda.synchronize()
    lv = leaves.cpu().numpy()
    vo = voff.cpu().numpy().astype(np.int64)
    io = ioff.cpu().numpy().astype(np.int64)
    sh = O.mandelbulb(15, 2.0)
    n1 = res + 1
    tot = dict(corners=0, c_diff=0, c_hard=0, verts=0, d_diff=0, n_diff=0, hard=0)
    for a in range(0, len(lv), 40000):
        b = min(len(lv), a + 40000)
        L = lv[a:b]
        d_gpu = dists[a:b].cpu().numpy()
        d_o, _, r, dr = O.de_points(sh, corner_points(L, res))
        dg = d_gpu.reshape(-1)
        assert (np.signbit(dg) == np.signbit(d_o)).all(), f"leaves {a}..{b}: a corner sign differs"
        ln_cr, hard = R.cr_logf(r)
        with np.errstate(all="ignore"):
            d_cr = (F32(0.5) * (ln_cr * r)) / dr
        same = same_bits(dg, d_cr) | (np.isnan(dg) & np.isnan(d_cr))
        assert (same | hard).all(), f"leaves {a}..{b}: {int((~same & ~hard).sum())} corner distances differ"
        tot["corners"] += dg.size
        tot["c_diff"] += int((~same).sum())
        tot["c_hard"] += int(hard.sum())
        del d_o, r, dr, ln_cr, hard, d_cr, same
        voff_o, ioff_o, v_o, ix_o = O.mesh_batch(sh, L, res, dists=d_gpu)
        assert (vo[a:b + 1] - vo[a] == voff_o).all() and (io[a:b + 1] - io[a] == ioff_o).all()
        assert (idx[io[a]:io[b]].cpu().numpy().view(np.uint32) == ix_o.view(np.uint32)).all()
        v = verts[vo[a]:vo[b]].cpu().numpy()
        assert same_bits(v[:, :3], v_o[:, :3]).all(), f"leaves {a}..{b}: a vertex position differs"
        d_cr, _, d_hard = _cr_distances(sh, v[:, :3])
        n_cr, n_hard = _cr_fd_normals(sh, L, voff_o, v, res)
        d_same = same_bits(v[:, 6], d_cr) | (np.isnan(v[:, 6]) & np.isnan(d_cr))
        n_same = (same_bits(v[:, 3:6], n_cr) | (np.isnan(v[:, 3:6]) & np.isnan(n_cr))).all(1)
        assert (d_same | d_hard).all() and (n_same | n_hard).all(), f"leaves {a}..{b}: a vertex differs"
        tot["verts"] += len(v)
        tot["d_diff"] += int((~d_same).sum())
        tot["n_diff"] += int((~n_same).sum())
        tot["hard"] += int((d_hard | n_hard).sum())
    print(f"{tot['corners']} corner distances: {tot['c_diff']} differ from the formula, all among the "
          f"{tot['c_hard']} hard cases for rounding; {tot['verts']} vertices (positions, indices identical): "
          f"{tot['d_diff']} distances and {tot['n_diff']} normals differ, all among the {tot['hard']} with a "
          f"hard case")
    assert tot["corners"] == 921604 * n1 ** 3 and tot["verts"] == int(vo[-1])


def same_bits(a, b):
    return np.asarray(a, F32).view(np.uint32) == np.asarray(b, F32).view(np.uint32)


def finite_normals(verts):
    import torch
    n = verts[:, 3:6]
    return float(torch.isfinite(n).all(1).to(torch.float64).mean())


def test_camera_trace_full_size(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    torch = torch_cuda
    tc = wl.TRACE_1080P
    W, H = tc["W"], tc["H"]
    cams = wl.orbit_poses(tc["poses"], tc["seed"], W, H, tc["fov_deg"])
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    depth, steps = dev.render_depth(shape, gpu(torch, cams), W, H)
    torch.cuda.synchronize()
    rows = np.random.default_rng(3).integers(0, H, len(cams))
    same, agree, n = 0, 0, 0
    for i, cam in enumerate(cams):
        dirs = O.ray_dirs(cam, W, H, int(rows[i]), 1)
        d_o, s_o = O.sphere_trace(O.mandelbulb(15, 2.0), cam[:3][None], dirs, 256, 1e-4, 10.0)
        d_g = depth[i, int(rows[i])].cpu().numpy()
        s_g = steps[i, int(rows[i])].cpu().numpy().astype(np.uint32)
        same += int((s_g == s_o).sum())
        agree += int((np.isfinite(d_g) == np.isfinite(d_o)).sum())
        n += W
    print(f"{n} sampled rays: steps equal {same / n:.5f}, hit agreement {agree / n:.5f}")
    assert same / n > 0.995 and agree / n > 0.999
    assert bool(torch.isfinite(depth).any()) and bool((steps > 0).all())


def test_camera_trace_every_ray(ctx, torch_cuda):
    """Every ray of the 32 x 1080p render (66.4 M rays) against the oracle's
    march with the final ln correctly rounded, as the device computes it:
    step counts and depths bit-identical on every ray whose march met no hard
    case for rounding (the oracle marks those)."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    torch = torch_cuda
    tc = wl.TRACE_1080P
    W, H = tc["W"], tc["H"]
    cams = wl.orbit_poses(tc["poses"], tc["seed"], W, H, tc["fov_deg"])
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    depth, steps = dev.render_depth(shape, gpu(torch, cams), W, H)
    torch.cuda.synchronize()
    sh = O.mandelbulb(15, 2.0)
    n = n_marked = n_diff = n_hit = 0
    with O.ln_mode("cr"):
        for i, cam in enumerate(cams):
            dirs = O.ray_dirs(cam, W, H)
            d_o, s_o, marked = O.sphere_trace(sh, cam[:3][None], dirs, 256, 1e-4, 10.0, marked=True)
            d_g = depth[i].reshape(-1).cpu().numpy()
            s_g = steps[i].reshape(-1).cpu().numpy().astype(np.uint32)
            same = (s_g == s_o) & same_bits(d_g, d_o)
            assert (same | marked).all(), f"pose {i}: {int((~same & ~marked).sum())} rays differ"
            n += len(dirs)
            n_marked += int(marked.sum())
            n_diff += int((~same).sum())
            n_hit += int(np.isfinite(d_o).sum())
    print(f"{n} rays ({n_hit} hits): steps and depth bit-identical on all but {n_diff}, all among the "
          f"{n_marked} whose march met a hard case for rounding")
    assert n == tc["poses"] * W * H


def test_shell_normals_full_size(ctx, torch_cuda):
    """2^22 shell points, power U[6,10] (the normals workload at full size):
    the forward-mode normals of every point against the FP64 dual-number
    gradient, with the 6-tap central difference (h = 1e-3) beside them.
    Points whose 6-tap stencil changes the iteration count (the distance
    estimate has a kink there, so no difference quotient can follow the
    gradient) are counted and the errors are also given without them."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    torch = torch_cuda
    sc = wl.SHELL_NORMALS
    pts, powers = wl.shell_points(sc["n"], sc["seed"])
    shape = ca.MandelbulbTrig(15, 2.0, 8.0, ctx=ctx)
    P, W = gpu(torch, pts), gpu(torch, powers)
    _, n_tan = dev.normals(shape, P, W)
    _, n_fd = dev.normals_fd(shape, P, sc["fd_h"], W)
    torch.cuda.synchronize()
    sel = np.arange(sc["n"])
    sel_t = torch.from_numpy(sel).cuda()
    orc = O.mandelbulb_trig(15, 2.0)
    _, g64, _ = O.de_grad_f64(orc, pts[sel], powers[sel])
    ok = np.isfinite(g64[:, :3]).all(1) & (np.linalg.norm(g64[:, :3], axis=1) > 0)
    # stencil points as the FD kernel forms them: p + unit_a * (+-h)
    _, it0, _, _ = O.de_points(orc, pts[sel], powers[sel])
    steady = np.ones(len(sel), bool)
    h = F32(sc["fd_h"])
    for a in range(3):
        unit = np.zeros(3, F32)
        unit[a] = 1.0
        for hh in (h, -h):
            _, it_q, _, _ = O.de_points(orc, pts[sel] + unit * hh, powers[sel])
            steady &= it_q == it0

    def angles(nrm):
        n = nrm[sel_t].cpu().numpy().astype(np.float64)[ok]
        g = g64[ok, :3]
        c = (n * g).sum(1) / (np.linalg.norm(n, axis=1) * np.linalg.norm(g, axis=1))
        return np.arccos(np.clip(c, -1, 1))

    st = steady[ok]
    stats = {}
    for tag, nrm in (("tangent", n_tan), ("fd", n_fd)):
        ang = angles(nrm)
        stats[tag] = (np.mean(ang <= 1e-3), np.mean(ang[st] <= 1e-3))
        print(f"{tag}: angle p50 {np.median(ang):.2e} p99 {np.percentile(ang, 99):.2e}, <=1e-3 on "
              f"{stats[tag][0]:.4f}; without the {int((~st).sum())} points whose stencil changes the "
              f"iteration count: p99 {np.percentile(ang[st], 99):.2e}, <=1e-3 on {stats[tag][1]:.4f}")
    assert stats["tangent"][0] > 0.99 and stats["tangent"][1] > 0.99


def test_mixed_jobs_full_size(ctx, torch_cuda):
    """The 64 mixed jobs in one persistent launch (56.4 M points): every job's
    completion stamp set, and every point of every job against the oracle and
    the FP64 dual gradient (the bar of check_trig_job)."""
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    torch = torch_cuda
    N, Pw, cap = wl.mixed_jobs(**{k: v for k, v in wl.MIXED_JOBS.items() if k != "bailout"})
    jobs = [dict(n=int(n), power=float(p), max_iters=int(c)) for n, p, c in zip(N, Pw, cap)]
    js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB_TRIG, 2.0, jobs)
    d = torch.empty(js.total, device="cuda")
    nr = torch.empty((js.total, 3), device="cuda")
    done = torch.zeros(js.njobs, dtype=torch.int64, device="cuda")
    start = torch.zeros(1, dtype=torch.int64, device="cuda")
    js.run(d, nr, done, start)
    torch.cuda.synchronize()
    assert js.total == int((np.asarray(N, np.int64) ** 3).sum())
    ticks = done.cpu().numpy().astype(np.uint64)
    assert (ticks > np.uint64(start.cpu().numpy()[0])).all()
    from test_gpu_jobs import check_trig_job
    d_all = d.cpu().numpy()
    nr_all = nr.cpu().numpy()
    for i, j in enumerate(jobs):
        a, b = js.offsets[i], js.offsets[i + 1]
        check_trig_job(d_all[a:b], nr_all[a:b], j, tag=f"job {i}: ")
    js.close()
