"""Batched jobs in one persistent launch, and HIP-graph capture/replay."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
F32 = np.float32


def angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = (a * b).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
    return np.arccos(np.clip(c, -1, 1))


def check_trig_job(d_job, nr_job, job, sample=None, seed=0, tag=""):
    """Parity bar of one real-power grid job (its points p = -1.5 + 3 i / (n - 1)
    in f32, GridTable order): on the points whose f32 reference distance is
    ulp-stable (a one-ulp move of any coordinate changes it by < 1e-6
    relative), fewer than 0.1% differ from the oracle by more than 1e-5
    relative -- the bar of the other trig tests; elsewhere the f32 orbit is
    chaotic and transcendentals that differ in the last ulp (ocml here,
    glibc in the reference) cannot agree.  The forward-mode normals of the
    same ulp-stable points are within 1e-3 rad of the FP64 dual-number
    gradient on at least 99% of them.  `sample` limits the check to that many
    seeded points of the job."""
    from test_gpu_de import grid_points, ulp_stable
    n = job["n"]
    pts = grid_points(n, n - 1.0)
    sel = np.arange(len(pts))
    if sample is not None and sample < len(pts):
        sel = np.sort(np.random.default_rng(seed).choice(len(pts), sample, replace=False))
    shape = O.mandelbulb_trig(job["max_iters"], 2.0, job["power"])
    q = pts[sel]
    dref = O.de_points(shape, q)[0]
    dg = np.asarray(d_job)[sel]
    rel = np.abs(dg.astype(np.float64) - dref) / np.maximum(np.abs(dref.astype(np.float64)), 1e-30)
    cond = ulp_stable(shape, q, dref)
    _, g64, _ = O.de_grad_f64(shape, q)
    ok = cond & np.isfinite(g64[:, :3]).all(1) & (np.linalg.norm(g64[:, :3], axis=1) > 0)
    ang = angle(np.asarray(nr_job)[sel][ok], g64[ok, :3])
    print(f"{tag}n={n} power={job['power']:.3f} cap={job['max_iters']}: {len(sel)} points, ulp-stable "
          f"{cond.mean():.3f}, of which rel>1e-5 {np.mean(rel[cond] > 1e-5):.5f}; normals <=1e-3 rad "
          f"{np.mean(ang <= 1e-3):.4f} (p99 {np.percentile(ang, 99) if ang.size else 0:.2e})")
    assert cond.mean() > 0.5
    assert np.mean(rel[cond] > 1e-5) < 1e-3
    assert np.mean(ang <= 1e-3) >= 0.99


def make_jobs(kind):
    if kind == "p8":
        return [dict(n=n, power=8.0, max_iters=cap) for n, cap in ((9, 15), (16, 8), (12, 30), (20, 11), (7, 1))]
    return [dict(n=n, power=p, max_iters=cap)
            for n, p, cap in ((9, 2.5, 15), (16, 7.25, 8), (12, 11.5, 30), (20, 4.0, 11), (13, 9.75, 20))]


def test_p8_jobs_match_oracle(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    torch = torch_cuda
    jobs = make_jobs("p8")
    js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB, 2.0, jobs)
    dists = torch.empty(js.total, device="cuda")
    nrm = torch.empty((js.total, 3), device="cuda")
    done = torch.zeros(js.njobs, dtype=torch.int64, device="cuda")
    start = torch.zeros(1, dtype=torch.int64, device="cuda")
    js.run(dists, nrm, done, start)
    torch.cuda.synchronize()
    d, nr = dists.cpu().numpy(), nrm.cpu().numpy()
    for i, j in enumerate(jobs):
        a, b = js.offsets[i], js.offsets[i + 1]
        dref, _ = O.de_grid(O.mandelbulb(j["max_iters"], 2.0), (-1.5,) * 3, (1.5,) * 3, j["n"], j["n"] - 1.0)
        same = (d[a:b].view(np.uint32) == dref.view(np.uint32)) | (np.isnan(d[a:b]) & np.isnan(dref))
        assert same.mean() > 0.99, (i, same.mean())
        # normals against the FP64 dual gradient of the same grid points
        t = np.arange(j["n"], dtype=F32) / F32(j["n"] - 1.0)
        ax = F32(-1.5) + F32(3.0) * t
        X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        _, g64, _ = O.de_grad_f64(O.mandelbulb(j["max_iters"], 2.0), pts)
        ok = np.isfinite(g64).all(1) & (np.linalg.norm(g64[:, :3], axis=1) > 0)
        ang = angle(nr[a:b][ok], g64[ok, :3])
        assert np.median(ang) < 1e-5 and np.mean(ang < 1e-3) > 0.95, (i, np.median(ang))
    ticks = done.cpu().numpy().astype(np.uint64)
    s0 = np.uint64(start.cpu().numpy()[0])
    assert (ticks > s0).all()
    js.close()


def test_trig_jobs_statistics(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    torch = torch_cuda
    jobs = make_jobs("trig")
    js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB_TRIG, 2.0, jobs)
    dists = torch.empty(js.total, device="cuda")
    nrm = torch.empty((js.total, 3), device="cuda")
    done = torch.zeros(js.njobs, dtype=torch.int64, device="cuda")
    start = torch.zeros(1, dtype=torch.int64, device="cuda")
    js.run(dists, nrm, done, start)
    torch.cuda.synchronize()
    d, nr = dists.cpu().numpy(), nrm.cpu().numpy()
    for i, j in enumerate(jobs):
        a, b = js.offsets[i], js.offsets[i + 1]
        dref, _ = O.de_grid(O.mandelbulb_trig(j["max_iters"], 2.0, j["power"]), (-1.5,) * 3, (1.5,) * 3, j["n"],
                            j["n"] - 1.0)
        from test_gpu_de import grid_points
        assert np.array_equal(O.de_points(O.mandelbulb_trig(j["max_iters"], 2.0, j["power"]),
                                          grid_points(j["n"], j["n"] - 1.0))[0].view(np.uint32),
                              dref.view(np.uint32))
        check_trig_job(d[a:b], nr[a:b], j, tag=f"job {i}: ")
    js.close()


def test_jobs_graph_replay(ctx, torch_cuda):
    """A captured jobs launch replays to the eager result every time, with
    eager work and synchronisation between replays."""
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    torch = torch_cuda
    jobs = make_jobs("p8")
    js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB, 2.0, jobs)
    ref = torch.empty(js.total, device="cuda")
    done = torch.zeros(js.njobs, dtype=torch.int64, device="cuda")
    start = torch.zeros(1, dtype=torch.int64, device="cuda")
    js.run(ref, None, done, start)
    torch.cuda.synchronize()
    out = torch.full((js.total,), float("nan"), device="cuda")
    s = torch.cuda.Stream()
    g = dev.Graph(lambda: js.run(out, None, done, start, stream=s.cuda_stream), s)
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    for rep in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        s.synchronize()
        assert torch.equal(out.nan_to_num(7.0), ref.nan_to_num(7.0)), rep
        dev.de_points(shape, torch.rand((1000, 3), device="cuda"))  # eager work between replays
        torch.cuda.synchronize()
    g.close()
    js.close()


def test_grid_graph_replay(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    ref = dev.de_grid(shape, (-1.5,) * 3, (1.5,) * 3, 64, 63.0)
    torch.cuda.synchronize()
    out = torch.empty_like(ref)
    s = torch.cuda.Stream()

    def body():
        import ctypes as C
        from cantucci_amd._lib import check, lib
        st = (C.c_float * 3)(-1.5, -1.5, -1.5)
        en = (C.c_float * 3)(1.5, 1.5, 1.5)
        for _ in range(4):
            check(lib.cth_de_grid_dev(ctx.handle, C.byref(shape.desc), st, en, 64, 63.0, out.data_ptr(), None,
                                      s.cuda_stream))

    g = dev.Graph(body, s)
    for rep in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.launch()
        s.synchronize()
        assert torch.equal(out.nan_to_num(7.0), ref.nan_to_num(7.0)), rep
    g.close()


def test_jobs_reject_bad_input(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    with pytest.raises(ca.CthError):
        dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB, 2.0, [dict(n=4, power=7.0, max_iters=3)])
    with pytest.raises(ca.CthError):
        dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB_TRIG, 2.0, [dict(n=4, power=7.0, max_iters=0)])
    with pytest.raises(ca.CthError):
        dev.JobSet(ctx, _lib.CTH_SHAPE_SPHERE, 2.0, [dict(n=4, power=7.0, max_iters=3)])


def test_graph_counters_not_shared_with_eager_launches(torch_cuda):
    """A captured launch keeps its work counters for every replay.  Those
    belong to the graph, not to the context's ring of 4096 pairs, so an eager
    launch on another stream that the ring gives the same slot (the 4096th
    launch after the capture) can run concurrently with a replay without
    either skipping or repeating work.  Span sampling claims chunks from its
    counter (dynamic schedule)."""
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import half_diagonals
    torch = torch_cuda
    c = ca.Context(0, "parity")
    c.set_launch(_lib.CTH_KERNEL_SPANS, 256, 8, 256, 32, dynamic=1)
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=c)
    leaves, _ = O.refine(O.mandelbulb(15, 2.0), [-1.5, -1.5, -1.5, 1.5, 1.5, 1.5], 2, 6, half_diagonals(3.0, 6))
    spans = torch.from_numpy(leaves[:3000].copy()).cuda()
    tiny = spans[:1].clone()
    ref, _ = dev.sample_spans(shape, spans, 8, with_cross=False)
    torch.cuda.synchronize()
    out_g = torch.empty_like(ref)
    out_e = torch.empty_like(ref)
    tiny_out = torch.empty((1, 729), device="cuda")
    s, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    g = dev.Graph(lambda: dev.sample_spans(shape, spans, 8, with_cross=False, out=out_g), s)
    for rep in range(2):
        out_g.fill_(float("nan"))
        out_e.fill_(float("nan"))
        torch.cuda.synchronize()
        # 4095 eager launches after the capture (or the previous overlap), so
        # that the next eager launch would take the graph's ring slot
        for _ in range(4095):
            dev.sample_spans(shape, tiny, 8, with_cross=False, out=tiny_out)
        torch.cuda.synchronize()
        g.launch()
        with torch.cuda.stream(s2):  # another non-blocking stream: overlaps the replay
            dev.sample_spans(shape, spans, 8, with_cross=False, out=out_e)
        torch.cuda.synchronize()
        assert torch.equal(out_g.nan_to_num(7.0), ref.nan_to_num(7.0)), rep
        assert torch.equal(out_e.nan_to_num(7.0), ref.nan_to_num(7.0)), rep
    g.close()
    c.close()
