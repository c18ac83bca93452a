"""Results do not depend on the launch shape.  Every persistent kernel family
(points and grids, span sampling, the mesh's vertex taps, tracing, gradients,
jobs) runs under extreme launch shapes and must give outputs bit-identical to
those of the default shapes:

* one 64-thread workgroup per CU, chunks of 64 indices, refill at 1 idle lane;
* 32 workgroups of 256 threads per CU (more than fit at once: the rest run
  after the first ones leave), chunks of 4096, refill only when all 64 lanes
  are idle;
* a static interleave with 128-thread workgroups and an odd threshold of 17.

What changes between shapes is the work queue, the refill events and the
lane bookkeeping, never a point's arithmetic."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(block=64, blocks_per_cu=1, chunk=64, refill_min=1, dynamic=1),
    dict(block=256, blocks_per_cu=32, chunk=4096, refill_min=64, dynamic=1),
    dict(block=128, blocks_per_cu=4, chunk=256, refill_min=17, dynamic=0),
    dict(block=256, blocks_per_cu=16, chunk=64, refill_min=64, dynamic=0),
]


def _run_all(torch, ctx):
    """Every family's outputs on small seeded inputs, as host arrays."""
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    from cantucci_amd import workloads as wl
    from test_gpu_mesh import octree_leaves
    out = {}
    p8 = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    trig = ca.MandelbulbTrig(15, 2.0, 7.5, ctx=ctx)
    rng = np.random.default_rng(21)
    pts = torch.from_numpy(rng.uniform(-1.3, 1.3, (50000, 3)).astype(np.float32)).cuda()
    out["p8_d"], out["p8_it"] = dev.de_points(p8, pts, with_iters=True)
    out["trig_d"] = dev.de_points(trig, pts)
    out["grid"] = dev.de_grid(p8, (-1.5,) * 3, (1.5,) * 3, 32, 31.0)
    leaves = torch.from_numpy(np.ascontiguousarray(octree_leaves(1000, seed=5))).cuda()
    dists, cross = dev.sample_spans(p8, leaves, 8)
    out["dists"], out["cross"] = dists, cross
    out["voff"], out["ioff"], out["verts"], out["idx"] = dev.mesh(p8, leaves, dists, 8, "fd6")
    cams = torch.from_numpy(wl.orbit_poses(2, 3, 160, 90, 60.0)).cuda()
    out["depth"], out["steps"] = dev.render_depth(p8, cams, 160, 90)
    sp, sw = wl.shell_points(20000, 4)
    sp, sw = torch.from_numpy(sp).cuda(), torch.from_numpy(sw).cuda()
    out["tan_d"], out["tan_n"] = dev.normals(trig, sp, sw)
    out["vjp_d"], out["vjp_g"] = dev.de_vjp(trig, sp, sw)
    out["grad_d"], out["grad_g"] = dev.de_grad(trig, sp, sw)
    jobs = [dict(n=n, power=p, max_iters=c) for n, p, c in ((9, 2.5, 15), (16, 7.25, 8), (12, 11.5, 30))]
    js = dev.JobSet(ctx, _lib.CTH_SHAPE_MANDELBULB_TRIG, 2.0, jobs)
    jd = torch.empty(js.total, device="cuda")
    jn = torch.empty((js.total, 3), device="cuda")
    js.run(jd, jn, torch.zeros(js.njobs, dtype=torch.int64, device="cuda"),
           torch.zeros(1, dtype=torch.int64, device="cuda"))
    out["jobs_d"], out["jobs_n"] = jd, jn
    torch.cuda.synchronize()
    js.close()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8) if a.dtype.itemsize == 1 else a.view(f"u{a.dtype.itemsize}")


def test_outputs_do_not_depend_on_launch_shape(torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import _lib
    torch = torch_cuda
    families = (_lib.CTH_KERNEL_DE, _lib.CTH_KERNEL_TRACE, _lib.CTH_KERNEL_ADJOINT, _lib.CTH_KERNEL_SPANS,
                _lib.CTH_KERNEL_JOBS, _lib.CTH_KERNEL_MESH)
    ctx = ca.Context(0, "parity")
    try:
        base = _run_all(torch, ctx)
        for shape in SHAPES:
            for f in families:
                ctx.set_launch(f, **shape)
            got = _run_all(torch, ctx)
            for k in base:
                assert got[k].shape == base[k].shape, (shape, k)
                assert (_bits(got[k]) == _bits(base[k])).all(), f"{k} differs under {shape}"
            print(f"{shape}: all {len(base)} outputs bit-identical")
    finally:
        ctx.close()
