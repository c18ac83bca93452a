"""Work counters of the refill kernels and the split surface-nets vertex pass."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
F32 = np.float32


def test_work_counters_match_iteration_counts(ctx, torch_cuda):
    """The counters a span-sampling launch adds equal what its iteration
    output implies: loop bodies = sum of the per-corner counts, bailed
    evaluations = corners below the cap, evaluations = corners.  A trace
    launch counts one evaluation per ray step."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from test_gpu_de import octree_leaf_sample
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    spans = torch.from_numpy(octree_leaf_sample(500)).cuda()
    cnt = torch.zeros(3, dtype=torch.int64, device="cuda")
    ctx.count_work(cnt)
    try:
        _, _, it = dev.sample_spans(shape, spans, 8, with_cross=False, with_iters=True)
        torch.cuda.synchronize()
        bodies, bails, evals = (int(v) for v in cnt.cpu())
        it = it.cpu().numpy().astype(np.int64)
        assert bodies == int(it.sum()) and evals == it.size and bails == int((it < 15).sum())
        cnt.zero_()
        from cantucci_amd.workloads import orbit_poses
        cams = torch.from_numpy(orbit_poses(2, seed=0, W=40, H=24)).cuda()
        _, steps = dev.render_depth(shape, cams, 40, 24)
        torch.cuda.synchronize()
        assert int(cnt[2]) == int(steps.to(torch.int64).sum())
        assert 0 < int(cnt[1]) <= int(cnt[2]) and int(cnt[0]) > 0
    finally:
        ctx.count_work(None)
    # counting off: the counters stay as they are
    before = cnt.clone()
    dev.sample_spans(shape, spans, 8, with_cross=False)
    torch.cuda.synchronize()
    assert torch.equal(cnt, before)


@pytest.mark.parametrize("normals", ["fd6", "tangent"])
def test_deferred_emit_then_vertex_pass(ctx, torch_cuda, normals):
    """cth_mesh_emit_dev with deferred normals followed by
    cth_mesh_vertices_dev gives the records of the one-call emit bit for bit;
    the deferred records hold the leaf's 6-tap half-steps
    0.7 (e' - s') / res (src/mesh/buffer.rs:256) and a zero distance."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from test_gpu_de import octree_leaf_sample
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    spans = torch.from_numpy(octree_leaf_sample(300, seed=3)).cuda()
    d, _ = dev.sample_spans(shape, spans, 8, with_cross=False)
    _, _, v_ref, i_ref = dev.mesh(shape, spans, d, 8, normals)
    _, _, v_def, i_def = dev.mesh(shape, spans, d, 8, "deferred")
    v_def = v_def.clone()
    torch.cuda.synchronize()
    vd = v_def.cpu().numpy()
    assert torch.equal(i_ref, i_def)
    assert (vd[:, 6] == 0).all()
    assert np.array_equal(vd[:, :3].view(np.uint32), v_ref.cpu().numpy()[:, :3].view(np.uint32))
    # half-steps of every leaf: (0.7 * (e' - s')) / res in f32, e', s' the adjusted span
    sp = spans.cpu().numpy()
    of = (sp[:, 3:] - sp[:, :3]) / F32(8)
    s2, e2 = sp[:, :3] + -of, sp[:, 3:] + of
    hs = (F32(0.7) * (e2 - s2)) / F32(8)
    assert np.isin(vd[:, 3:6].view(np.uint32), hs.view(np.uint32)).all()
    dev.mesh_vertices(shape, v_def, normals)
    torch.cuda.synchronize()
    assert np.array_equal(v_def.cpu().numpy().view(np.uint32), v_ref.cpu().numpy().view(np.uint32))
    with pytest.raises(ca.CthError):
        dev.mesh_vertices(shape, v_def, "deferred")
