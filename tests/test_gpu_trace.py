"""GPU sphere tracing against the oracle."""
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
F32 = np.float32


def gpu(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def oracle_render(shape, cams, W, H, max_steps=256, eps_rel=1e-4, t_max=10.0):
    depth, steps = [], []
    for cam in cams:
        dirs = O.ray_dirs(cam, W, H)
        d, s = O.sphere_trace(shape, cam[:3][None], dirs, max_steps, eps_rel, t_max)
        depth.append(d)
        steps.append(s)
    return np.stack(depth).reshape(len(cams), H, W), np.stack(steps).reshape(len(cams), H, W)


def test_sphere_rays_bitexact(ctx, torch_cuda):
    """A sphere's distance is exact arithmetic, so the GPU trace equals the
    oracle's bit for bit (depth, steps and final position)."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    rng = np.random.default_rng(1)
    dirs = rng.normal(size=(20000, 3))
    dirs[:, 1] = np.abs(dirs[:, 1]) * 3
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(F32)
    origin = np.array([[0.1, -3.0, 0.2]], F32)
    shape = ca.Sphere((0.0, 0.0, 0.0), 0.8, ctx=ctx)
    d, s, p = dev.sphere_trace(shape, gpu(torch, origin), gpu(torch, dirs), 256, 0.0, 1e-4, 10.0, with_pos=True)
    torch.cuda.synchronize()
    d_o, s_o, p_o = O.sphere_trace(O.sphere((0, 0, 0), 0.8), origin, dirs, 256, 1e-4, 10.0, 0.0, with_pos=True)
    assert (s.cpu().numpy().astype(np.uint32) == s_o).all()
    assert (d.cpu().numpy().view(np.uint32) == d_o.view(np.uint32)).all()
    assert (p.cpu().numpy().view(np.uint32) == p_o.view(np.uint32)).all()
    assert np.isfinite(d_o).any() and np.isinf(d_o).any()


def test_camera_render_matches_oracle(ctx, torch_cuda):
    """Primary rays of the 32 seeded orbit poses at reduced resolution, P8
    Mandelbulb, 256 steps, hit when d < 1e-4 t, escape at t > 10."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import orbit_poses
    torch = torch_cuda
    W, H = 48, 27
    cams = orbit_poses(32, seed=0, W=W, H=H)
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    depth, steps = dev.render_depth(shape, gpu(torch, cams), W, H)
    torch.cuda.synchronize()
    depth, steps = depth.cpu().numpy(), steps.cpu().numpy().astype(np.uint32)
    d_o, s_o = oracle_render(O.mandelbulb(15, 2.0), cams, W, H)
    hit_g, hit_o = np.isfinite(depth), np.isfinite(d_o)
    same_steps = steps == s_o
    both = hit_g & hit_o & same_steps
    rel = np.abs(depth[both].astype(np.float64) - d_o[both]) / d_o[both]
    bit = depth.view(np.uint32) == d_o.view(np.uint32)
    print(f"rays {depth.size}: hits {hit_o.mean():.3f}, hit agreement {np.mean(hit_g == hit_o):.5f}, "
          f"steps equal {same_steps.mean():.5f}, depth bit-equal {bit.mean():.5f}, "
          f"max rel depth (equal steps) {rel.max() if rel.size else 0:.2e}")
    assert hit_o.mean() > 0.05
    assert np.mean(hit_g == hit_o) > 0.999
    # A step-count difference needs a hit test d < 1e-4 t decided differently
    # by an ulp of the final ln; it is rare and is the only source of larger
    # depth differences.
    assert same_steps.mean() > 0.995
    assert np.percentile(rel, 99.9) < 1e-5


def test_render_graph_replay(ctx, torch_cuda):
    """The 32-pose render captured in a HIP graph replays to the eager result
    every time, with eager work and synchronisation between the replays."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import orbit_poses
    torch = torch_cuda
    W, H = 96, 54
    cams = gpu(torch, orbit_poses(32, seed=0, W=W, H=H))
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    ref_d, ref_s = dev.render_depth(shape, cams, W, H)
    torch.cuda.synchronize()
    depth = torch.empty_like(ref_d)
    steps = torch.empty_like(ref_s)
    s = torch.cuda.Stream()
    g = dev.Graph(lambda: dev.render_depth(shape, cams, W, H, depth=depth, steps=steps), s)
    for rep in range(3):
        depth.fill_(-1.0)
        steps.fill_(-1)
        torch.cuda.synchronize()
        g.launch()
        s.synchronize()
        assert torch.equal(depth.nan_to_num(-2.0), ref_d.nan_to_num(-2.0)), rep
        assert torch.equal(steps, ref_s), rep
        dev.de_points(shape, torch.rand((5000, 3), device="cuda"))  # eager work between replays
        torch.cuda.synchronize()
    g.close()


def _ref_cameras(n, seed, fov=1.0, aspect=16 / 9, near=4e-5):
    """Reference cameras (position, unit direction, fov, aspect, near, far)
    around the Mandelbulb: positions at radius 1.6-3.5 looking near the origin."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        v = rng.normal(size=3)
        pos = v / np.linalg.norm(v) * rng.uniform(1.6, 3.5)
        d = -pos + rng.uniform(-0.2, 0.2, 3)
        d = (d / np.linalg.norm(d)).astype(F32)
        out.append(np.concatenate([pos.astype(F32), d, np.array([fov, aspect, near, 10.0], F32)]))
    return np.array(out, F32)


@pytest.mark.parametrize("near,fp", [(4e-5, 5), (0.1, 5), (0.01, 4), (4e-5, 1)])
def test_focus_rays_bitexact(ctx, torch_cuda, near, fp):
    """get_focii's rays built on the device from the reference camera
    (src/mesh/mod.rs:205-231, src/camera.rs:69-103,186-189) equal the
    oracle's bit for bit (fov 1 rad, the app's: its tangent is correctly
    rounded by glibc too); on a sphere, whose distance is exact arithmetic,
    the marched focus points, depths and step counts are bit-identical."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    cams = _ref_cameras(16, 5 + fp, near=near)
    shape = ca.Sphere((0.05, -0.02, 0.01), 0.9, ctx=ctx)
    focii, depth, steps, dirs = dev.get_focii(shape, gpu(torch, cams), fp, with_dirs=True)
    torch.cuda.synchronize()
    focii, depth, steps, dirs = (t.cpu().numpy() for t in (focii, depth, steps, dirs))
    sh = O.sphere((0.05, -0.02, 0.01), 0.9)
    for i, cam in enumerate(cams):
        pos_o, depth_o, steps_o, dirs_o = O.get_focii(sh, cam, fp)
        assert (dirs[i].view(np.uint32) == dirs_o.view(np.uint32)).all(), i
        assert (steps[i].astype(np.uint32) == steps_o).all(), i
        assert (depth[i].view(np.uint32) == depth_o.view(np.uint32)).all(), i
        hit = np.isfinite(depth_o)
        assert (focii[i][hit].view(np.uint32) == pos_o[hit].view(np.uint32)).all(), i
        # a miss (get_focii drops it) marches off to +-inf and NaN within its
        # 100 steps; a NaN the arithmetic creates has the sign bit set on x86
        # and clear on the GPU, so misses compare up to the NaN sign
        fm, om = focii[i][~hit], pos_o[~hit]
        assert (np.isnan(fm) == np.isnan(om)).all() and (fm[~np.isnan(fm)] == om[~np.isnan(om)]).all(), i
    assert np.isfinite(depth).any()
    if fp > 1:
        assert np.isinf(depth).any()


def test_get_focii_mandelbulb(ctx, torch_cuda):
    """get_focii with the app's shape and projection (classic(6, 2.5), fov
    1 rad, near 4e-5: src/app.rs:98,105), 64 cameras x 25 rays: directions
    bit-identical to the oracle; the marched points agree except where the
    final ln (correctly rounded here, glibc's logf in the reference) differs
    by an ulp on some step, which can move a ray by an ulp from then on."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    cams = _ref_cameras(64, 11)
    shape = ca.Mandelbulb.classic(6, 2.5, ctx=ctx)
    focii, depth, steps, dirs = dev.get_focii(shape, gpu(torch, cams), 5, with_dirs=True)
    torch.cuda.synchronize()
    focii, depth, steps, dirs = (t.cpu().numpy() for t in (focii, depth, steps, dirs))
    sh = O.mandelbulb(6, 2.5)
    res = [O.get_focii(sh, cam, 5) for cam in cams]
    pos_o = np.stack([r[0] for r in res])
    depth_o = np.stack([r[1] for r in res])
    steps_o = np.stack([r[2] for r in res])
    dirs_o = np.stack([r[3] for r in res])
    assert (dirs.view(np.uint32) == dirs_o.view(np.uint32)).all()
    hit, hit_o = np.isfinite(depth), np.isfinite(depth_o)
    same = steps.astype(np.uint32) == steps_o
    both = hit & hit_o
    dpos = np.linalg.norm(focii[both].astype(np.float64) - pos_o[both], axis=1)
    bit = (focii.view(np.uint32) == pos_o.view(np.uint32)).all(-1)
    print(f"focus rays {depth.size}: hits {hit_o.mean():.3f}, hit agreement {np.mean(hit == hit_o):.4f}, "
          f"steps equal {same.mean():.4f}, points bit-equal {bit.mean():.4f}, "
          f"max |dpos| (both hit) {dpos.max() if dpos.size else 0:.2e}")
    assert hit_o.mean() > 0.2
    assert np.mean(hit == hit_o) > 0.97
    assert same.mean() > 0.9
    assert np.percentile(dpos, 99) < 1e-4


def test_get_focii_host_mirror(ctx):
    """ShapeMesh::get_focii through the host entry point: the hit points in
    ray order (the reference's filter_map), equal to the device batch."""
    import cantucci_amd as ca
    from cantucci_amd.mesh import get_focii
    shape = ca.Sphere((0.0, 0.0, 0.0), 0.8, ctx=ctx)
    pts = get_focii(shape, (2.5, 0.4, 0.3), (-2.5, -0.4, -0.3), 0.5, 1.5, 0.1, 10.0, 5)
    d = np.array([-2.5, -0.4, -0.3], F32)
    d = d * (F32(1.0) / np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
    cam = np.concatenate([np.array([2.5, 0.4, 0.3], F32), d, np.array([0.5, 1.5, 0.1, 10.0], F32)])
    pos_o, depth_o, _, _ = O.get_focii(O.sphere((0, 0, 0), 0.8), cam, 5)
    exp = pos_o[np.isfinite(depth_o)]
    assert 0 < len(pts) < 25
    assert (pts.view(np.uint32) == exp.view(np.uint32)).all()
    with pytest.raises(ValueError):
        get_focii(shape, (1, 0, 0), (0, 0, 0), 1.0, 1.0, 0.1, 10.0)


def test_trace_rejects_bad_arguments(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    with pytest.raises(ca.CthError):
        dev.sphere_trace(shape, torch.zeros((1, 3), device="cuda"), torch.ones((4, 3), device="cuda"), 0)
    d, s = dev.sphere_trace(shape, torch.zeros((1, 3), device="cuda"), torch.ones((0, 3), device="cuda"), 10)
    assert d.numel() == 0
    cams = torch.zeros((1, 10), device="cuda")
    for fp, steps in ((256, 100), (5, 0)):
        with pytest.raises(ca.CthError):
            dev.get_focii(shape, cams, fp, max_steps=steps)
    # no focus points: no rays and no hits, as the reference's iter::square(0)
    f, d, s = dev.get_focii(shape, cams, 0)
    assert f.shape == (1, 0, 3) and d.numel() == 0
    from cantucci_amd.mesh import get_focii
    assert get_focii(shape, (-3, 0, 0), (1, 0, 0), 1.0, 16 / 9, 0.00004, 10.0, focus_points=0).shape == (0, 3)
    with pytest.raises(ValueError):
        dev.get_focii(shape, torch.zeros((1, 9), device="cuda"))
