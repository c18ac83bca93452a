"""The reference's own split rule (ShapeMesh::update, src/mesh/mod.rs:91-104)
on the GPU against the oracle, fed by the focus rays' hits."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
F32 = np.float32


def gpu(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_focus_split_matches_oracle(ctx, torch_cuda):
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import half_diagonals
    torch = torch_cuda
    root = [-1.2, -1.2, -1.2, 1.2, 1.2, 1.2]
    leaves, _ = O.refine(O.mandelbulb(6, 2.5), root, 2, 5, half_diagonals(2.4, 5))
    rng = np.random.default_rng(11)
    # the eye in a surface leaf, focus points around it and far away
    eye = ((leaves[100, :3] + leaves[100, 3:]) / 2).astype(F32)
    foci = np.concatenate([eye + rng.uniform(-0.2, 0.2, (3000, 3)), rng.uniform(-1.3, 1.3, (2000, 3))]).astype(F32)
    # points on leaf faces and corners: the half-open rule decides
    near = np.where(np.abs(leaves[:, :3] - eye).max(1) < 0.3)[0]
    corners = leaves[rng.choice(near, 1000), :3]
    foci = np.concatenate([foci, corners, leaves[rng.choice(near, 500), 3:]]).astype(F32)
    depth = np.where(rng.uniform(size=len(foci)) < 0.1, np.inf, 1.0).astype(F32)
    flags = dev.focus_split(ctx, gpu(torch, leaves), gpu(torch, foci), eye, depth=gpu(torch, depth))
    torch.cuda.synchronize()
    ref = O.focus_split(leaves, foci, eye, depth)
    got = flags.cpu().numpy().astype(np.uint32)
    print(f"{len(leaves)} leaves, {int(ref.sum())} split")
    assert ref.sum() > 10 and (got == ref).all()
    # in-place replacement by create_spans children
    out = dev.split_leaves(ctx, gpu(torch, leaves), flags).cpu().numpy()
    expect = []
    for i, sp in enumerate(leaves):
        if ref[i]:
            expect.extend(O.create_spans(sp))
        else:
            expect.append(sp)
    expect = np.array(expect, F32)
    assert out.shape == expect.shape and (out.view(np.uint32) == expect.view(np.uint32)).all()


def test_focus_rays_drive_the_split(ctx, torch_cuda):
    """get_focii's 25 rays traced on the device feed the rule directly."""
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    from cantucci_amd.workloads import half_diagonals, look_at_basis
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(6, 2.5, ctx=ctx)
    eye = np.array([-1.3, 0.05, 0.1], F32)
    s, u, f = look_at_basis(eye.astype(np.float64), (0.0, 0.0, 0.0))
    g = (np.arange(5) + 0.5) / 5 * 2 - 1
    dirs = np.array([f + 0.3 * (a * s + b * u) for b in g for a in g])
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(F32)
    d, steps, pos = dev.sphere_trace(shape, gpu(torch, eye[None]), gpu(torch, dirs), 100, 1e-6, 0.0, float("inf"),
                                     with_pos=True)
    leaves, _ = O.refine(O.mandelbulb(6, 2.5), [-1.2, -1.2, -1.2, 1.2, 1.2, 1.2], 2, 4, half_diagonals(2.4, 4))
    flags = dev.focus_split(ctx, gpu(torch, leaves), pos, eye, depth=d)
    torch.cuda.synchronize()
    ref = O.focus_split(leaves, pos.cpu().numpy(), eye, d.cpu().numpy())
    assert np.isfinite(d.cpu().numpy()).sum() >= 5 and ref.sum() >= 1
    assert (flags.cpu().numpy().astype(np.uint32) == ref).all()
