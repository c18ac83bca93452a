"""A point batch beyond 2^31 points (24 GiB in, 8 GiB out).
`cth_min_distance_dev` runs it as launches of 2^30 points, so this batch of
2^31 + 12,345 points takes three.  The points are a seeded 2^20-point tile
repeated, so every copy of the tile must give the tile's own distances bit
for bit: index or offset arithmetic near 2^31, or at a launch boundary, that
dropped, repeated or moved a point would break that."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_points_beyond_2_31(ctx, torch_cuda):
    import cantucci_amd as ca
    from cantucci_amd import device as dev
    torch = torch_cuda
    shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
    n_tile = 1 << 20
    n = (1 << 31) + 12345
    tile = torch.from_numpy(np.random.default_rng(9).uniform(-1.3, 1.3, (n_tile, 3)).astype(np.float32)).cuda()
    d_tile = dev.de_points(shape, tile)
    reps = (n + n_tile - 1) // n_tile
    pts = tile.repeat(reps, 1)[:n]
    d = dev.de_points(shape, pts)
    torch.cuda.synchronize()
    del pts
    full = (n // n_tile) * n_tile
    bits = d[:full].view(torch.int32).view(-1, n_tile)
    same = bits == d_tile.view(torch.int32)[None, :]
    assert bool(same.all()), f"{int((~same).sum())} distances differ from their tile's"
    assert torch.equal(d[full:].view(torch.int32), d_tile[: n - full].view(torch.int32))
