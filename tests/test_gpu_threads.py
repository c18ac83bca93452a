"""One context per host thread, as the reference shares its shape among pool
threads (src/mesh/mod.rs:61-62,141-142): four threads, each with its own
context and its own stream, run point queries, span sampling and meshing at
the same time, and one of them also hits an error.  Each thread's outputs are
bit-identical to the same work run alone, and each thread reads its own
error message (cth_last_error is per thread)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _work(torch, i, fail):
    import cantucci_amd as ca
    from cantucci_amd import _lib
    from cantucci_amd import device as dev
    from test_gpu_mesh import octree_leaves
    ctx = ca.Context(0, "parity")
    s = torch.cuda.Stream()
    out = {}
    try:
        with torch.cuda.stream(s):
            shape = ca.Mandelbulb.classic(15, 2.0, ctx=ctx)
            rng = np.random.default_rng(100 + i)
            pts = torch.from_numpy(rng.uniform(-1.3, 1.3, (40000, 3)).astype(np.float32)).cuda()
            out["d"] = dev.de_points(shape, pts)
            leaves = torch.from_numpy(np.ascontiguousarray(octree_leaves(400, seed=30 + i))).cuda()
            dists, cross = dev.sample_spans(shape, leaves, 8)
            out["dists"], out["cross"] = dists, cross
            out["voff"], out["ioff"], out["verts"], out["idx"] = dev.mesh(shape, leaves, dists, 8, "fd6")
            if fail:
                with pytest.raises(ca.CthError):
                    dev.sample_spans(shape, leaves, 3)  # not a power of two
                out["error"] = _lib.lib.cth_last_error().decode()
        s.synchronize()
        return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()}
    finally:
        ctx.close()


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(f"u{a.dtype.itemsize}")


def test_contexts_on_host_threads(torch_cuda):
    torch = torch_cuda
    alone = [_work(torch, i, False) for i in range(4)]
    results, errors = [None] * 4, []

    def run(i):
        try:
            results[i] = _work(torch, i, i == 2)
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    assert all(not t.is_alive() for t in threads)
    for i in range(4):
        for k, v in alone[i].items():
            assert v.shape == results[i][k].shape and (_bits(v) == _bits(results[i][k])).all(), (i, k)
    assert "power of two" in results[2]["error"]
