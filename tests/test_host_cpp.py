"""The C++ host mirror of the `Shape` trait, `MeshBuffer::generate_for_box` and `ShapeMesh`
(cantucci_amd/host/cantucci.hpp), exercised by the compiled program
tests/cpp/host_parity against the oracle."""
import subprocess
from pathlib import Path

import pytest

PROG = Path(__file__).resolve().parent / "cpp" / "host_parity"


def _run(timeout):
    if not PROG.exists():
        pytest.fail(f"{PROG} is not built (run __graft_entry__.build())")
    return subprocess.run([str(PROG)], capture_output=True, text=True, timeout=timeout)


@pytest.mark.gpu
def test_cpp_host_mirror_parity():
    res = _run(300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr


def test_cpp_host_refuses_without_device():
    """Without a gfx950 device the host layer fails loudly (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    res = _run(120)
    assert res.returncode == 2 and "cantucci error 3" in res.stdout, res.stdout + res.stderr
