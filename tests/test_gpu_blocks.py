"""Device building blocks checked exhaustively on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
HIP = Path(__file__).resolve().parent / "hip"


def test_wave_sqrt_matches_correctly_rounded_sqrt_on_all_floats():
    """sqrt_rn_wave (the hot loop's square root) is bit-identical to the
    correctly rounded __builtin_sqrtf on all 2^32 f32 patterns, with waves of
    one range and with waves mixing both paths."""
    prog = HIP / "sqrt_exhaustive"
    assert prog.exists(), "tests/hip is not built (run __graft_entry__.build())"
    res = subprocess.run([str(prog)], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr


def test_unscaled_division_matches_correctly_rounded_division():
    """div_rn_unscaled (the P8 loop body's division when every lane is in
    range) is bit-identical to the correctly rounded `a / b` on all 2^46
    pairs of significands, and on 2^32 random pairs over the body's operand
    range (there, 1 + a/b identical always; a/b itself wherever a and a/b
    are at least 2^-100)."""
    prog = HIP / "div_check"
    assert prog.exists(), "tests/hip is not built (run __graft_entry__.build())"
    res = subprocess.run([str(prog)], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
