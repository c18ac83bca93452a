"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, and its host-only entry points behave."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(cth_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "cth_min_distance_batch" in names and "cth_sample_spans_dev" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    from cantucci_amd import _lib
    missing = [n for n in declared_symbols() if not hasattr(_lib.lib, n)]
    assert not missing, f"not exported: {missing}"


def _prototypes():
    """(return type, name, [parameter types]) of every cth_* prototype in the
    public headers."""
    out = []
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        pat = r"\b((?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**)\s*\b(cth_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;"
        for ret, name, args in re.findall(pat, text, flags=re.S):
            a = " ".join(args.split())
            out.append((" ".join(ret.split()), name, [] if a in ("", "void") else [p.strip() for p in a.split(",")]))
    return out


def _c_kind(t):
    """Kind of a C parameter or return type: 'ptr' for pointers and arrays,
    else the scalar's width and class (size_t is 64-bit on this ABI)."""
    t = " ".join(t.replace("const", " ").split())
    if "*" in t or "[" in t:
        return "ptr"
    base = t.split()[0] if t.split()[0] != "unsigned" else t
    return {"int": "i32", "int32_t": "i32", "float": "f32", "uint32_t": "u32", "size_t": "u64", "uint64_t": "u64",
            "double": "f64", "void": "void"}.get(base, base)


def _py_kind(a):
    if a is None:
        return "void"
    if a in (C.c_void_p, C.c_char_p) or (isinstance(a, type) and issubclass(a, C._Pointer)):
        return "ptr"
    return {C.c_int: "i32", C.c_float: "f32", C.c_uint32: "u32", C.c_uint64: "u64", C.c_double: "f64"}.get(a, repr(a))


def test_ctypes_signatures_match_header():
    """Every prototype of the header has a ctypes signature in the Python
    binding with the same number of parameters and the same kind of each
    parameter and of the result (a mismatch would pass wrong values to the
    device entry points)."""
    from cantucci_amd import _lib
    protos = _prototypes()
    assert len(protos) == len(declared_symbols())
    assert set(_lib._SIGS) == {name for _, name, _ in protos}
    for ret, name, params in protos:
        res, args = _lib._SIGS[name]
        assert [_c_kind(p) for p in params] == [_py_kind(a) for a in args], name
        assert _c_kind(ret) == _py_kind(res), name


def test_header_has_no_torch_or_cpp_types():
    text = (ROOT / "include" / "cantucci_hip.h").read_text()
    for bad in ("torch", "at::", "std::", "Tensor"):
        assert bad not in text


def test_struct_layout_matches_oracle():
    import oracle
    from cantucci_amd import _lib
    assert C.sizeof(_lib.CthShape) == oracle.binding._lib.orc_shape_size() == 32
    for f in ("kind", "power", "max_iters", "bailout", "center", "radius"):
        assert getattr(_lib.CthShape, f).offset == getattr(oracle.OrcShape, f).offset


def test_bounding_box_host_only():
    from cantucci_amd import _lib
    span = _lib.CthSpan()
    s = _lib.make_shape(_lib.CTH_SHAPE_MANDELBULB, 8.0, 15, 2.0)
    assert _lib.lib.cth_bounding_box(C.byref(s), C.byref(span)) == 0
    assert list(span.start) == pytest.approx([-1.2] * 3) and list(span.end) == pytest.approx([1.2] * 3)
    sp = _lib.make_shape(_lib.CTH_SHAPE_SPHERE, 0, 1, 0, (1.0, 2.0, 3.0), 0.5)
    assert _lib.lib.cth_bounding_box(C.byref(sp), C.byref(span)) == 0
    assert list(span.start) == [0.5, 1.5, 2.5] and list(span.end) == [1.5, 2.5, 3.5]


def test_invalid_shapes_rejected_without_gpu():
    from cantucci_amd import _lib
    span = _lib.CthSpan()
    bad_iters = _lib.make_shape(_lib.CTH_SHAPE_MANDELBULB, 8.0, 0, 2.0)  # assert!(max_iters >= 1)
    assert _lib.lib.cth_bounding_box(C.byref(bad_iters), C.byref(span)) == _lib.CTH_E_INVALID
    assert b"max_iters" in _lib.lib.cth_last_error()
    bad_power = _lib.make_shape(_lib.CTH_SHAPE_MANDELBULB, 8.5, 15, 2.0)  # Mandelbulb<P: u8>
    assert _lib.lib.cth_bounding_box(C.byref(bad_power), C.byref(span)) == _lib.CTH_E_INVALID
    assert _lib.lib.cth_min_distance_batch(None, C.byref(bad_power), None, 0, None) == _lib.CTH_E_INVALID


def test_context_creation_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from cantucci_amd import _lib
    h = C.c_void_p()
    rc = _lib.lib.cth_ctx_create(0, C.byref(h))
    assert rc == _lib.CTH_E_DEVICE and not h.value


def test_fastdiv_magic_numbers():
    """Host-side magic numbers of the device's divide-by-constant: check the
    formula the kernels use against exact division on sampled numerators."""
    import numpy as np
    for d in list(range(1, 300)) + [729, 4096, 16641, 2146689, 16777216]:
        s = 0
        while (1 << s) < d:
            s += 1
        m = ((((1 << s) - d) << 32) // d + 1) & 0xFFFFFFFF
        if d == 1:
            m = 0
        n = np.unique(np.concatenate([np.arange(0, 5000, dtype=np.uint64),
                                      np.random.default_rng(d).integers(0, 2 ** 31, 5000, dtype=np.uint64),
                                      np.array([2 ** 31 - 1], np.uint64)]))
        q = (((n * np.uint64(m)) >> np.uint64(32)) + n) >> np.uint64(s)
        assert (q == n // np.uint64(d)).all(), d


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before
    touching any device."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    res = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2 and "one GPU" in res.stderr, res.stderr


def test_device_wrapper_argument_checks():
    """The torch wrappers refuse point arrays that are not (n, 3) float32 and
    per-point arrays (powers, incoming gradients) of the wrong size or
    dtype, on the host, before any pointer reaches a kernel."""
    import torch
    from cantucci_amd import device as dev
    with pytest.raises(ValueError):
        dev._points(torch.zeros((4, 2)))
    with pytest.raises(ValueError):
        dev._points(torch.zeros(12))
    with pytest.raises(TypeError):
        dev._points(torch.zeros((4, 3), dtype=torch.float64))
    assert dev._points(torch.zeros((4, 3)).t().contiguous().t()).is_contiguous()
    with pytest.raises(ValueError):
        dev._per_point(torch.zeros(3), 4, "powers")
    with pytest.raises(TypeError):
        dev._per_point(torch.zeros(4, dtype=torch.float64), 4, "grad_d")
    assert dev._per_point(None, 4, "powers") is None
    assert dev._per_point(torch.zeros((2, 2)), 4, "powers").numel() == 4


def test_create_spans_mirror_matches_oracle():
    """The Python mirror's create_spans (used by ShapeMesh) against the
    oracle's, on dyadic and non-dyadic spans."""
    import numpy as np
    import oracle as O
    from cantucci_amd.mesh import create_spans
    for span in ([-1.2] * 3 + [1.2] * 3, [0.1, -0.7, 0.3, 0.35, -0.2, 0.9], [-1e-3, 2.0, -5.0, 1e-3, 2.5, 7.0]):
        got = np.array([[*a, *b] for a, b in create_spans((span[:3], span[3:]))], np.float32)
        assert (got.view(np.uint32) == O.create_spans(span).view(np.uint32)).all()
