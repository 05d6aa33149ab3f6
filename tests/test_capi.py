"""The C-ABI library loads and exports every symbol declared in include/*.h (no GPU calls)."""
import re
from pathlib import Path

from xped_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_in_headers():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b((?:xh|xt)_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(xlib):
    names = declared_in_headers()
    assert len(names) > 60
    missing = [n for n in sorted(names) if not hasattr(xlib, n)]
    assert not missing, missing


def test_python_binding_covers_the_headers(xlib):
    assert set(_lib.declared_symbols()) <= declared_in_headers()


def test_structure_only_runtime_rejects_data_ops(rt_struct):
    import pytest

    import xped_amd.tensor as xt
    from xped_amd._lib import XpedNoDevice

    b = xt.Qbasis(xt.U1).push_back(0, 2).push_back(1, 1)
    t = xt.Tensor(rt_struct, [b], [b])
    assert t.sectors() == [((0,), 0, 2, 2), ((1,), 4, 1, 1)]
    with pytest.raises(XpedNoDevice):
        t.download()
    with pytest.raises(XpedNoDevice):
        t.tSVD(2, 0.0)


def test_runtime_destroy_is_refused_while_tensors_live(xlib):
    import ctypes as C

    import xped_amd.tensor as xt

    rt = C.c_void_p()
    assert xlib.xt_rt_create(-1, C.byref(rt)) == 0
    b = xt.Qbasis(xt.U1).push_back(0, 2).push_back(1, 1)
    arr = (C.c_void_p * 2)(b.h.value, b.h.value)
    t1, t2 = C.c_void_p(), C.c_void_p()
    assert xlib.xt_tensor_create(rt, 1, 1, arr, C.byref(t1)) == 0
    assert xlib.xt_tensor_copy(t1, C.byref(t2)) == 0
    assert xlib.xt_rt_destroy(rt) == _lib.XH_EINVAL
    assert b"still alive" in xlib.xt_last_error(rt)
    # the runtime is intact: structure queries on its tensors still work
    r, c, n, s = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
    assert xlib.xt_tensor_shape(t2, C.byref(r), C.byref(c), C.byref(n), C.byref(s)) == 0 and s.value == 5
    assert xlib.xt_tensor_destroy(t1) == 0
    assert xlib.xt_rt_destroy(rt) == _lib.XH_EINVAL
    assert xlib.xt_tensor_destroy(t2) == 0
    assert xlib.xt_rt_destroy(rt) == 0


def test_cpp_facade_compiles_and_runs_structure_queries(xlib):
    """include/xped_amd.hpp (the C++ facade a maintainer binds Xped's Tensor to) compiled against
    the headers and linked with the library: structure queries on a structure-only runtime match
    the oracle, and a data operation without a device surfaces as std::runtime_error."""
    from oracle import tensor as oten
    from oracle.structure import Qbasis
    from oracle.symmetry import U1, Sym

    from .cpp_adapter import build_and_run

    out = build_and_run("struct")
    s = Sym(U1)
    virt, phys = Qbasis(s), Qbasis(s)
    for q in range(-2, 3):
        virt.push_back((q,), 3 - abs(q))
    phys.push_back((-1,), 1)
    phys.push_back((1,), 1)
    fused = virt.combine(phys)
    assert out["virt_dim"] == virt.dim() and out["fused_dim"] == fused.dim() and out["fused_full"] == fused.full_dim()
    assert out["fused"] == [[e[0][0], e[1], e[2]] for e in fused.data]
    A = oten.Tensor(s, [virt, phys], [virt])
    assert out["A"] == [[q[0], o, r, c] for q, o, r, c in A.structure()]
    assert out["error_mapped"] is True
