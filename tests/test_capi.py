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
