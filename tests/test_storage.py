"""The storage wire format (xt_tensor_save/load/serialize/deserialize): the fields of the
reference's StorageType serialisation (StorageType_contiguous.hpp:191-204) in one binary layout.
CPU: structures of U(1), SU(2) and fermionic U(1)xU(1) tensors, including conjugated legs after
permutes, survive a round trip on a structure-only runtime; corrupted archives are refused."""
import itertools
import random

import pytest

import xped_amd.tensor as xt
from xped_amd._lib import XpedError

from .helpers import rand_entries


def _t(rt, pcode, seed, rank=2, corank=2):
    r = random.Random(seed)
    legs = []
    for _ in range(rank + corank):
        b = xt.Qbasis(pcode)
        for q, d in rand_entries(pcode, r, 3, 4):
            b.push_back(q, d)
        legs.append(b)
    return xt.Tensor(rt, legs[:rank], legs[rank:])


@pytest.mark.parametrize("pcode", [xt.U1, xt.SU2, xt.FU1xU1], ids=["U1", "SU2", "FU1xU1"])
def test_structure_round_trip(rt_struct, pcode, tmp_path):
    t = _t(rt_struct, pcode, 5)
    for shift, p in [(0, (0, 1, 2, 3)), (1, (2, 0, 3, 1)), (-2, (3, 2, 1, 0))]:
        u = t.permute(shift, p) if (shift, p) != (0, (0, 1, 2, 3)) else t
        a = u.serialize()
        v = xt.Tensor.deserialize(rt_struct, a)
        assert v.sectors() == u.sectors() and (v.rank, v.corank) == (u.rank, u.corank)
        for i in range(u.rank + u.corank):
            assert v.leg(i).entries() == u.leg(i).entries() and v.leg(i).info() == u.leg(i).info()
        assert v.serialize() == a
        f = tmp_path / f"t{shift}.xt"
        u.save(f)
        assert xt.Tensor.load(rt_struct, f).sectors() == u.sectors()


def test_corrupted_archives_are_refused(rt_struct):
    a = bytearray(_t(rt_struct, xt.U1, 6).serialize())
    with pytest.raises(XpedError):
        xt.Tensor.deserialize(rt_struct, bytes(a[:-9]))  # truncated
    with pytest.raises(XpedError):
        xt.Tensor.deserialize(rt_struct, b"NOTATENS" + bytes(a[8:]))
    b = bytearray(a)
    b[-9] ^= 0x40  # a stored block shape that the legs do not give
    with pytest.raises(XpedError):
        xt.Tensor.deserialize(rt_struct, bytes(b))
