"""TlmDevice.set_tuning and TlmBatch.set_tuning set only the fields named and
keep the others (host logic, no device): a stand-in for the library's
tlm_(batch_)get_tuning / tlm_(batch_)set_tuning holds the tuning struct."""
import ctypes as C

import pytest

from emsolver_amd import _native, tlm


class _FakeLib:
    """Holds one tuning struct; get copies it out, set copies it in."""

    def __init__(self, struct, prefix, init):
        self.cur = struct(*init)
        self.sets = 0
        setattr(self, prefix + "get_tuning", self._get)
        setattr(self, prefix + "set_tuning", self._set)

    def _get(self, h, ref):
        C.memmove(C.addressof(ref._obj), C.addressof(self.cur), C.sizeof(self.cur))
        return 0

    def _set(self, h, ref):
        C.memmove(C.addressof(self.cur), C.addressof(ref._obj), C.sizeof(self.cur))
        self.sets += 1
        return 0


def _fields(s):
    return {f: getattr(s, f) for f, _ in s._fields_}


def test_device_set_tuning_keeps_unnamed_fields():
    lib = _FakeLib(_native.Tuning, "tlm_", (0, 3, 0, 0, 0, 2, -1))  # a 16-bit context's planes
    dev = object.__new__(tlm.TlmDevice)
    dev.lib, dev.h = lib, None
    dev.set_tuning(occupancy=5)
    assert _fields(lib.cur) == {"kernel": 0, "planes": 3, "block_waves": 0, "nontemporal": 0,
                                "grid_blocks": 0, "fold": 2, "occupancy": 5}
    dev.set_tuning(planes=1, block_waves=2)
    assert (lib.cur.planes, lib.cur.block_waves, lib.cur.occupancy) == (1, 2, 5)
    assert dev.get_tuning() == _fields(lib.cur)
    with pytest.raises(TypeError, match="unknown fields"):
        dev.set_tuning(slabs=2)
    assert lib.sets == 2


def test_batch_set_tuning_keeps_unnamed_fields():
    lib = _FakeLib(_native.BatchTuning, "tlm_batch_", (1, 4, 0, 32, 4))
    b = object.__new__(tlm.TlmBatch)
    b.lib, b.h, b.members = lib, None, {}
    b.set_tuning(graph_ticks=0)
    assert _fields(lib.cur) == {"planes": 1, "block_waves": 4, "grid_blocks": 0,
                                "graph_ticks": 0, "planes16": 4}
    b.set_tuning(planes16=2, grid_blocks=7)
    assert _fields(lib.cur) == {"planes": 1, "block_waves": 4, "grid_blocks": 7,
                                "graph_ticks": 0, "planes16": 2}
    with pytest.raises(TypeError, match="unknown fields"):
        b.set_tuning(occupancy=5)
