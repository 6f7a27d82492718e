"""DLPack export of device buffers (cnhip.dlpack), checked on host memory: the capsule layout,
zero-copy aliasing through torch.from_dlpack, the owner lease and the reference's view checks
(python/bindings.cpp:282-340)."""
import ctypes as C
import gc

import numpy as np
import pytest

from cnhip import dlpack

torch = pytest.importorskip("torch")


class Owner:
    def __init__(self):
        self.n = 0
        self.released = 0

    def _acquire_view(self):
        self.n += 1

    def _release_view(self):
        self.n -= 1
        self.released += 1


def test_capsule_aliases_host_buffer_and_releases_owner():
    buf = (C.c_float * 24)(*range(24))
    owner = Owner()
    cap = dlpack.to_dlpack(C.addressof(buf), (6, 4), "float32", dlpack.KDL_CPU, 0, owner=owner, size_bytes=96)
    assert owner.n == 1
    t = torch.from_dlpack(cap)
    assert t.shape == (6, 4) and t.dtype == torch.float32
    np.testing.assert_array_equal(t.numpy(), np.arange(24, dtype=np.float32).reshape(6, 4))
    buf[5] = 100.0  # the tensor aliases the buffer
    assert t[1, 1].item() == 100.0
    del t, cap
    gc.collect()
    assert owner.n == 0 and owner.released == 1


def test_unconsumed_capsule_releases_owner():
    buf = (C.c_uint32 * 8)()
    owner = Owner()
    cap = dlpack.to_dlpack(C.addressof(buf), (8,), "uint32", dlpack.KDL_CPU, 0, owner=owner)
    assert owner.n == 1
    del cap
    gc.collect()
    assert owner.n == 0


def test_strided_view():
    buf = (C.c_float * 24)(*range(24))
    cap = dlpack.to_dlpack(C.addressof(buf), (6,), "float32", dlpack.KDL_CPU, 0, strides=(4,), size_bytes=96)
    t = torch.from_dlpack(cap)
    np.testing.assert_array_equal(t.numpy(), np.arange(0, 24, 4, dtype=np.float32))


@pytest.mark.parametrize("kw", [dict(shape=(0, 4)), dict(shape=(6, 4), strides=(4, 0)), dict(shape=(7, 4)),
                                dict(shape=(6, 4), dtype="float64")])
def test_view_checks(kw):
    buf = (C.c_float * 24)()
    args = dict(shape=(6, 4), dtype="float32")
    args.update(kw)
    with pytest.raises(ValueError):
        dlpack.to_dlpack(C.addressof(buf), args["shape"], args["dtype"], dlpack.KDL_CPU, 0,
                         strides=args.get("strides"), size_bytes=96)
