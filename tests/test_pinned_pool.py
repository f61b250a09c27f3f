"""
The pool of pinned output arrays (kernelhip.PinnedPool), on the host: the
library's pinned allocator is replaced by ordinary ctypes buffers, so the
pool's bookkeeping runs without a device.  A buffer returns to the pool
when its last view is gone, the first buffer of a size comes with a spare,
and past the cap the pool hands out ordinary arrays.
"""

import ctypes as ct
import gc

import numpy as np

from sasmodels_amd import kernelhip


class _FakeLib:
    """saship_host_alloc over ctypes buffers kept alive here."""

    def __init__(self):
        self.bufs = []

    def saship_host_alloc(self, nbytes, ptr_ref):
        buf = (ct.c_char*int(nbytes))()
        self.bufs.append(buf)
        ct.cast(ptr_ref, ct.POINTER(ct.c_void_p))[0] = ct.addressof(buf)
        return 0


def _pool(monkeypatch, cap):
    fake = _FakeLib()
    monkeypatch.setattr(kernelhip._lib, "lib", lambda: fake)
    return kernelhip.PinnedPool(cap_bytes=cap), fake


def test_first_buffer_comes_with_a_spare_and_views_keep_it(monkeypatch):
    pool, fake = _pool(monkeypatch, 1 << 20)
    a = pool.empty(100)
    assert len(fake.bufs) == 2 and pool._held == 2*800
    b = pool.empty(100)                 # the spare: no new allocation
    assert len(fake.bufs) == 2
    assert a.__array_interface__["data"][0] != b.__array_interface__["data"][0]
    rows = a.reshape(4, 25)
    addr = a.__array_interface__["data"][0]
    del a
    gc.collect()
    c = pool.empty(100)                 # a's buffer is still viewed by rows: a new one
    assert c.__array_interface__["data"][0] != addr and len(fake.bufs) == 3
    del rows
    gc.collect()
    d = pool.empty(100)                 # now a's buffer is back
    assert d.__array_interface__["data"][0] == addr and len(fake.bufs) == 3
    d[:] = np.arange(100.0)
    assert d.dtype == np.float64 and d.size == 100 and d[99] == 99.0


def test_past_the_cap_arrays_are_ordinary(monkeypatch):
    pool, fake = _pool(monkeypatch, 1000)
    a = pool.empty(100)                 # 800 bytes: no room for the spare
    assert len(fake.bufs) == 1 and pool._held == 800
    b = pool.empty(100)                 # over the cap: plain host memory
    assert len(fake.bufs) == 1 and b.size == 100
    assert b.__array_interface__["data"][0] != a.__array_interface__["data"][0]
