# SPDX-License-Identifier: 0BSD
"""PooledDevice (pacuda_amd/runtime.py) on the CPU, over a stand-in for the library's allocation
entry points: a freed buffer goes back to the pool and serves the next request of at least half
its size, a buffer in use is never handed out twice, and close() frees what the pool holds."""
import numpy as np

from pacuda_amd.runtime import PooledDevice


class FakeLib:
    """pacuda_device_alloc / _free / _synchronize as counters over fake addresses"""

    def __init__(self):
        self.next, self.live, self.allocs = 1 << 20, {}, 0

    def pacuda_device_alloc(self, nbytes):
        p = self.next
        self.next += (nbytes + 4095) // 4096 * 4096 + 4096
        self.live[p] = nbytes
        self.allocs += 1
        return p

    def pacuda_device_free(self, p):
        del self.live[p]

    def pacuda_synchronize(self):
        return 0


def test_freed_buffers_serve_later_requests():
    lib = FakeLib()
    dev = PooledDevice(lib)
    a = dev.empty(1 << 20)  # 8 MiB
    pa = a.ptr
    a.free()
    assert a.ptr is None
    b = dev.empty(900_000)  # (7.2 MB from the 8 MiB buffer)
    assert b.ptr == pa and b.nbytes == 7_200_000 and lib.allocs == 1
    c = dev.empty(900_000)  # the first is in use: a new one
    assert c.ptr != pa and lib.allocs == 2
    b.free()
    b.free()  # (a second free of the same handle is a no-op)
    d = dev.empty(100_000)  # 0.8 MB: far smaller than the free buffers, a new one
    assert d.ptr not in (pa, c.ptr) and lib.allocs == 3
    e = dev.empty(np.uint64(1 << 20), np.float64)
    assert e.ptr == pa


def test_arrays_go_through_the_pool_and_close_frees_it():
    lib = FakeLib()
    dev = PooledDevice(lib)
    x = dev.empty(64, np.uint32)
    y = dev.empty(64, np.uint32)
    x.free()
    y.free()
    assert len(lib.live) == 2
    z = dev.empty(60, np.uint32)
    assert z.ptr in lib.live
    late = dev.empty(64, np.uint32)  # (in use when the pool closes)
    z.free()
    dev.close()
    assert list(lib.live) == [late.ptr]
    late.free()
    assert lib.live == {}
