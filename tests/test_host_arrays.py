"""The per-step host output reserve (radtranbook_amd/_lib.py HostArrays):
arrays handed out are fresh, distinct, of the asked shape, never shared with
a later take, and the reserve refills in the background."""
import time

import numpy as np

from radtranbook_amd._lib import HostArrays


def _wait_for(cond, timeout=5.0):
    t0 = time.time()
    while not cond() and time.time() - t0 < timeout:
        time.sleep(0.005)
    return cond()


def test_take_gives_distinct_arrays_and_refills():
    h = HostArrays((200, 200))
    assert h.depth == 12
    first = h.take()
    assert first.shape == (200, 200) and first.dtype == np.float64
    assert _wait_for(lambda: len(h._q) == h.depth)
    got = [h.take() for _ in range(2 * h.depth)]  # more than the reserve: falls back to np.empty
    got.append(first)
    for k, a in enumerate(got):
        a[...] = k
    # no two handed-out arrays share memory, so a kept step output is never overwritten
    assert all(float(a[0, 0]) == k and float(a[-1, -1]) == k for k, a in enumerate(got))
    ptrs = {a.ctypes.data for a in got}
    assert len(ptrs) == len(got)
    assert _wait_for(lambda: len(h._q) == h.depth)
    h.close()
    assert len(h._q) == 0
    assert h.take().shape == (200, 200)  # still usable after close, without a reserve


def test_small_and_huge_shapes_have_no_reserve():
    assert HostArrays((100,)).depth == 0
    assert HostArrays((4096, 4096)).depth == 0  # 128 MiB per array exceeds the 64 MiB reserve cap
    assert HostArrays((1000, 1000), max_bytes=32 << 20).depth == 4
