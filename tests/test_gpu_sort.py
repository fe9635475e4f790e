"""The builder's device sort (ag_sort_pairs_host: the radix passes of
ag_csr_build, exposed for sortRunAlgorithms) against numpy's stable argsort.

Every key width from 1 to 32 bits, so one to four passes and every split of
the key into digits of at most 8 bits, on a length that leaves a partial last
tile, with many equal keys (stability: values are the input positions), and
the empty and one-element cases."""
import ctypes as C

import numpy as np
import pytest

from accelgraph_express_amd import _lib as L

pytestmark = pytest.mark.gpu


def _sort(ctx, keys, vals, limit):
    n = keys.size
    ko = np.zeros(max(n, 1), np.uint32)
    vo = np.zeros(max(n, 1), np.uint32)
    u32 = C.POINTER(C.c_uint32)
    rc = L.load().ag_sort_pairs_host(ctx.h, keys.ctypes.data_as(u32), vals.ctypes.data_as(u32), n,
                                     limit, ko.ctypes.data_as(u32), vo.ctypes.data_as(u32))
    assert rc == L.AG_OK, L.load().ag_last_error()
    return ko[:n], vo[:n]


@pytest.mark.parametrize("bits", list(range(1, 33)))
def test_stable_sort_every_key_width(ctx, bits):
    n = 300_007
    rng = np.random.RandomState(bits)
    limit = (1 << bits) if bits < 32 else 0xFFFFFFFF
    # a pool of n / 8 distinct keys spread over the whole range: every key
    # repeats about 8 times, and every digit of a wide key varies
    pool = rng.randint(0, limit, size=max(1, n // 8), dtype=np.uint64).astype(np.uint32)
    pool[0] = limit - 1
    keys = pool[rng.randint(0, pool.size, n)]
    vals = np.arange(n, dtype=np.uint32)
    print(f"case: {bits}-bit keys, n={n}", flush=True)
    ko, vo = _sort(ctx, keys, vals, limit)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(vo, order.astype(np.uint32))
    np.testing.assert_array_equal(ko, keys[order])


def test_sort_edge_cases(ctx):
    for keys in (np.zeros(0, np.uint32), np.array([7], np.uint32), np.full(5000, 3, np.uint32)):
        vals = np.arange(keys.size, dtype=np.uint32)
        ko, vo = _sort(ctx, keys, vals, 8)
        np.testing.assert_array_equal(ko, keys)
        np.testing.assert_array_equal(vo, vals)
