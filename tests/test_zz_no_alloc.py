"""nts_sort_pairs_u32_noalloc (the sort for callers that cannot allocate)
against nts_sort_pairs_u32 through the C ABI: two calls in a row, the
second smaller and with an odd pass count (the result then comes back from
the static workspace), bit-exact against the allocating entry point and the
closed form (ascending key, ties by input position); sizes beyond the
stated maximum are refused."""
import numpy as np
import pytest

from nanotracer_amd import _lib as L
from nanotracer_amd.api import RadixSort

pytestmark = pytest.mark.gpu


def _sorted_both(ctx, keys, vals, key_bits):
    n = keys.size
    ka, va = ctx.storage_buffer(keys), ctx.storage_buffer(vals)
    kb, vb = ctx.storage_buffer(keys), ctx.storage_buffer(vals)
    RadixSort().sort_noalloc(ctx, ka, va, n, key_bits)
    RadixSort().sort(ctx, kb, vb, n, key_bits)
    out = [b.download(np.uint32, n) for b in (ka, va, kb, vb)]
    for b in (ka, va, kb, vb):
        b.free()
    return out


@pytest.mark.parametrize("order", [(1_500_000, 32, 300_000, 24), (2_000_000, 20, 5, 32)])
def test_noalloc_sort_equals_the_allocating_sort(ctx, order):
    n1, bits1, n2, bits2 = order
    rng = np.random.default_rng(n1 + n2)
    for n, bits in ((n1, bits1), (n2, bits2)):
        keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
        keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # many equal keys: stability matters
        vals = np.arange(n, dtype=np.uint32)
        k_na, v_na, k_al, v_al = _sorted_both(ctx, keys, vals, bits)
        assert np.array_equal(k_na, k_al) and np.array_equal(v_na, v_al), (n, bits)
        idx = np.argsort(keys, kind="stable")
        assert np.array_equal(k_na, keys[idx]) and np.array_equal(v_na, vals[idx]), (n, bits)


def test_noalloc_sort_limits(ctx):
    n = L.SORT_NOALLOC_MAX_PAIRS + 1
    kb, vb = ctx.storage_buffer_zeroed(4 * n), ctx.storage_buffer_zeroed(4 * n)
    with pytest.raises(L.NtsError) as e:
        RadixSort().sort_noalloc(ctx, kb, vb, n)
    assert e.value.code == L.NTS_E_CAPACITY
    RadixSort().sort_noalloc(ctx, kb, vb, 1)  # trivial sizes are no-ops
    RadixSort().sort_noalloc(ctx, kb, vb, 0)
    kb.free()
    vb.free()
