"""Scan, radix sort and tile binning on the GPU against the oracle and the
reference's known-answer tests: bit-exact."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from nanotracer_amd.api import PrefixScan, RadixSort, TileBinner, TilingParams

from .helpers import lcg_u32, projected_record

pytestmark = pytest.mark.gpu


def _scan(ctx, a):
    buf = ctx.storage_buffer(a)
    PrefixScan().scan(ctx, buf, a.size)
    return buf.download(np.uint32, a.size)


def test_scan_all_ones_yields_indices(ctx):
    # prefix_scan.rs:291-307
    out = _scan(ctx, np.ones(1000, np.uint32))
    assert np.array_equal(out, np.arange(1000, dtype=np.uint32))


def test_scan_random_matches_cpu_reference(ctx):
    # prefix_scan.rs:309-333 (LCG seed 0x12345678, mod 17, 70,000 values)
    a = lcg_u32(0x12345678, 70_000, 17)
    assert np.array_equal(_scan(ctx, a), O.exclusive_scan(a))


@pytest.mark.parametrize("n", [1, 255, 4096, 4097, 1 << 20, 3_000_001])
def test_scan_sizes_and_wrapping(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)  # wraps
    assert np.array_equal(_scan(ctx, a), O.exclusive_scan(a))


def _sort(ctx, keys, vals, bits=None):
    kb, vb = ctx.storage_buffer(keys), ctx.storage_buffer(vals)
    if keys.dtype == np.uint32:
        RadixSort().sort(ctx, kb, vb, keys.size, 32 if bits is None else bits)
    else:
        RadixSort().sort_u64(ctx, kb, vb, keys.size, 64 if bits is None else bits)
    return kb.download(keys.dtype, keys.size), vb.download(np.uint32, keys.size)


def test_sort_small_array_stable(ctx):
    # radix_sort.rs:357-375
    sk, sv = _sort(ctx, np.array([3, 1, 4, 1, 5, 9, 2, 6], np.uint32), np.arange(8, dtype=np.uint32))
    assert list(sk) == [1, 1, 2, 3, 4, 5, 6, 9]
    assert sv[0] == 1 and sv[1] == 3


def test_sort_random_u32_with_payload(ctx):
    # radix_sort.rs:377-408 (10,000 LCG keys, seed 0xdeadbeef)
    keys = lcg_u32(0xDEADBEEF, 10_000)
    sk, sv = _sort(ctx, keys, np.arange(10_000, dtype=np.uint32))
    assert np.array_equal(sk, keys[sv])
    assert np.all(sk[:-1] <= sk[1:])
    ok, ov = O.sort_pairs(keys, np.arange(10_000, dtype=np.uint32))
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)


# key widths that leave the top 8-bit digit partly empty, or exactly full
@pytest.mark.parametrize("n,bits", [(5, 32), (4096, 32), (4097, 20), (1_000_003, 32), (2_500_000, 28), (4097, 25),
                                    (1_000_003, 27), (300_000, 26), (70_001, 17), (1_000_003, 30), (5000, 29),
                                    (600_000, 24), (4096, 16)])
def test_sort_u32_stable_vs_oracle(ctx, n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # plenty of ties: stability matters
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = _sort(ctx, keys, vals, bits)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)


@pytest.mark.parametrize("n,bits", [(10_000, 40), (1_500_000, 36)])
def test_sort_u64_stable_vs_oracle(ctx, n, bits):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    keys[::7] = keys[3]
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = _sort(ctx, keys, vals, bits)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)


def _tile_bits_strict(num_tiles):
    b = 1
    while (1 << b) <= num_tiles:
        b += 1
    return b


def _tile_keys(rng, n, num_tiles, hot, cull, dtype):
    """(tile << 16 | depth16) keys: a fraction `hot` of them in three tiles (so
    groups and single tiles outgrow a workgroup's LDS), depths from a small
    set for a third of them (ties, which must keep their input order), and a
    fraction `cull` all ones (dropped pairs)."""
    tiles = rng.integers(0, num_tiles, n).astype(np.uint64)
    nh = int(hot * n)
    tiles[:nh] = rng.choice(np.array([0, num_tiles // 2, num_tiles - 1], np.uint64), nh)
    rng.shuffle(tiles)
    depth = rng.integers(0, 1 << 16, n).astype(np.uint64)
    tie = rng.uniform(size=n) < 0.33
    depth[tie] = rng.choice(np.array([0, 7, 0xFFFF], np.uint64), int(tie.sum()))
    keys = (tiles << np.uint64(16)) | depth
    keys[rng.uniform(size=n) < cull] = np.uint64((1 << (8 * np.dtype(dtype).itemsize)) - 1)
    return keys.astype(dtype)


def _tile_sort(ctx, keys, vals, tile_bits, num_tiles):
    kb, vb = ctx.storage_buffer(keys), ctx.storage_buffer(vals)
    rb = ctx.storage_buffer(np.full(2 * max(num_tiles, 1), 0xDEAD, np.uint32))
    RadixSort().sort_tiles(ctx, kb, vb, keys.size, tile_bits, rb, num_tiles, keys.dtype.itemsize)
    return (kb.download(keys.dtype, keys.size), vb.download(np.uint32, keys.size),
            rb.download(np.uint32, 2 * num_tiles))


def _ranges_u64(sorted_keys, num_tiles):
    """[begin, end) of each tile by binary search over the sorted tile ids."""
    tiles = (sorted_keys >> np.uint64(16)).astype(np.int64)
    ids = np.arange(num_tiles)
    lo = np.searchsorted(tiles, ids, "left")
    hi = np.searchsorted(tiles, ids, "right")
    empty = lo == hi
    lo[empty] = 0
    hi[empty] = 0
    return np.stack([lo, hi], 1).reshape(-1).astype(np.uint32)


# tile counts: 200 (8 tile bits, one tile per group), 2040 (1920x1080 at
# 32-px tiles: groups of 8 tiles), 8160 (1920x1080 at 16 px: groups of 32),
# 32400 (3840x2160 at 16 px: every tile bit device-wide, then per tile)
@pytest.mark.parametrize("num_tiles,n,hot,cull", [
    (200, 60_000, 0.0, 0.0), (200, 300_000, 0.5, 0.05),
    (2040, 400_000, 0.0, 0.05), (2040, 400_000, 0.6, 0.0),
    (8160, 300_000, 0.3, 0.05), (32400, 300_000, 0.3, 0.05), (32400, 50_000, 0.0, 0.0),
])
def test_tile_sort_equals_full_sort_u32(ctx, num_tiles, n, hot, cull):
    rng = np.random.default_rng(num_tiles * 7 + n + int(100 * hot))
    tb = _tile_bits_strict(num_tiles)
    keys = _tile_keys(rng, n, num_tiles, hot, cull, np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv, sr = _tile_sort(ctx, keys, vals, tb, num_tiles)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok)
    assert np.array_equal(sv, ov)
    assert np.array_equal(sr, O.tile_ranges(ok, num_tiles))


def test_tile_sort_all_tile_bits_used_without_dropped_pairs(ctx):
    # 2048 tiles in 11 bits: tile 2047 exists, allowed with no all-ones key
    rng = np.random.default_rng(11)
    keys = _tile_keys(rng, 200_000, 2048, 0.2, 0.0, np.uint32)
    keys[:100] = np.uint32((2047 << 16) | 0xFFFF)
    vals = rng.integers(0, 2**32, keys.size, dtype=np.uint64).astype(np.uint32)
    sk, sv, sr = _tile_sort(ctx, keys, vals, 11, 2048)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)
    assert np.array_equal(sr, O.tile_ranges(ok, 2048))


def test_tile_sort_u64_keys(ctx):
    # 3840x2160 at 8-px tiles: 129,600 tiles, 17 tile bits, 64-bit keys
    rng = np.random.default_rng(17)
    num_tiles, n = 129_600, 400_000
    keys = _tile_keys(rng, n, num_tiles, 0.3, 0.05, np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv, sr = _tile_sort(ctx, keys, vals, 17, num_tiles)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)
    live = ok[ok != np.uint64(2**64 - 1)]
    assert np.array_equal(sr, _ranges_u64(live, num_tiles))


# Sorts of 16M pairs or more take their onesweep offsets from per-partition
# digit-count tables (passes 0 and 1) instead of the look-back; the smaller
# cases above keep the look-back in every pass.
TABLE_MIN_PAIRS = 16 << 20


def test_tile_sort_count_tables_u32(ctx):
    # 31,100 tiles (15 tile bits: a 7-bit first pass, an 8-bit second)
    rng = np.random.default_rng(31100)
    num_tiles, n = 31_100, TABLE_MIN_PAIRS + 700_001
    keys = _tile_keys(rng, n, num_tiles, 0.2, 0.05, np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv, sr = _tile_sort(ctx, keys, vals, 15, num_tiles)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)
    assert np.array_equal(sr, O.tile_ranges(ok, num_tiles))


def test_tile_sort_count_tables_u64(ctx):
    # 17 tile bits in 64-bit keys: a 1-bit first pass and two 8-bit passes,
    # the last with the look-back
    rng = np.random.default_rng(129_600)
    num_tiles, n = 129_600, TABLE_MIN_PAIRS + 3
    keys = _tile_keys(rng, n, num_tiles, 0.1, 0.05, np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv, sr = _tile_sort(ctx, keys, vals, 17, num_tiles)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)
    live = ok[ok != np.uint64(2**64 - 1)]
    assert np.array_equal(sr, _ranges_u64(live, num_tiles))


def test_sort_u32_count_tables_vs_oracle(ctx):
    n = TABLE_MIN_PAIRS + 4097
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    keys[: n // 4] = keys[n // 4: 2 * (n // 4)]  # ties
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = _sort(ctx, keys, vals, 32)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)


@pytest.mark.parametrize("n", [0, 1, 2, 4097])
def test_tile_sort_small_and_edge_sizes(ctx, n):
    rng = np.random.default_rng(n)
    keys = _tile_keys(rng, max(n, 1), 100, 0.0, 0.1, np.uint32)[:n]
    vals = np.arange(n, dtype=np.uint32)
    sk, sv, sr = _tile_sort(ctx, keys, vals, 7, 100)
    ok, ov = O.sort_pairs(keys, vals)
    assert np.array_equal(sk, ok) and np.array_equal(sv, ov)
    assert np.array_equal(sr, O.tile_ranges(ok, 100))


def test_tile_sort_rejects_bad_widths(ctx):
    keys = np.zeros(16, np.uint32)
    vals = np.zeros(16, np.uint32)
    kb, vb = ctx.storage_buffer(keys), ctx.storage_buffer(vals)
    rb = ctx.storage_buffer(np.zeros(8, np.uint32))
    for bits in (0, 17):
        with pytest.raises(Exception):
            RadixSort().sort_tiles(ctx, kb, vb, 16, bits, rb, 4, 4)
    with pytest.raises(Exception):  # 5 tiles do not fit 2 tile bits
        RadixSort().sort_tiles(ctx, kb, vb, 16, 2, rb, 5, 4)


def _bin_gpu(ctx, proj, params):
    buf = ctx.storage_buffer(np.ascontiguousarray(proj, np.float32))
    res = TileBinner().bin(ctx, buf, proj.shape[0], params)
    return res


def test_four_splats_bin_to_expected_tiles(ctx):
    # tile_binner.rs:443-501
    proj = np.stack([projected_record(8, 8, 10, 4), projected_record(40, 8, 20, 4), projected_record(8, 56, 30, 4),
                     projected_record(56, 56, 5, 4)])
    res = _bin_gpu(ctx, proj, TilingParams(64, 64, 16, 100.0))
    assert res.total_pairs == 4 and res.num_tiles == 16
    keys, pay, ranges = res.sorted_keys(), res.sorted_payloads(), res.tile_ranges()
    assert list(keys >> 16) == [0, 2, 12, 15]
    nz = [(t, ranges[2 * t], ranges[2 * t + 1]) for t in range(16) if ranges[2 * t + 1] > ranges[2 * t]]
    assert nz == [(0, 0, 1), (2, 1, 2), (12, 2, 3), (15, 3, 4)]
    assert list(pay) == [0, 1, 2, 3]


def test_two_splats_same_tile_sorted_by_depth(ctx):
    # tile_binner.rs:503-541
    proj = np.stack([projected_record(24, 24, 80, 4), projected_record(24, 24, 5, 4)])
    res = _bin_gpu(ctx, proj, TilingParams(64, 64, 16, 100.0))
    assert res.total_pairs == 2
    assert list(res.sorted_keys() >> 16) == [5, 5]
    assert list(res.sorted_payloads()) == [1, 0]
    r = res.tile_ranges()
    assert r[10] == 0 and r[11] == 2


def test_empty_scene_fast_path(ctx):
    # tile_binner.rs:247-259: nothing visible -> zero pairs, ranges (0, 0)
    proj = np.stack([projected_record(24, 24, 80, 4, visible=0.0)])
    res = _bin_gpu(ctx, proj, TilingParams(64, 64, 16, 100.0))
    assert res.total_pairs == 0
    assert not res.tile_ranges().any()


@pytest.mark.parametrize("w,h,ts", [(256, 256, 16), (1280, 720, 16), (333, 177, 8), (1920, 1080, 32)])
def test_bin_random_projected_matches_oracle(ctx, w, h, ts):
    rng = np.random.default_rng(w * h + ts)
    n = 20_000
    proj = np.zeros((n, 12), np.float32)
    proj[:, 0] = rng.uniform(-40, w + 40, n)
    proj[:, 1] = rng.uniform(-40, h + 40, n)
    proj[:, 2] = rng.uniform(0.0, 14.0, n)
    proj[:, 3] = np.ceil(rng.exponential(6.0, n))
    proj[:, 7] = (rng.uniform(size=n) > 0.1).astype(np.float32)
    params = TilingParams(w, h, ts, 12.0)
    res = _bin_gpu(ctx, proj, params)
    ok, op, orr = O.bin(proj, O.tiling(w, h, ts, 12.0))
    assert res.total_pairs == ok.size
    assert np.array_equal(res.sorted_keys(), ok)
    assert np.array_equal(res.sorted_payloads(), op)
    assert np.array_equal(res.tile_ranges(), orr)


def test_bin_more_than_65536_tiles_uses_wide_keys(ctx):
    # 3840x2160 at 8-px tiles = 129,600 tiles: the reference's 16-bit tile id
    # wraps (tile_emit.wgsl:60-61); here the key is 64-bit and stays ordered.
    w, h, ts = 3840, 2160, 8
    rng = np.random.default_rng(7)
    n = 30_000
    proj = np.zeros((n, 12), np.float32)
    proj[:, 0] = rng.uniform(0, w, n)
    proj[:, 1] = rng.uniform(0, h, n)
    proj[:, 2] = rng.uniform(0.0, 10.0, n)
    proj[:, 3] = np.ceil(rng.exponential(4.0, n))
    proj[:, 7] = 1.0
    res = _bin_gpu(ctx, proj, TilingParams(w, h, ts, 12.0))
    k64 = res.sorted_keys_u64()
    assert np.all(k64[:-1] <= k64[1:])
    tiles = (k64 >> np.uint64(16)).astype(np.int64)
    assert tiles.max() >= 65536
    pay = res.sorted_payloads()
    # each pair's tile lies in its splat's bbox, and every bbox tile appears once
    counts, total = O.tile_count(proj, O.tiling(w, h, ts, 12.0))
    assert total == res.total_pairs
    assert np.array_equal(np.bincount(pay, minlength=n), counts)
