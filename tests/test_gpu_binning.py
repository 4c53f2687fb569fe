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
