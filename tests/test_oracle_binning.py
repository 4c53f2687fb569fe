"""Pins the checker's scan, sort, count/emit and ranges: the reference's
known-answer tests (prefix_scan.rs, radix_sort.rs, tile_binner.rs) and
closed-form properties."""
import math

import numpy as np
import pytest

from oracle import oracle as O

from .helpers import lcg_u32, projected_record


def test_scan_all_ones_yields_indices():
    assert np.array_equal(O.exclusive_scan(np.ones(1000, np.uint32)), np.arange(1000, dtype=np.uint32))


def test_scan_random_matches_cumsum():
    a = lcg_u32(0x12345678, 70_000, 17)
    ref = np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(O.exclusive_scan(a), ref)


def test_scan_wraps_like_u32():
    a = np.full(5, 0xFFFFFFFF, np.uint32)
    assert list(O.exclusive_scan(a)) == [0, 0xFFFFFFFF, 0xFFFFFFFE, 0xFFFFFFFD, 0xFFFFFFFC]


def test_sort_small_array_stable():
    k, v = O.sort_pairs(np.array([3, 1, 4, 1, 5, 9, 2, 6], np.uint32), np.arange(8, dtype=np.uint32))
    assert list(k) == [1, 1, 2, 3, 4, 5, 6, 9] and v[0] == 1 and v[1] == 3


def test_sort_random_is_a_stable_permutation():
    keys = lcg_u32(0xDEADBEEF, 10_000) % np.uint32(997)
    k, v = O.sort_pairs(keys, np.arange(keys.size, dtype=np.uint32))
    assert np.array_equal(k, keys[v])
    assert np.all(k[:-1] <= k[1:])
    same = k[:-1] == k[1:]
    assert np.all(v[:-1][same] < v[1:][same])  # ties keep input order
    assert np.array_equal(np.sort(v), np.arange(keys.size))


def test_four_splats_bin_to_expected_tiles():
    proj = np.stack([projected_record(8, 8, 10, 4), projected_record(40, 8, 20, 4), projected_record(8, 56, 30, 4),
                     projected_record(56, 56, 5, 4)])
    keys, pay, ranges = O.bin(proj, O.tiling(64, 64, 16, 100.0))
    assert list(keys >> 16) == [0, 2, 12, 15] and list(pay) == [0, 1, 2, 3]
    nz = [(t, ranges[2 * t], ranges[2 * t + 1]) for t in range(16) if ranges[2 * t + 1] > ranges[2 * t]]
    assert nz == [(0, 0, 1), (2, 1, 2), (12, 2, 3), (15, 3, 4)]


def test_two_splats_same_tile_sorted_by_depth():
    proj = np.stack([projected_record(24, 24, 80, 4), projected_record(24, 24, 5, 4)])
    keys, pay, ranges = O.bin(proj, O.tiling(64, 64, 16, 100.0))
    assert list(keys >> 16) == [5, 5] and list(pay) == [1, 0] and ranges[10] == 0 and ranges[11] == 2


def test_depth_quantisation_closed_form():
    # d16 = trunc(clamp(depth / depth_max, 0, 1) * 65535), measured from 0
    depths = np.array([0.0, 0.01, 1.0, 49.999, 50.0, 75.0, -3.0], np.float32)
    proj = np.stack([projected_record(8, 8, d, 2) for d in depths])
    keys, pay, _ = O.bin(proj, O.tiling(16, 16, 16, 50.0))
    d16 = {int(p): int(k & 0xFFFF) for k, p in zip(keys, pay)}
    for i, d in enumerate(depths):
        want = int(np.float32(min(max(np.float32(d) / np.float32(50.0), 0.0), 1.0)) * np.float32(65535.0))
        assert d16[i] == want


@pytest.mark.parametrize("w,h,ts", [(64, 48, 16), (100, 37, 8), (96, 96, 32)])
def test_bin_properties_brute_force(w, h, ts):
    """Each visible splat appears once in every tile its (mean +- r) box
    touches and nowhere else; the stream is ordered by (tile, depth16,
    splat index); ranges delimit exactly each tile's entries."""
    rng = np.random.default_rng(w + h + ts)
    n = 300
    proj = np.zeros((n, 12), np.float32)
    proj[:, 0] = rng.uniform(-20, w + 20, n)
    proj[:, 1] = rng.uniform(-20, h + 20, n)
    proj[:, 2] = rng.uniform(0, 12, n)
    proj[:, 3] = np.ceil(rng.uniform(0, 30, n))
    proj[:, 7] = (rng.uniform(size=n) > 0.2).astype(np.float32)
    t = O.tiling(w, h, ts, 10.0)
    keys, pay, ranges = O.bin(proj, t)
    tx_n, ty_n = -(-w // ts), -(-h // ts)
    expect = set()
    for i in range(n):
        if proj[i, 7] == 0:
            continue
        x, y, r = float(proj[i, 0]), float(proj[i, 1]), float(proj[i, 3])
        for ty in range(ty_n):
            for tx in range(tx_n):
                # tile [tx*ts, (tx+1)*ts) intersects [x - r, x + r] in tile units
                if (math.floor((x - r) / ts) <= tx <= math.floor((x + r) / ts)
                        and math.floor((y - r) / ts) <= ty <= math.floor((y + r) / ts)):
                    expect.add((ty * tx_n + tx, i))
    got = [(int(k >> 16), int(p)) for k, p in zip(keys, pay)]
    assert len(got) == len(set(got)) and set(got) == expect
    triples = [(int(k >> 16), int(k & 0xFFFF), int(p)) for k, p in zip(keys, pay)]
    assert triples == sorted(triples)
    for tile in range(tx_n * ty_n):
        b, e = ranges[2 * tile], ranges[2 * tile + 1]
        idx = [j for j, (tt, _) in enumerate(got) if tt == tile]
        if idx:
            assert (b, e) == (idx[0], idx[-1] + 1)
        else:
            assert (b, e) == (0, 0)


def test_empty_scene():
    proj = np.stack([projected_record(24, 24, 80, 4, visible=0.0)])
    keys, pay, ranges = O.bin(proj, O.tiling(64, 64, 16, 100.0))
    assert keys.size == 0 and not ranges.any()
