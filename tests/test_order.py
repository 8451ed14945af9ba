"""The spatial ion order of mixed-precision engines (sbk_spatial_order, a
host-only export of libsimbuca_hip.so, so these run without a GPU).

The order only affects speed: the Coulomb kernel forms separations from FP32
offsets for source sub-tiles far from a target block, and a block is compact
when its ions are consecutive in this order.  What is pinned here: it is a
permutation, it depends only on the positions (not on the input order), it is
idempotent (re-sorting a sorted cloud moves nothing, so sbk_reorder of an
unmoved cloud leaves results bit-identical), and aligned blocks are compact.
"""
import numpy as np
import pytest

import simbuca_amd as S
from simbuca_amd import synth


def _block_half_extents(p, blk):
    out = []
    for b in range(0, len(p) - blk + 1, blk):
        s = p[b:b + blk]
        out.append(np.abs(s - 0.5 * (s.min(0) + s.max(0))).max())
    return np.array(out)


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000, 4097, 10000])
def test_order_is_a_canonical_permutation(n):
    rng = np.random.default_rng(n)
    p = rng.normal(0, 1e-3, (n, 3))
    o = S.spatial_order(p)
    assert np.array_equal(np.sort(o), np.arange(n))
    assert np.array_equal(S.spatial_order(p[o]), np.arange(n))          # idempotent
    sh = rng.permutation(n)
    assert np.array_equal(p[sh][S.spatial_order(p[sh])], p[o])        # input order irrelevant


def test_aligned_blocks_are_compact():
    p = synth.workload(3).clouds[0].pos                                # 65,536 ions
    sigma = np.sqrt((p ** 2).mean())
    q = p[S.spatial_order(p)]
    for blk, med_max in ((64, 0.35), (256, 0.75)):
        sorted_ext = _block_half_extents(q, blk) / sigma
        caller_ext = _block_half_extents(p, blk) / sigma
        assert np.median(sorted_ext) < med_max
        assert sorted_ext.max() < 2.0
        assert np.median(caller_ext) > 2.0                              # what it replaces


def test_leaf_boundaries_are_aligned():
    # bisection splits at multiples of the largest power of two A <= 1024 with
    # 2A <= part size: every aligned run of 64 ions lies inside one leaf, so
    # its bounding box is no larger than the leaf's (checked via the sorted
    # x-coordinate inside each run, which is how a leaf is ordered)
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1e-3, (5000, 3))
    q = p[S.spatial_order(p)]
    for b in range(0, 5000 - 64 + 1, 64):
        x = q[b:b + 64, 0]
        assert np.all(np.diff(x) >= 0)


@pytest.mark.parametrize("n", [50, 3000])
def test_order_with_non_finite_positions_is_a_permutation(n):
    # NaN sorts as +infinity, so the sort and the partition keep a strict
    # weak order: the result is still a permutation of the ions
    rng = np.random.default_rng(7)
    pos = rng.normal(0, 1e-3, (n, 3))
    bad = rng.choice(n, max(1, n // 20), replace=False)
    pos[bad, rng.integers(0, 3, bad.size)] = np.nan
    pos[bad[:1], 0] = np.inf
    order = S.spatial_order(pos)
    assert np.array_equal(np.sort(order), np.arange(n))
