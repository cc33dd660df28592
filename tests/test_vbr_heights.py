"""Variable-height (VBR) blocking against fixed 16-row blocking, the
measurement behind building no variable-height SpMM kernel.

SPARTA's VBR format (VBR::fill_from_CSR_inplace, preprocess/src/general/
vbr.cpp:135-237) stores each group of rows as one block row of its own
height; the SMaT pipeline instead cuts the Keeper's grouping into fixed
16-row groups (get_fixed_size_grouping, utilities.cpp:8-20;
reorder.py:131 runs the Keeper with -F 1).  On the matrix cores a block row
of height h occupies ceil(h / 16) 16-row MFMA tiles per stored block column,
so the work of a grouping is sum over its stored blocks of ceil(h / 16).
Variable heights come from the reference's algorithm 0
(IterativeBlockingPattern, blocking.cpp:87-152): the Keeper
(blocking.cpp:431-547) tops every group up to 16 rows while free rows
remain, so nearly all its groups are 16 high.  On an R-MAT graph and a band matrix the Keeper's
fixed 16-row blocking needs fewer MFMA tiles than algorithm 0's variable
heights at every tau tried, although the variable blocks store fewer
elements."""
import numpy as np
import pytest

import oracle_ffi as of
from smat import gen


def _blocking(rp, ci, grouping, b=16):
    """(stored blocks, stored elements, 16-row MFMA tiles) of a grouping
    with b-wide column blocks: a group of h rows with c distinct column
    blocks stores c blocks of h x b elements and needs c * ceil(h / 16)
    tiles."""
    _, g = np.unique(np.asarray(grouping), return_inverse=True)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    keys = np.unique(g[rows].astype(np.int64) * (1 << 32) + ci // b)
    per = np.bincount((keys >> 32).astype(np.int64), minlength=g.max() + 1)
    h = np.bincount(g, minlength=g.max() + 1)
    return int(keys.size), int((per * h).sum() * b), int((per * -(-h // 16)).sum())


def test_blocking_counts_against_collect_blocking_info():
    # the counts above are CollectBlockingInfo's VBR_nzblocks_count and VBR_nzcount
    M, K, rp, ci, v = gen.rmat(scale=9, edge_factor=8, seed=5)
    rp64, ci64 = rp.astype(np.int64), ci.astype(np.int64)
    g, _ = of.algo0(rp64, ci64, 0.5, 16)
    (nzcount, nzblocks, _), _ = of.collect_blocking_info(rp64, ci64, g, K, 16)
    nb, stored, _ = _blocking(rp, ci, g)
    assert (nb, stored) == (nzblocks, nzcount)


@pytest.mark.parametrize("name,make", [
    ("rmat11", lambda: gen.rmat(scale=11, edge_factor=16, seed=3)),
    ("band2048", lambda: gen.band(n=2048, hb=64, p=0.5, seed=1001)),
])
@pytest.mark.parametrize("tau", [0.3, 0.5, 0.7])
def test_keeper_fixed_16_needs_fewer_mfma_tiles_than_variable_heights(name, make, tau):
    M, K, rp, ci, v = make()
    rp64, ci64 = rp.astype(np.int64), ci.astype(np.int64)
    g0, _ = of.algo0(rp64, ci64, tau, 16)
    gk, _ = of.keeper(rp64, ci64, tau, 16, 16)
    heights = np.bincount(np.unique(gk, return_inverse=True)[1])
    assert heights[heights == 16].sum() >= 0.95 * M  # the Keeper's groups are (almost) all 16 high
    var = _blocking(rp, ci, g0)
    fixed = _blocking(rp, ci, of.fixed_size_grouping(gk, 16))
    print(f"{name} tau {tau}: variable {var}, Keeper fixed-16 {fixed}")
    assert fixed[2] < var[2]
