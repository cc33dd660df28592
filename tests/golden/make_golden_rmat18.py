"""Writes tests/golden/keeper_rmat18_golden.npz: the reference's own Keeper
(IterativeBlockingKeeper, preprocess/src/general/blocking.cpp:431-547,
compiled from the reference sources into oracle/_ref/ by `make -C oracle ref`)
on the R-MAT scale-18 matrix of the workload (262144 rows, 16 edge draws per
row, (a,b,c,d) = (.57,.19,.19,.05), seed 1002), 16-column blocks, 16-row
groups, for tau in {0.1, 0.3, 0.5, 0.7}, followed by get_fixed_size_grouping,
get_permutation (utilities.cpp:8-54) and CollectBlockingInfo
(blocking.cpp:562-614), as the threshold search of reorder_matrix.cpp:17-40
runs them.

Per tau the file keeps the counters [comparison_counter, merge_counter], the
blocking info of the fixed-size grouping [VBR_nzcount, VBR_nzblocks_count,
VBR_longest_row, round(1e6 * VBR_average_height)], the Keeper grouping
(int32) and the SHA-256 of the fixed-size permutation (int64 bytes; the
permutation itself is 262144 incompressible indices).  The matrix itself is
re-generated from its seed by the tests.  One reference run takes about 13
minutes on one core, so the four taus run as four processes.

Run:  python tests/golden/make_golden_rmat18.py   (this container only)
"""
from __future__ import annotations

import hashlib
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "sptc-smat_amd"))

TAUS = (0.1, 0.3, 0.5, 0.7)
SCALE, SEED, COL_BLOCK, ROW_BLOCK = 18, 1002, 16, 16


def matrix():
    from smat import gen
    M, K, rp, ci, _ = gen.rmat(scale=SCALE, seed=SEED)
    rows = 1 << SCALE
    return rows, K, rp[: rows + 1].astype(np.int64), ci.astype(np.int64)


def one(tau):
    import oracle_ffi as of
    rows, K, rp, ci = matrix()
    t0 = time.time()
    g, cnt, t = of.keeper(rp, ci, np.float32(tau), COL_BLOCK, ROW_BLOCK, cols=K, use_ref=True)
    fixed = of.fixed_size_grouping(g, ROW_BLOCK, use_ref=True)
    perm = of.get_permutation(fixed, use_ref=True)
    info, avg = of.collect_blocking_info(rp, ci, fixed, K, COL_BLOCK, use_ref=True)
    print(f"tau {tau}: {time.time() - t0:.0f} s, counters {cnt.tolist()}, info {info.tolist()}", flush=True)
    digest = np.frombuffer(hashlib.sha256(perm.astype(np.int64).tobytes()).digest(), np.uint8)
    return tau, cnt, np.append(info, np.int64(round(avg * 1e6))), g.astype(np.int32), digest, t


def main():
    import oracle_ffi as of
    assert of.ref_available(), "build oracle/_ref first: make -C oracle ref"
    with mp.get_context("spawn").Pool(len(TAUS)) as pool:
        res = pool.map(one, TAUS)
    arrays = {"taus": np.array(TAUS, np.float32), "params": np.array([SCALE, SEED, COL_BLOCK, ROW_BLOCK], np.int64)}
    for tau, cnt, info, g, perm, t in res:
        key = f"t{tau}"
        arrays[key + "__counters"] = cnt
        arrays[key + "__info"] = info
        arrays[key + "__grouping"] = g
        arrays[key + "__perm_sha256"] = perm
        arrays[key + "__ref_seconds"] = np.array([t], np.float64)
    np.savez_compressed(os.path.join(HERE, "keeper_rmat18_golden.npz"), **arrays)
    print("written", os.path.join(HERE, "keeper_rmat18_golden.npz"))


if __name__ == "__main__":
    main()
