"""The Jaccard Keeper (IterativeBlockingKeeper, blocking.cpp:431-547) on the
R-MAT scale-18 workload matrix (262144 rows, seed 1002), 16-column blocks and
16-row groups, at tau 0.1 / 0.3 / 0.5 / 0.7, against the reference's own
compiled code (tests/golden/keeper_rmat18_golden.npz, written by
tests/golden/make_golden_rmat18.py): the grouping and both counters bit for
bit, the fixed-size permutation (get_fixed_size_grouping + get_permutation,
utilities.cpp:8-54) by its SHA-256, the blocking info of the fixed-size
grouping (CollectBlockingInfo, blocking.cpp:562-614), and the threshold
search's choice (reorder_matrix.cpp:17-40: fewest nonzero blocks, ties to
the later tau)."""
import hashlib
import os
import time

import numpy as np
import pytest

from smat import gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TAUS = (0.1, 0.3, 0.5, 0.7)


@pytest.mark.timeout(900)
def test_keeper_rmat18_tau_sweep_bit_exact(golden_dir):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import smat
    gold = np.load(os.path.join(golden_dir, "keeper_rmat18_golden.npz"))
    scale, seed, cb, rb = (int(x) for x in gold["params"])
    M, K, rp, ci, _ = gen.rmat(scale=scale, seed=seed)
    rows = 1 << scale
    rp = rp[: rows + 1]
    d_rp, d_ci = torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda()
    blocks = {}
    for tau in TAUS:
        key = f"t{tau}"
        t0 = time.perf_counter()
        g, cnt = smat.reorder_grouping(rp, ci, np.float32(tau), cb, rb)
        dt = time.perf_counter() - t0
        print(f"Keeper s{scale} tau {tau}: {dt:.1f} s (reference {float(gold[key + '__ref_seconds'][0]):.0f} s on one "
              f"core), counters {cnt.tolist()}", flush=True)
        assert np.array_equal(cnt, gold[key + "__counters"]), tau
        assert np.array_equal(g, gold[key + "__grouping"].astype(np.int64)), tau
        fixed = smat.fixed_size_grouping(g, rb)
        perm = smat.grouping_to_permutation(fixed)
        digest = np.frombuffer(hashlib.sha256(perm.astype(np.int64).tobytes()).digest(), np.uint8)
        assert np.array_equal(digest, gold[key + "__perm_sha256"]), tau
        info, avg = smat.blocking_info(rows, K, d_rp, d_ci, torch.from_numpy(fixed).cuda(), cb)
        assert np.array_equal(info, gold[key + "__info"][:3]), tau
        assert int(round(avg * 1e6)) == int(gold[key + "__info"][3]), tau
        blocks[tau] = int(info[1])
    # the threshold search: start from the nonzero count, `<=` keeps the later tau on ties
    best_tau, best = -1.0, int(ci.size)
    ref_tau, ref_best = -1.0, int(ci.size)
    for tau in TAUS:
        if blocks[tau] <= best:
            best_tau, best = tau, blocks[tau]
        if int(gold[f"t{tau}__info"][1]) <= ref_best:
            ref_tau, ref_best = tau, int(gold[f"t{tau}__info"][1])
    assert (best_tau, best) == (ref_tau, ref_best)
