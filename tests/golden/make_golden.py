"""Writes the golden vectors under tests/golden/ by running the reference's own
serial preprocessing code (compiled from the reference sources by
`make -C oracle ref` into oracle/_ref/, in this container only).

Inputs:
  * the reference's fixture matrices preprocess/data/TEST_matrix_weighted.el
    and data/real_world/{cs_department,seventh_graders}.el (their edges are
    stored with the vectors, so the tests never need the reference checkout);
  * seeded synthetic band and R-MAT matrices (re-generated from their seeds
    by the tests).

Outputs (all numpy .npz without pickles, or JSON):
  kat_test_matrix.json     the reference's recorded TEST_results.txt KAT plus
                           the algo-0 grouping the compiled reference prints
  keeper_golden.npz        IterativeBlockingKeeper groupings + counters,
                           fixed-size groupings, permutations and blocking
                           info for every (matrix, tau, col_block, max_rows)
  jaccard_merge_golden.npz JaccardDistanceGroup and merge_rows on random rows
  csr_multiply_golden.npz  CSR::multiply on a small square matrix

Run:  python tests/golden/make_golden.py   (needs oracle/_ref/libsmat_ref.so)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "sptc-smat_amd"))

import oracle_ffi as of  # noqa: E402
from smat import gen  # noqa: E402

REF = "/root/reference"


def read_text(path):
    with open(path) as f:
        return f.read()


def synthetic_inputs():
    out = {}
    M, K, rp, ci, _ = gen.band(n=1024, hb=24, p=0.3, seed=11)
    out["band1024"] = (rp.astype(np.int64), ci.astype(np.int64), K)
    M, K, rp, ci, _ = gen.rmat(scale=11, edge_factor=8, seed=12)
    out["rmat11"] = (rp.astype(np.int64), ci.astype(np.int64), K)
    return out


def main():
    assert of.ref_available(), "build oracle/_ref first: make -C oracle ref"
    # ---- KAT: preprocess/results/TEST_results.txt on TEST_matrix_weighted.el
    t = read_text(f"{REF}/preprocess/data/TEST_matrix_weighted.el")
    rows, cols, rp, ci, vals = of.read_edgelist(t, weighted=True)
    g, cnt = of.algo0(rp, ci, 0.5, 3, 0, cols=cols, use_ref=True)
    info, avg = of.collect_blocking_info(rp, ci, g, cols, 3, use_ref=True)
    kat = {
        "source": "preprocess/results/TEST_results.txt:1-2 (algo 0, tau 0.5, b=B=3)",
        "edges": [[int(r), int(c), float(v)] for r, c, v in
                  zip(np.repeat(np.arange(rows), np.diff(rp)), ci, vals)],
        "rows": rows, "cols": cols, "nnz": int(ci.size),
        "VBR_nzcount": 36, "VBR_nzblocks_count": 5, "VBR_average_height": 2.4,
        "VBR_longest_row": 3, "merge_counter": 5, "comparison_counter": 17,
        "compiled_reference": {"grouping": g.tolist(), "counters": cnt.tolist(),
                               "info": info.tolist(), "avg_height": avg},
    }
    with open(os.path.join(HERE, "kat_test_matrix.json"), "w") as f:
        json.dump(kat, f, indent=1)

    # ---- Keeper goldens
    mats = {}
    for name in ("cs_department", "seventh_graders"):
        t = read_text(f"{REF}/preprocess/data/real_world/{name}.el")
        r, c, rp, ci, _ = of.read_edgelist(t, weighted=False)
        mats[name] = (rp, ci, c)
    r, c, rp, ci, _ = of.read_edgelist(read_text(f"{REF}/preprocess/data/TEST_matrix_weighted.el"), True)
    mats["test_matrix"] = (rp, ci, c)
    mats.update(synthetic_inputs())
    arrays = {}
    cases = []
    for name, (rp, ci, cols) in mats.items():
        arrays[f"{name}__rowptr"] = rp
        arrays[f"{name}__colidx"] = ci
        arrays[f"{name}__cols"] = np.array([cols], np.int64)
        for tau in (0.1, 0.3, 0.5, 0.7):
            for cb, mr in ((8, 16), (16, 16), (3, 3)):
                g, cnt, _ = of.keeper(rp, ci, tau, cb, mr, cols=cols, use_ref=True)
                fsg = of.fixed_size_grouping(g, mr, use_ref=True)
                perm = of.get_permutation(fsg, use_ref=True)
                info, avg = of.collect_blocking_info(rp, ci, fsg, cols, cb, use_ref=True)
                key = f"{name}__t{tau}_b{cb}_B{mr}"
                arrays[key + "__grouping"] = g
                arrays[key + "__counters"] = cnt
                arrays[key + "__fixed"] = fsg
                arrays[key + "__perm"] = perm
                arrays[key + "__info"] = np.append(info, np.int64(round(avg * 1e6)))
                cases.append(key)
    arrays["cases"] = np.array(cases)
    np.savez_compressed(os.path.join(HERE, "keeper_golden.npz"), **arrays)

    # ---- Jaccard / merge_rows on random sorted rows
    rng = np.random.default_rng(99)
    jm = {}
    A_list, B_list, jac, mer, params = [], [], [], [], []
    for i in range(400):
        na, nb = rng.integers(0, 12, size=2)
        a = np.unique(rng.integers(0, 64, size=na)).astype(np.int64)
        b = np.unique(rng.integers(0, 64, size=nb)).astype(np.int64)
        ga, blk = int(rng.integers(1, 16)), int(rng.choice([1, 3, 4, 8, 16]))
        jac.append(of.jaccard(a, ga, b, 1, blk, use_ref=True))
        m = of.merge_rows(a, b, use_ref=True)
        A_list.append(a)
        B_list.append(b)
        mer.append(m)
        params.append((ga, blk))
    def pack(lst):
        lens = np.array([x.size for x in lst], np.int64)
        return np.concatenate([np.zeros(0, np.int64)] + lst), lens
    jm["a_vals"], jm["a_lens"] = pack(A_list)
    jm["b_vals"], jm["b_lens"] = pack(B_list)
    jm["m_vals"], jm["m_lens"] = pack(mer)
    jm["jaccard"] = np.array(jac, np.float32)
    jm["params"] = np.array(params, np.int64)
    np.savez_compressed(os.path.join(HERE, "jaccard_merge_golden.npz"), **jm)

    # ---- CSR::multiply
    rng = np.random.default_rng(5)
    n, bc = 48, 3
    dense = (rng.random((n, n)) < 0.15) * rng.uniform(-1, 1, (n, n))
    rp = np.concatenate([[0], np.cumsum((dense != 0).sum(1))]).astype(np.int64)
    ci = np.nonzero(dense)[1].astype(np.int64)
    vals = dense[dense != 0].astype(np.float32)
    B = rng.uniform(-1, 1, n * bc).astype(np.float32)
    C = np.zeros(n * bc, np.float32)
    of.ref().ref_csr_multiply(n, n, of.P(rp), of.P(ci), of.P(vals), of.P(B), bc, of.P(C))
    np.savez_compressed(os.path.join(HERE, "csr_multiply_golden.npz"), rowptr=rp, colidx=ci, vals=vals,
                        B=B, C=C, n=np.array([n, bc]))
    print("golden vectors written:", len(cases), "keeper cases")


if __name__ == "__main__":
    main()
