"""Files end to end on the GPU: the reference's own edge-list fixtures
(tests/golden/*.el, read as CSR::read_from_edgelist does), a Matrix Market
file (read as mmio_allinone does) and an MTX file read in the edge-list
reader's MTX mode, each padded (rows to 16, columns to the block width, as
padSparseMatrix, mmio_highlevel.h:514-520), built into BCSR on the GPU and
multiplied, against the oracle's FP64 product of the same FP16 values."""
import os

import numpy as np
import pytest

import oracle_ffi as of
from smat import gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _pad(c, k=32):
    """Pads a reader's CSR: rows to a multiple of 16, columns to a multiple of k."""
    rows, cols = int(c["rows"]), int(c["cols"])
    M, K = gen.pad_to(max(rows, 1), 16), gen.pad_to(max(cols, 1), k)
    rp = np.concatenate([c["rowptr"], np.full(M - rows, c["rowptr"][-1], np.int32)]).astype(np.int32)
    return M, K, rp, c["colidx"].astype(np.int32), c["values_f16"]


def _spmm_check(M, K, rp, ci, v, N=64):
    import smat
    B = gen.dense_b(K, N, seed=12)
    R = of.spmm_csr_f64(M, rp, ci, v, B)
    for bk in (1, 16):
        A = smat.Bcsr(M, K, _dev(rp), _dev(ci), _dev(v.view(np.int16)), bk=bk)
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        smat.SpmmPlan(A, N)(_dev(B.view(np.int16)), C)
        torch.cuda.synchronize()
        Cn = C.cpu().numpy().astype(np.float64)
        assert np.all(np.isfinite(Cn))
        assert np.linalg.norm(Cn - R) <= 1e-6 * max(np.linalg.norm(R), 1e-30), bk
        A.close()


@pytest.mark.parametrize("name", ["TEST_matrix_weighted.el", "cs_department.el", "seventh_graders.el"])
@pytest.mark.parametrize("pattern_only", [False, True])
def test_reference_edgelist_files_through_spmm(name, pattern_only):
    _need_gpu()
    import smat
    c = smat.read_edgelist(os.path.join(GOLDEN, name), pattern_only=pattern_only)
    print(f"{name}: {c['rows']} x {c['cols']}, {c['nnz']} entries", flush=True)
    _spmm_check(*_pad(c))


def test_matrix_market_file_through_spmm(tmp_path):
    _need_gpu()
    import smat
    rng = np.random.default_rng(4)
    n = 300
    cells = {tuple(sorted(rng.integers(1, n + 1, 2).tolist(), reverse=True)) for _ in range(1500)}
    lines = [f"{i} {j} {rng.uniform(-2, 2):.5f}" for i, j in sorted(cells, key=lambda t: (t[1], t[0]))]
    text = "%%MatrixMarket matrix coordinate real symmetric\n% test\n" + f"{n} {n} {len(lines)}\n" + "\n".join(lines)
    p = tmp_path / "sym.mtx"
    p.write_text(text + "\n")
    for k in (16, 32):
        c = smat.read_mtx(str(p), k=k)
        assert c["symmetric"] and c["rows"] % 16 == 0 and c["cols"] % k == 0
        _spmm_check(int(c["rows"]), int(c["cols"]), c["rowptr"], c["colidx"], c["values_f16"])
    # the edge-list reader's MTX mode on the same file: the index fields swapped, indices kept
    e = smat.read_edgelist(str(p), mat_fmt=smat.SMAT_FMT_MTX)
    _spmm_check(*_pad(e))
