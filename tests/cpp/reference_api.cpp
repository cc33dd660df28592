// Drives the reference-named C++ API (include/smat_reference.hpp) the way
// the reference's own code calls it, and writes what it produced as raw
// little-endian arrays for tests/test_gpu_reference_api.py to check:
//   reference_api <matrix.mtx> <graph.el> <tau> <col_block> <row_block> <out_dir>
//   1. SparseMatrix A("A", mtx, 16): its BCSR, grid maps and filtered lists
//      (host copies), then mmaOBTKernel on its device arrays with a B of
//      K x 64 FP16 values (B[n*K + k], written to out_dir/B.bin) into C.
//   2. A.bcsrBlocking(): the merged 64x64 tiles (host copies) and
//      mmaOBTKernel_tiled on them into C_tiled; SparseMatrix A32("A32", mtx,
//      32), bcsrBlocking(), preprocessing_mmaOBTSKernel_tiled_large and
//      mmaOBTSKernel_tiled_large (buffers sized as tester.h:385-394) into
//      C_tiled_large, with its own B (B32.bin).
//   3. CSR(std::ifstream&) of the edge list (pattern only, el format), a
//      BlockingEngine with the Keeper (iterative_max_size), GetGrouping,
//      get_fixed_size_grouping, CollectBlockingInfo on the fixed grouping,
//      CSR::reorder with it; the reordered pattern is written out.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "smat_reference.hpp"

static std::vector<uint16_t> random_b(size_t n, uint32_t s) {
  std::vector<uint16_t> b(n);
  for (auto& x : b) {  // FP16 values in [-1, 1): sign, exponent 14 or 13, random mantissa
    s = s * 1664525u + 1013904223u;
    x = (uint16_t)(((s >> 31) << 15) | ((13u + ((s >> 30) & 1u)) << 10) | ((s >> 8) & 0x3ffu));
  }
  return b;
}

template <typename T>
static T* dev_alloc(size_t n) {
  T* p = nullptr;
  smat_ref_detail::hip_check(hipMalloc(reinterpret_cast<void**>(&p), (n ? n : 1) * sizeof(T)), "hipMalloc");
  smat_ref_detail::hip_check(hipMemset(p, 0, (n ? n : 1) * sizeof(T)), "hipMemset");
  return p;
}

static void dump(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || (bytes && std::fwrite(p, 1, bytes, f) != bytes)) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s matrix.mtx graph.el tau col_block row_block out_dir\n", argv[0]);
    return 2;
  }
  const std::string out = argv[6];
  try {
    // ---- 1. SparseMatrix + mmaOBTKernel
    SparseMatrix A("A", argv[1], 16);
    const size_t M = A.getRow(), K = A.getCol(), nb = A.getNonzeroblocks(), N = 64;
    const size_t grid = (M / 16) * (K / 16);
    std::printf("SparseMatrix %zu x %zu nnz %zu blocks %zu sparse %zu dense %zu\n", M, K, A.getNnz(), nb,
                A.getSparseBlocks(), A.getDenseBlocks());
    dump(out + "/bcsr_rowptr.bin", A.getBcsrRowPtrHost(), (M / 16 + 1) * 4);
    dump(out + "/bcsr_colidx.bin", A.getBcsrColIdxHost(), nb * 4);
    dump(out + "/bcsr_values.bin", A.getBcsrValuesHost(), nb * 256 * 2);
    dump(out + "/block_info.bin", A.getBlockInfo_host(), grid * 4);
    dump(out + "/rel_map.bin", A.getRelativeBlockIndexMapping_host(), grid * 4);
    dump(out + "/sparse_rowptr.bin", A.getSparseBcsrRowPtrHost(), (M / 16 + 1) * 4);
    dump(out + "/sparse_colidx.bin", A.getSparseBcsrColIdxHost(), A.getSparseBlocks() * 4);
    dump(out + "/dense_rowptr.bin", A.getDenseBcsrRowPtrHost(), (M / 16 + 1) * 4);
    dump(out + "/dense_colidx.bin", A.getDenseBcsrColIdxHost(), A.getDenseBlocks() * 4);
    dump(out + "/sparse_rel_map.bin", A.getSparseRelativeBlockIndexMappingHost(), grid * 4);
    const std::vector<uint16_t> Bh = random_b(K * N, 12345u);
    dump(out + "/B.bin", Bh.data(), Bh.size() * 2);
    half *dB = nullptr, *dC = nullptr;
    smat_ref_detail::hip_check(hipMalloc(reinterpret_cast<void**>(&dB), Bh.size() * 2), "hipMalloc");
    smat_ref_detail::hip_check(hipMalloc(reinterpret_cast<void**>(&dC), M * N * 2), "hipMalloc");
    smat_ref_detail::hip_check(hipMemcpy(dB, Bh.data(), Bh.size() * 2, hipMemcpyHostToDevice), "hipMemcpy");
    mmaOBTKernel(A.getBcsrValues(), A.getBcsrRowPtr(), A.getBcsrColIdx(), dB, dC, M, N, K, nb, A.getBlockInfo_dev(),
                 A.getRelativeBlockIndexMapping_dev());
    if (smat_last_status() != SMAT_OK) throw std::runtime_error(smat_last_error());
    std::vector<uint16_t> Ch(M * N);
    smat_ref_detail::hip_check(hipMemcpy(Ch.data(), dC, Ch.size() * 2, hipMemcpyDeviceToHost), "hipMemcpy");
    dump(out + "/C.bin", Ch.data(), Ch.size() * 2);
    // ---- 2. merged tiles + the tiled launchers
    A.bcsrBlocking();
    const size_t nt = A.getMergedNonzeroBlocks(), ntr = (M + 63) / 64, ntc = (K + 63) / 64;
    std::printf("bcsrBlocking: %zu tiles of 64 x 64\n", nt);
    dump(out + "/tiles_rowptr.bin", A.getMergedBcsrRowPtrHost(), (ntr + 1) * 4);
    dump(out + "/tiles_colidx.bin", A.getMergedBcsrColIdxHost(), nt * 4);
    dump(out + "/tiles_values.bin", A.getMergedBcsrValuesHost(), nt * 16 * 256 * 2);
    dump(out + "/tiles_info.bin", A.getMergedTileInfo_host(), nt * 16 * 4);
    dump(out + "/tiles_block_info.bin", A.getMergedBlockInfo_host(), ntr * ntc * 4);
    dump(out + "/tiles_rel_map.bin", A.getMergedRelativeBlockIndexMapping_host(), ntr * ntc * 4);
    mmaOBTKernel_tiled(A.getMergedBcsrValues(), A.getMergedBcsrRowPtr(), A.getMergedBcsrColIdx(), dB, dC, M, N, K,
                       nb, A.getMergedBlockInfo_dev(), A.getMergedRelativeBlockIndexMapping_dev());
    if (smat_last_status() != SMAT_OK) throw std::runtime_error(smat_last_error());
    smat_ref_detail::hip_check(hipMemcpy(Ch.data(), dC, Ch.size() * 2, hipMemcpyDeviceToHost), "hipMemcpy");
    dump(out + "/C_tiled.bin", Ch.data(), Ch.size() * 2);
    (void)hipFree(dB);
    (void)hipFree(dC);

    SparseMatrix A32("A32", argv[1], 32);
    A32.bcsrBlocking();
    const size_t M32 = A32.getRow(), K32 = A32.getCol(), nb32 = A32.getNonzeroblocks();
    std::printf("SparseMatrix k=32 %zu x %zu blocks %zu tiles %zu\n", M32, K32, nb32, A32.getMergedNonzeroBlocks());
    const std::vector<uint16_t> B32 = random_b(K32 * N, 777u);
    dump(out + "/B32.bin", B32.data(), B32.size() * 2);
    half* dB32 = dev_alloc<half>(B32.size());
    half* dC32 = dev_alloc<half>(M32 * N);
    smat_ref_detail::hip_check(hipMemcpy(dB32, B32.data(), B32.size() * 2, hipMemcpyHostToDevice), "hipMemcpy");
    char* meta = dev_alloc<char>(nb32 * 16 * 16 * 4);          // tester.h:385-388 (bytes)
    half* sparseA = dev_alloc<half>(nb32 * 16 * 4 * 16 * 4);   // tester.h:393-394 (halves)
    preprocessing_mmaOBTSKernel_tiled_large(A32.getMergedBcsrValues(), meta, sparseA, M32, N, K32, nb32,
                                            A32.getMergedBlockInfo_dev(), A32.getMergedRelativeBlockIndexMapping_dev(),
                                            A32.getMergedTileInfo_dev());
    if (smat_last_status() != SMAT_OK) throw std::runtime_error(smat_last_error());
    mmaOBTSKernel_tiled_large(A32.getMergedBcsrValues(), A32.getMergedBcsrRowPtr(), A32.getMergedBcsrColIdx(), meta,
                              sparseA, dB32, dC32, M32, N, K32, nb32, A32.getMergedBlockInfo_dev(),
                              A32.getMergedRelativeBlockIndexMapping_dev(), A32.getMergedTileInfo_dev());
    if (smat_last_status() != SMAT_OK) throw std::runtime_error(smat_last_error());
    std::vector<uint16_t> C32(M32 * N);
    smat_ref_detail::hip_check(hipMemcpy(C32.data(), dC32, C32.size() * 2, hipMemcpyDeviceToHost), "hipMemcpy");
    dump(out + "/C_tiled_large.bin", C32.data(), C32.size() * 2);
    for (void* p : {(void*)dB32, (void*)dC32, (void*)meta, (void*)sparseA}) (void)hipFree(p);

    // ---- 3. CSR + BlockingEngine (Keeper) + reorder
    std::ifstream fin(argv[2]);
    CSR cmat(fin, " ", true, el);
    BlockingEngine bmat;
    bmat.tau = std::stof(argv[3]);
    bmat.col_block_size = std::stol(argv[4]);
    bmat.row_block_size = std::stol(argv[5]);
    bmat.blocking_algo = iterative_max_size;
    std::vector<intT> grouping = bmat.GetGrouping(cmat);
    std::vector<intT> fixed = get_fixed_size_grouping(grouping, bmat.row_block_size);
    bmat.grouping_result = fixed;
    bmat.CollectBlockingInfo(cmat);
    std::printf("CSR %ld x %ld nnz %ld comparisons %ld merges %ld VBR_nzcount %ld VBR_nzblocks_count %ld "
                "VBR_longest_row %ld VBR_average_height %.6f\n",
                cmat.rows, cmat.cols, cmat.nztot(), bmat.comparison_counter, bmat.merge_counter, bmat.VBR_nzcount,
                bmat.VBR_nzblocks_count, bmat.VBR_longest_row, bmat.VBR_average_height);
    const int64_t counters[6] = {bmat.comparison_counter, bmat.merge_counter, bmat.VBR_nzcount,
                                 bmat.VBR_nzblocks_count, bmat.VBR_longest_row,
                                 (int64_t)std::llround(bmat.VBR_average_height * 1e6)};
    dump(out + "/counters.bin", counters, sizeof(counters));
    dump(out + "/grouping.bin", grouping.data(), grouping.size() * 8);
    dump(out + "/fixed.bin", fixed.data(), fixed.size() * 8);
    const std::vector<intT> perm = get_permutation(fixed);
    dump(out + "/perm.bin", perm.data(), perm.size() * 8);
    cmat.reorder(fixed);
    std::vector<int64_t> rp(cmat.rows + 1, 0), ci;
    for (intT i = 0; i < cmat.rows; i++) {
      rp[i + 1] = rp[i] + cmat.nzcount[i];
      for (intT q = 0; q < cmat.nzcount[i]; q++) ci.push_back(cmat.ja[i][q]);
    }
    dump(out + "/reordered_rowptr.bin", rp.data(), rp.size() * 8);
    dump(out + "/reordered_colidx.bin", ci.data(), ci.size() * 8);
    const std::vector<intT> part = get_partition(fixed);
    std::printf("partition blocks %zu\n", part.size() - 1);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::printf("done\n");
  return 0;
}
