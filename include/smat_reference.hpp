// smat_reference.hpp — the reference's C++ format-conversion and
// preprocessing API, header-only, over libsmat's C ABI (smat.h) and the HIP
// runtime.  Code written against the reference's headers
//   src/cuda_hgemm/src/common/matrix.h   (SparseMatrix)
//   preprocess/include/matrices.h         (CSR)
//   preprocess/include/blocking.h         (BlockingEngine)
//   preprocess/include/utilities.h        (get_permutation, get_partition,
//                                          get_fixed_size_grouping)
//   preprocess/include/definitions.h      (intT, DataT, MatrixFormat, BlockingType)
//   src/cuda_hgemm/src/main.cu:10-77      (the SpMM launchers, taking half*)
// includes this header instead and keeps its calls.  The work runs in
// libsmat: file parsing (smat_read_mtx, smat_parse_edgelist), CSR -> BCSR and
// the block filter on the GPU (smat_bcsr_from_csr, smat_bcsr_grid_maps,
// smat_bcsr_filter), the Keeper grouping with distances on the GPU
// (smat_reorder_grouping), the permutations (smat_grouping_to_permutation,
// smat_fixed_size_grouping) and the blocking statistics on the GPU
// (smat_blocking_info).
//
// Errors: where the reference throws std::invalid_argument (csr.cpp:69,101,
// 234) this header throws it too; where the reference prints and exits on a
// CUDA error (common.h:43-58), this header throws std::runtime_error with
// the library's message.  Only the blocking algorithms the SMaT pipeline
// runs are provided: iterative_max_size (IterativeBlockingKeeper, algorithm 5,
// reorder.py:131) and fixed_size; the others throw std::invalid_argument.
// Link with -lsmat and the HIP runtime (libamdhip64).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "smat.h"

// ---- definitions.h ----------------------------------------------------------
typedef long int intT;
typedef float DataT;
typedef float DataT_C;
enum MatrixFormat { el, mtx };
enum BlockingType { iterative, iterative_structured, fixed_size, iterative_clocked, iterative_queue, iterative_max_size };

namespace smat_ref_detail {

inline void check(smat_status st) {
  if (st != SMAT_OK) throw std::runtime_error(std::string("libsmat: ") + smat_last_error());
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// A device array owned by the caller of the C ABI (freed with hipFree).
template <class T>
struct DeviceArray {
  T* p = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { alloc(count); }
  DeviceArray(const T* host, size_t count) {
    alloc(count);
    if (count) hip_check(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  }
  void alloc(size_t count) {
    reset();
    n = count;
    hip_check(hipMalloc(reinterpret_cast<void**>(&p), (count ? count : 1) * sizeof(T)), "hipMalloc");
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* release() {
    T* q = p;
    p = nullptr;
    n = 0;
    return q;
  }
  ~DeviceArray() { reset(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
};

template <class T>
inline T* to_host(const T* dev, size_t count) {
  T* h = static_cast<T*>(std::malloc((count ? count : 1) * sizeof(T)));
  if (!h) throw std::runtime_error("host allocation failed");
  if (count) hip_check(hipMemcpy(h, dev, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  return h;
}

template <class T>
inline T* to_device(const T* host, size_t count) {
  DeviceArray<T> d(host, count);
  return d.release();
}

}  // namespace smat_ref_detail

// ---- utilities.h (utilities.cpp:8-54) ---------------------------------------
inline std::vector<intT> get_permutation(const std::vector<intT>& grouping) {
  std::vector<intT> v(grouping.size());
  static_assert(sizeof(intT) == sizeof(int64_t), "intT is 64-bit");
  smat_ref_detail::check(smat_grouping_to_permutation((int64_t)grouping.size(),
                                                      reinterpret_cast<const int64_t*>(grouping.data()),
                                                      reinterpret_cast<int64_t*>(v.data())));
  return v;
}

inline std::vector<intT> get_fixed_size_grouping(const std::vector<intT>& grouping, intT row_block_size) {
  std::vector<intT> out(grouping.size());
  smat_ref_detail::check(smat_fixed_size_grouping((int64_t)grouping.size(),
                                                  reinterpret_cast<const int64_t*>(grouping.data()), row_block_size,
                                                  reinterpret_cast<int64_t*>(out.data())));
  return out;
}

// Start of every group in sorted grouping order, then the row count.
inline std::vector<intT> get_partition(const std::vector<intT>& grouping) {
  std::vector<intT> sorted(grouping);
  std::sort(sorted.begin(), sorted.end());
  std::vector<intT> partition;
  for (size_t i = 0; i < sorted.size(); i++)
    if (i == 0 || sorted[i] != sorted[i - 1]) partition.push_back((intT)i);
  partition.push_back((intT)sorted.size());
  return partition;
}

// ---- matrices.h: CSR --------------------------------------------------------
struct CSR {
  intT rows = 0;
  intT cols = 0;
  intT* nzcount = nullptr;
  intT** ja = nullptr;
  DataT** ma = nullptr;
  bool pattern_only = true;

  // csr.cpp:166-290 (the parser is smat_parse_edgelist)
  void read_from_edgelist(std::ifstream& infile, std::string delimiter = " ", bool pattern_only = true,
                          MatrixFormat mat_fmt = mtx, bool symmetrize = false) {
    clean();
    std::string text((std::istreambuf_iterator<char>(infile)), std::istreambuf_iterator<char>());
    smat_csr_host h{};
    const smat_status st = smat_parse_edgelist(text.data(), (int64_t)text.size(), delimiter.c_str(),
                                               pattern_only ? 1 : 0, mat_fmt == mtx ? SMAT_FMT_MTX : SMAT_FMT_EL,
                                               symmetrize ? 1 : 0, &h);
    if (st != SMAT_OK) throw std::invalid_argument(std::string("CSR.read_from_edgelist: ") + smat_last_error());
    this->pattern_only = pattern_only;
    rows = h.rows;
    cols = h.cols;
    nzcount = new intT[rows];
    ja = new intT*[rows];
    ma = pattern_only ? nullptr : new DataT*[rows];
    for (intT i = 0; i < rows; i++) {
      const int32_t b = h.rowptr[i], e = h.rowptr[i + 1];
      nzcount[i] = e - b;
      ja[i] = new intT[e - b];
      for (int32_t q = b; q < e; q++) ja[i][q - b] = h.colidx[q];
      if (!pattern_only) {
        ma[i] = new DataT[e - b];
        std::copy(h.values_f32 + b, h.values_f32 + e, ma[i]);
      }
    }
    smat_csr_host_free(&h);
  }

  // csr.cpp:65-74: new row i is old row permutation[i]
  void permute_rows(std::vector<intT> permutation) {
    if ((intT)permutation.size() != rows)
      throw std::invalid_argument("CSR.permute_rows argument must have the same length as rows");
    std::vector<intT*> j2(rows);
    std::vector<DataT*> m2(pattern_only ? 0 : rows);
    std::vector<intT> n2(rows);
    for (intT i = 0; i < rows; i++) {
      j2[i] = ja[permutation[i]];
      n2[i] = nzcount[permutation[i]];
      if (!pattern_only) m2[i] = ma[permutation[i]];
    }
    std::copy(j2.begin(), j2.end(), ja);
    std::copy(n2.begin(), n2.end(), nzcount);
    if (!pattern_only) std::copy(m2.begin(), m2.end(), ma);
  }

  // csr.cpp:97-105: rows of equal grouping value become adjacent
  void reorder(std::vector<intT> grouping) {
    if ((intT)grouping.size() != rows)
      throw std::invalid_argument("CSR.reorder argument must have the same length as rows");
    permute_rows(get_permutation(grouping));
  }

  intT nztot() {
    intT t = 0;
    for (intT i = 0; i < rows; i++) t += nzcount[i];
    return t;
  }

  void clean() {
    if (ja)
      for (intT i = 0; i < rows; i++) delete[] ja[i];
    if (ma)
      for (intT i = 0; i < rows; i++) delete[] ma[i];
    delete[] ja;
    delete[] ma;
    delete[] nzcount;
    ja = nullptr;
    ma = nullptr;
    nzcount = nullptr;
    rows = 0;
    cols = 0;
  }

  CSR(std::ifstream& infile, std::string delimiter = " ", bool pattern_only = true, MatrixFormat mat_fmt = mtx) {
    read_from_edgelist(infile, delimiter, pattern_only, mat_fmt);
  }
  ~CSR() { clean(); }
  CSR(const CSR&) = delete;
  CSR& operator=(const CSR&) = delete;

  // The pattern as int32 CSR arrays (what the C ABI takes).
  void pattern_arrays(std::vector<int32_t>& rowptr, std::vector<int32_t>& colidx) const {
    rowptr.assign(rows + 1, 0);
    for (intT i = 0; i < rows; i++) rowptr[i + 1] = rowptr[i] + (int32_t)nzcount[i];
    colidx.resize(rowptr[rows]);
    for (intT i = 0; i < rows; i++)
      for (intT q = 0; q < nzcount[i]; q++) colidx[rowptr[i] + q] = (int32_t)ja[i][q];
  }
};

// ---- blocking.h: BlockingEngine ---------------------------------------------
class BlockingEngine {
 public:
  float tau = 0.5;
  intT col_block_size = 1;
  intT row_block_size = 1;
  bool use_groups = false;
  bool use_pattern = true;
  bool force_fixed_size = false;
  int structured_m = 2;
  int structured_n = 4;
  BlockingType blocking_algo = iterative_clocked;

  intT comparison_counter = 0;
  intT merge_counter = 0;
  float timer_total = 0;
  float timer_comparisons = 0;
  float timer_merges = 0;
  float average_row_distance = 0;
  float average_merge_tau = 0;

  intT VBR_nzcount = 0;
  intT VBR_nzblocks_count = 0;
  float VBR_average_height = 0;
  intT VBR_longest_row = 0;

  std::vector<intT> grouping_result;

  BlockingEngine() {}

  // blocking.cpp:616-657
  std::vector<intT> GetGrouping(const CSR& cmat) {
    comparison_counter = merge_counter = 0;
    timer_total = timer_comparisons = timer_merges = 0;
    average_row_distance = average_merge_tau = 0;
    grouping_result.assign(cmat.rows, 0);
    if (blocking_algo == iterative_max_size) {
      if (!use_pattern)
        throw std::invalid_argument("BlockingEngine: the Keeper is provided with use_pattern = true only");
      std::vector<int32_t> rp, ci;
      cmat.pattern_arrays(rp, ci);
      smat_ref_detail::DeviceArray<int32_t> d_rp(rp.data(), rp.size()), d_ci(ci.data(), ci.size());
      int64_t counters[2] = {0, 0};
      if (cmat.rows > 0)
        smat_ref_detail::check(smat_reorder_grouping(cmat.rows, d_rp.p, d_ci.p, rp.data(),
                                                     ci.empty() ? rp.data() : ci.data(), tau, col_block_size,
                                                     row_block_size,
                                                     reinterpret_cast<int64_t*>(grouping_result.data()), counters,
                                                     nullptr));
      comparison_counter = counters[0];
      merge_counter = counters[1];
    } else if (blocking_algo == fixed_size) {
      for (intT i = 0; i < cmat.rows; i++) grouping_result[i] = i / row_block_size;  // FixedBlocking
    } else {
      throw std::invalid_argument("BlockingEngine: only iterative_max_size and fixed_size are provided");
    }
    if (force_fixed_size && blocking_algo != fixed_size)
      grouping_result = get_fixed_size_grouping(grouping_result, row_block_size);
    return grouping_result;
  }

  // blocking.cpp:562-614, on grouping_result (smat_blocking_info on the GPU)
  void CollectBlockingInfo(const CSR& cmat) {
    std::vector<int32_t> rp, ci;
    cmat.pattern_arrays(rp, ci);
    smat_ref_detail::DeviceArray<int32_t> d_rp(rp.data(), rp.size()), d_ci(ci.data(), ci.size());
    static_assert(sizeof(intT) == sizeof(int64_t), "intT is 64-bit");
    smat_ref_detail::DeviceArray<int64_t> d_g(reinterpret_cast<const int64_t*>(grouping_result.data()),
                                              grouping_result.size());
    int64_t info[3] = {0, 0, 0};
    float avg = 0.f;
    smat_ref_detail::check(smat_blocking_info(cmat.rows, cmat.cols, (int64_t)ci.size(), d_rp.p, d_ci.p, d_g.p,
                                              col_block_size, info, &avg, nullptr));
    VBR_nzcount = info[0];
    VBR_nzblocks_count = info[1];
    VBR_longest_row = info[2];
    VBR_average_height = avg;
  }
};

// ---- matrix.h: SparseMatrix -------------------------------------------------
// Built from a Matrix Market file as the reference's constructor does
// (matrix.h:161-246): read and pad (mmio_allinone + padSparseMatrix), upload
// the CSR, csrToBcsr() and filterBcsrBlocks() (on the GPU here), keeping
// host and device copies of every array the reference exposes.
class SparseMatrix {
 public:
  SparseMatrix(const std::string& name = "Matrix", const char* file = "./data/cop20k_A.mtx", int k = 16)
      : m_name(name), filename(file ? file : ""), mMMA_K(k) {
    smat_csr_host h{};
    smat_ref_detail::check(smat_read_mtx(filename.c_str(), k, &h));
    m_row = (size_t)h.rows;
    m_col = (size_t)h.cols;
    nnz = (size_t)h.nnz;
    csrRowPtr_host = h.rowptr;
    csrColIdx_host = h.colidx;
    csrVal_host = reinterpret_cast<half*>(h.values_f16);
    std::free(h.values_f32);
    csrRowPtr_dev = smat_ref_detail::to_device(csrRowPtr_host, m_row + 1);
    csrColIdx_dev = smat_ref_detail::to_device(csrColIdx_host, nnz);
    csrVal_dev = smat_ref_detail::to_device(csrVal_host, nnz);
    csrToBcsr();
    filterBcsrBlocks();
  }

  // matrix.h:666-792: BCSR of 16 x k blocks, the reference's block tags and
  // its dense-grid maps
  void csrToBcsr() {
    release_bcsr();
    smat_ref_detail::check(smat_bcsr_from_csr((int64_t)m_row, (int64_t)m_col, (int64_t)nnz, csrRowPtr_dev,
                                              csrColIdx_dev, reinterpret_cast<const uint16_t*>(csrVal_dev), mMMA_K,
                                              SMAT_TAG_REFERENCE, nullptr, &bcsr_));
    smat_ref_detail::check(smat_bcsr_describe(bcsr_, &desc_));
    nonzeroBlocks = (size_t)desc_.nnzb;
    numberOfBlocks = (m_row / 16) * ((m_col + mMMA_K - 1) / mMMA_K);
    blockInfo_dev = smat_ref_detail::DeviceArray<int>(numberOfBlocks).release();
    relativeBlockIndexMapping_dev = smat_ref_detail::DeviceArray<int>(numberOfBlocks).release();
    smat_ref_detail::check(smat_bcsr_grid_maps(&desc_, blockInfo_dev, relativeBlockIndexMapping_dev, nullptr));
    hip_sync();
    bcsrRowPtr_host = smat_ref_detail::to_host(desc_.rowptr, m_row / 16 + 1);
    bcsrColIdx_host = smat_ref_detail::to_host(desc_.colidx, nonzeroBlocks);
    bcsrVal_host = reinterpret_cast<half*>(smat_ref_detail::to_host(desc_.values, nonzeroBlocks * 16 * mMMA_K));
    blockInfo_host = smat_ref_detail::to_host(blockInfo_dev, numberOfBlocks);
    relativeBlockIndexMapping_host = smat_ref_detail::to_host(relativeBlockIndexMapping_dev, numberOfBlocks);
  }

  // matrix.h:1140-1288: the 2:4 (tag 1) and dense (tag 2) block lists, the
  // all-ones sparse block info and the sparse blocks' index over the grid
  void filterBcsrBlocks() {
    release_filter();
    int64_t ns = 0, nd = 0;
    smat_ref_detail::check(smat_bcsr_filter(&desc_, &ns, &nd, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                            nullptr));
    sparseBlocks = (size_t)ns;
    denseBlocks = (size_t)nd;
    const size_t nb1 = m_row / 16 + 1, bs = 16 * (size_t)mMMA_K;
    sparseBcsrRowPtr_dev = smat_ref_detail::DeviceArray<int>(nb1).release();
    sparseBcsrColIdx_dev = smat_ref_detail::DeviceArray<int>(sparseBlocks).release();
    sparseBcsrVal_dev = smat_ref_detail::DeviceArray<half>(sparseBlocks * bs).release();
    denseBcsrRowPtr_dev = smat_ref_detail::DeviceArray<int>(nb1).release();
    denseBcsrColIdx_dev = smat_ref_detail::DeviceArray<int>(denseBlocks).release();
    denseBcsrVal_dev = smat_ref_detail::DeviceArray<half>(denseBlocks * bs).release();
    smat_ref_detail::check(smat_bcsr_filter(&desc_, nullptr, nullptr, sparseBcsrRowPtr_dev, sparseBcsrColIdx_dev,
                                            reinterpret_cast<uint16_t*>(sparseBcsrVal_dev), denseBcsrRowPtr_dev,
                                            denseBcsrColIdx_dev, reinterpret_cast<uint16_t*>(denseBcsrVal_dev),
                                            nullptr));
    hip_sync();
    sparseBcsrRowPtr_host = smat_ref_detail::to_host(sparseBcsrRowPtr_dev, nb1);
    sparseBcsrColIdx_host = smat_ref_detail::to_host(sparseBcsrColIdx_dev, sparseBlocks);
    sparseBcsrVal_host = smat_ref_detail::to_host(sparseBcsrVal_dev, sparseBlocks * bs);
    denseBcsrRowPtr_host = smat_ref_detail::to_host(denseBcsrRowPtr_dev, nb1);
    denseBcsrColIdx_host = smat_ref_detail::to_host(denseBcsrColIdx_dev, denseBlocks);
    denseBcsrVal_host = smat_ref_detail::to_host(denseBcsrVal_dev, denseBlocks * bs);
    sparseBlockInfo_host = static_cast<int*>(std::malloc((sparseBlocks ? sparseBlocks : 1) * sizeof(int)));
    std::fill(sparseBlockInfo_host, sparseBlockInfo_host + sparseBlocks, 1);
    sparseBlockInfo_dev = smat_ref_detail::to_device(sparseBlockInfo_host, sparseBlocks);
    sparseRelativeBlockIndexMapping_host =
        static_cast<int*>(std::malloc((numberOfBlocks ? numberOfBlocks : 1) * sizeof(int)));
    int counter = 0;
    for (size_t i = 0; i < numberOfBlocks; i++)
      sparseRelativeBlockIndexMapping_host[i] = blockInfo_host[i] == 1 ? counter++ : -1;
    sparseRelativeBlockIndexMapping_dev = smat_ref_detail::to_device(sparseRelativeBlockIndexMapping_host,
                                                                     numberOfBlocks);
  }

  // matrix.h:794-1137: merged 64 x 4k tiles (BLOCK = 4) of the BCSR on the
  // GPU (smat_tiles_from_bcsr), the tile-grid maps and one type per 16 x k
  // sub-block.  The tiles are stored in (tile row, tile column) order with a
  // row pointer per tile row and types computed per sub-block, the layout
  // the reference's consumers index (bcsrBlocking itself stores tiles in
  // order of first appearance); see smat.h.
  void bcsrBlocking() {
    release_tiles();
    smat_ref_detail::check(smat_tiles_from_bcsr(&desc_, nullptr, &tiles_));
    smat_ref_detail::check(smat_tiles_describe(tiles_, &tdesc_));
    mergedNonzeroBlocks = (size_t)tdesc_.ntiles;
    numberOfMergedRegions = (size_t)(tdesc_.ntrows * tdesc_.ntcols);
    mergedBlockInfo_dev = smat_ref_detail::DeviceArray<int>(numberOfMergedRegions).release();
    mergedRelativeBlockIndexMapping_dev = smat_ref_detail::DeviceArray<int>(numberOfMergedRegions).release();
    smat_ref_detail::check(
        smat_tiles_grid_maps(&tdesc_, mergedBlockInfo_dev, mergedRelativeBlockIndexMapping_dev, nullptr));
    hip_sync();
    const size_t tsz = 16 * 16 * (size_t)mMMA_K;
    mergedBcsrRowPtr_host = smat_ref_detail::to_host(tdesc_.rowptr, (size_t)tdesc_.ntrows + 1);
    mergedBcsrColIdx_host = smat_ref_detail::to_host(tdesc_.colidx, mergedNonzeroBlocks);
    mergedBcsrVal_host = reinterpret_cast<half*>(smat_ref_detail::to_host(tdesc_.values, mergedNonzeroBlocks * tsz));
    mergedTileInfo_host = smat_ref_detail::to_host(tdesc_.info, mergedNonzeroBlocks * 16);
    mergedBlockInfo_host = smat_ref_detail::to_host(mergedBlockInfo_dev, numberOfMergedRegions);
    mergedRelativeBlockIndexMapping_host =
        smat_ref_detail::to_host(mergedRelativeBlockIndexMapping_dev, numberOfMergedRegions);
  }

  ~SparseMatrix() {
    release_tiles();
    release_filter();
    release_bcsr();
    std::free(csrRowPtr_host);
    std::free(csrColIdx_host);
    std::free(csrVal_host);
    (void)hipFree(csrRowPtr_dev);
    (void)hipFree(csrColIdx_dev);
    (void)hipFree(csrVal_dev);
  }
  SparseMatrix(const SparseMatrix&) = delete;
  SparseMatrix& operator=(const SparseMatrix&) = delete;

  size_t getRow() { return m_row; }
  size_t getCol() { return m_col; }
  size_t getElemNum() { return m_row * m_col; }
  size_t getNnz() { return nnz; }
  half* getBcsrValues() { return reinterpret_cast<half*>(desc_.values); }
  int* getBcsrRowPtr() { return desc_.rowptr; }
  int* getBcsrColIdx() { return desc_.colidx; }
  half* getBcsrValuesHost() { return bcsrVal_host; }
  int* getBcsrRowPtrHost() { return bcsrRowPtr_host; }
  int* getBcsrColIdxHost() { return bcsrColIdx_host; }
  size_t getNonzeroblocks() { return nonzeroBlocks; }
  size_t getSparseBlocks() { return sparseBlocks; }
  size_t getDenseBlocks() { return denseBlocks; }
  int* getBlockInfo_dev() { return blockInfo_dev; }
  int* getBlockInfo_host() { return blockInfo_host; }
  int* getRelativeBlockIndexMapping_dev() { return relativeBlockIndexMapping_dev; }
  int* getRelativeBlockIndexMapping_host() { return relativeBlockIndexMapping_host; }
  half* getSparseBcsrValues() { return sparseBcsrVal_dev; }
  int* getSparseBcsrRowPtr() { return sparseBcsrRowPtr_dev; }
  int* getSparseBcsrColIdx() { return sparseBcsrColIdx_dev; }
  half* getDenseBcsrValues() { return denseBcsrVal_dev; }
  int* getDenseBcsrRowPtr() { return denseBcsrRowPtr_dev; }
  int* getDenseBcsrColIdx() { return denseBcsrColIdx_dev; }
  int* getSparseBlockInfo_dev() { return sparseBlockInfo_dev; }
  int* getSparseRelativeBlockIndexMapping_dev() { return sparseRelativeBlockIndexMapping_dev; }
  half* getSparseBcsrValuesHost() { return sparseBcsrVal_host; }
  int* getSparseBcsrRowPtrHost() { return sparseBcsrRowPtr_host; }
  int* getSparseBcsrColIdxHost() { return sparseBcsrColIdx_host; }
  int* getSparseBlockInfoHost() { return sparseBlockInfo_host; }
  int* getSparseRelativeBlockIndexMappingHost() { return sparseRelativeBlockIndexMapping_host; }
  half* getDenseBcsrValuesHost() { return denseBcsrVal_host; }
  int* getDenseBcsrRowPtrHost() { return denseBcsrRowPtr_host; }
  int* getDenseBcsrColIdxHost() { return denseBcsrColIdx_host; }
  half* getMergedBcsrValues() { return reinterpret_cast<half*>(tdesc_.values); }
  int* getMergedBcsrRowPtr() { return tdesc_.rowptr; }
  int* getMergedBcsrColIdx() { return tdesc_.colidx; }
  half* getMergedBcsrValuesHost() { return mergedBcsrVal_host; }
  int* getMergedBcsrRowPtrHost() { return mergedBcsrRowPtr_host; }
  int* getMergedBcsrColIdxHost() { return mergedBcsrColIdx_host; }
  int* getMergedRelativeBlockIndexMapping_dev() { return mergedRelativeBlockIndexMapping_dev; }
  int* getMergedRelativeBlockIndexMapping_host() { return mergedRelativeBlockIndexMapping_host; }
  int* getMergedBlockInfo_dev() { return mergedBlockInfo_dev; }
  int* getMergedBlockInfo_host() { return mergedBlockInfo_host; }
  int* getMergedTileInfo_dev() { return tdesc_.info; }
  int* getMergedTileInfo_host() { return mergedTileInfo_host; }
  size_t getMergedNonzeroBlocks() { return mergedNonzeroBlocks; }
  const smat_bcsr_desc& desc() const { return desc_; }  // the native view of the same BCSR
  const smat_tiles_desc& tiles_desc() const { return tdesc_; }  // and of its merged tiles

 private:
  static void hip_sync() { smat_ref_detail::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

  void release_bcsr() {
    if (bcsr_) smat_bcsr_destroy(bcsr_);
    bcsr_ = nullptr;
    desc_ = smat_bcsr_desc{};
    (void)hipFree(blockInfo_dev);
    (void)hipFree(relativeBlockIndexMapping_dev);
    std::free(bcsrRowPtr_host);
    std::free(bcsrColIdx_host);
    std::free(bcsrVal_host);
    std::free(blockInfo_host);
    std::free(relativeBlockIndexMapping_host);
    blockInfo_dev = relativeBlockIndexMapping_dev = nullptr;
    bcsrRowPtr_host = bcsrColIdx_host = blockInfo_host = relativeBlockIndexMapping_host = nullptr;
    bcsrVal_host = nullptr;
  }

  void release_tiles() {
    if (tiles_) smat_tiles_destroy(tiles_);
    tiles_ = nullptr;
    tdesc_ = smat_tiles_desc{};
    (void)hipFree(mergedBlockInfo_dev);
    (void)hipFree(mergedRelativeBlockIndexMapping_dev);
    for (int* p : {mergedBcsrRowPtr_host, mergedBcsrColIdx_host, mergedTileInfo_host, mergedBlockInfo_host,
                   mergedRelativeBlockIndexMapping_host})
      std::free(p);
    std::free(mergedBcsrVal_host);
    mergedBlockInfo_dev = mergedRelativeBlockIndexMapping_dev = nullptr;
    mergedBcsrRowPtr_host = mergedBcsrColIdx_host = mergedTileInfo_host = mergedBlockInfo_host = nullptr;
    mergedRelativeBlockIndexMapping_host = nullptr;
    mergedBcsrVal_host = nullptr;
    mergedNonzeroBlocks = numberOfMergedRegions = 0;
  }

  void release_filter() {
    for (int* p : {sparseBcsrRowPtr_dev, sparseBcsrColIdx_dev, denseBcsrRowPtr_dev, denseBcsrColIdx_dev,
                   sparseBlockInfo_dev, sparseRelativeBlockIndexMapping_dev})
      (void)hipFree(p);
    (void)hipFree(sparseBcsrVal_dev);
    (void)hipFree(denseBcsrVal_dev);
    for (int* p : {sparseBcsrRowPtr_host, sparseBcsrColIdx_host, denseBcsrRowPtr_host, denseBcsrColIdx_host,
                   sparseBlockInfo_host, sparseRelativeBlockIndexMapping_host})
      std::free(p);
    std::free(sparseBcsrVal_host);
    std::free(denseBcsrVal_host);
    sparseBcsrRowPtr_dev = sparseBcsrColIdx_dev = denseBcsrRowPtr_dev = denseBcsrColIdx_dev = nullptr;
    sparseBlockInfo_dev = sparseRelativeBlockIndexMapping_dev = nullptr;
    sparseBcsrVal_dev = denseBcsrVal_dev = nullptr;
    sparseBcsrRowPtr_host = sparseBcsrColIdx_host = denseBcsrRowPtr_host = denseBcsrColIdx_host = nullptr;
    sparseBlockInfo_host = sparseRelativeBlockIndexMapping_host = nullptr;
    sparseBcsrVal_host = denseBcsrVal_host = nullptr;
  }

  std::string m_name, filename;
  int mMMA_K;
  size_t m_row = 0, m_col = 0, nnz = 0;
  size_t numberOfBlocks = 0, nonzeroBlocks = 0, sparseBlocks = 0, denseBlocks = 0;
  int *csrRowPtr_host = nullptr, *csrColIdx_host = nullptr;
  half* csrVal_host = nullptr;
  int *csrRowPtr_dev = nullptr, *csrColIdx_dev = nullptr;
  half* csrVal_dev = nullptr;
  smat_bcsr bcsr_ = nullptr;
  smat_bcsr_desc desc_{};
  int *bcsrRowPtr_host = nullptr, *bcsrColIdx_host = nullptr;
  half* bcsrVal_host = nullptr;
  int *blockInfo_dev = nullptr, *relativeBlockIndexMapping_dev = nullptr;
  int *blockInfo_host = nullptr, *relativeBlockIndexMapping_host = nullptr;
  int *sparseBcsrRowPtr_dev = nullptr, *sparseBcsrColIdx_dev = nullptr, *denseBcsrRowPtr_dev = nullptr,
      *denseBcsrColIdx_dev = nullptr;
  half *sparseBcsrVal_dev = nullptr, *denseBcsrVal_dev = nullptr;
  int *sparseBcsrRowPtr_host = nullptr, *sparseBcsrColIdx_host = nullptr, *denseBcsrRowPtr_host = nullptr,
      *denseBcsrColIdx_host = nullptr;
  half *sparseBcsrVal_host = nullptr, *denseBcsrVal_host = nullptr;
  int *sparseBlockInfo_dev = nullptr, *sparseRelativeBlockIndexMapping_dev = nullptr;
  int *sparseBlockInfo_host = nullptr, *sparseRelativeBlockIndexMapping_host = nullptr;
  smat_tiles tiles_ = nullptr;
  smat_tiles_desc tdesc_{};
  size_t mergedNonzeroBlocks = 0, numberOfMergedRegions = 0;
  int *mergedBlockInfo_dev = nullptr, *mergedRelativeBlockIndexMapping_dev = nullptr;
  int *mergedBcsrRowPtr_host = nullptr, *mergedBcsrColIdx_host = nullptr, *mergedTileInfo_host = nullptr;
  int *mergedBlockInfo_host = nullptr, *mergedRelativeBlockIndexMapping_host = nullptr;
  half* mergedBcsrVal_host = nullptr;
};

// ---- main.cu:10-77: the launchers with the reference's half* parameters -----
inline void mmaOBTKernel(half* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, half* B, half* C, size_t M, size_t N,
                         size_t K, size_t nonzeroBlocks, int* blockInfo, int* relativeBlockIndexMapping) {
  mmaOBTKernel(reinterpret_cast<uint16_t*>(bcsrValuesA), bcsrRowPtrA, bcsrColIdxA, reinterpret_cast<uint16_t*>(B),
               reinterpret_cast<uint16_t*>(C), M, N, K, nonzeroBlocks, blockInfo, relativeBlockIndexMapping);
}

inline void mmaOBTKernel_large(half* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, half* B, half* C, size_t M,
                               size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                               int* relativeBlockIndexMapping) {
  mmaOBTKernel_large(reinterpret_cast<uint16_t*>(bcsrValuesA), bcsrRowPtrA, bcsrColIdxA,
                     reinterpret_cast<uint16_t*>(B), reinterpret_cast<uint16_t*>(C), M, N, K, nonzeroBlocks,
                     blockInfo, relativeBlockIndexMapping);
}

inline void mmaOBTSKernel(half* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata, half* sparseMatrixA,
                          half* B, half* C, size_t M, size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                          int* relativeBlockIndexMapping) {
  mmaOBTSKernel(reinterpret_cast<uint16_t*>(bcsrValuesA), bcsrRowPtrA, bcsrColIdxA, metadata,
                reinterpret_cast<uint16_t*>(sparseMatrixA), reinterpret_cast<uint16_t*>(B),
                reinterpret_cast<uint16_t*>(C), M, N, K, nonzeroBlocks, blockInfo, relativeBlockIndexMapping);
}

inline void mmaOBTSKernel_large(half* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                                half* sparseMatrixA, half* B, half* C, size_t M, size_t N, size_t K,
                                size_t nonzeroBlocks, int* blockInfo, int* relativeBlockIndexMapping) {
  mmaOBTSKernel_large(reinterpret_cast<uint16_t*>(bcsrValuesA), bcsrRowPtrA, bcsrColIdxA, metadata,
                      reinterpret_cast<uint16_t*>(sparseMatrixA), reinterpret_cast<uint16_t*>(B),
                      reinterpret_cast<uint16_t*>(C), M, N, K, nonzeroBlocks, blockInfo, relativeBlockIndexMapping);
}

inline void preprocessing_mmaSTKernel(half* bcsrValuesA, char* metadata, half* sparseMatrixA, size_t M, size_t N,
                                      size_t K, size_t nonzeroBlocks, int* blockInfo,
                                      int* relativeBlockIndexMapping) {
  preprocessing_mmaSTKernel(reinterpret_cast<uint16_t*>(bcsrValuesA), metadata,
                            reinterpret_cast<uint16_t*>(sparseMatrixA), M, N, K, nonzeroBlocks, blockInfo,
                            relativeBlockIndexMapping);
}

inline void preprocessing_mmaSTKernel_large(half* bcsrValuesA, char* metadata, half* sparseMatrixA, size_t M,
                                            size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                                            int* relativeBlockIndexMapping) {
  preprocessing_mmaSTKernel_large(reinterpret_cast<uint16_t*>(bcsrValuesA), metadata,
                                  reinterpret_cast<uint16_t*>(sparseMatrixA), M, N, K, nonzeroBlocks, blockInfo,
                                  relativeBlockIndexMapping);
}

inline void mmaOBTKernel_tiled(half* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, half* B, half* C, size_t M,
                               size_t N, size_t K, size_t nonzeroBlocks, int* blockInfo,
                               int* relativeBlockIndexMapping) {
  mmaOBTKernel_tiled(reinterpret_cast<uint16_t*>(bcsrValuesA), bcsrRowPtrA, bcsrColIdxA,
                     reinterpret_cast<uint16_t*>(B), reinterpret_cast<uint16_t*>(C), M, N, K, nonzeroBlocks,
                     blockInfo, relativeBlockIndexMapping);
}

inline void preprocessing_mmaOBTSKernel_tiled_large(half* bcsrValuesA, char* metadata, half* sparseMatrixA,
                                                    size_t M, size_t N, size_t K, size_t nonzeroBlocks,
                                                    int* blockInfo, int* relativeBlockIndexMapping,
                                                    int* tileInfo) {
  preprocessing_mmaOBTSKernel_tiled_large(reinterpret_cast<uint16_t*>(bcsrValuesA), metadata,
                                          reinterpret_cast<uint16_t*>(sparseMatrixA), M, N, K, nonzeroBlocks,
                                          blockInfo, relativeBlockIndexMapping, tileInfo);
}

inline void mmaOBTSKernel_tiled_large(half* bcsrValuesA, int* bcsrRowPtrA, int* bcsrColIdxA, char* metadata,
                                      half* sparseMatrixA, half* B, half* C, size_t M, size_t N, size_t K,
                                      size_t nonzeroBlocks, int* blockInfo, int* relativeBlockIndexMapping,
                                      int* tileInfo) {
  mmaOBTSKernel_tiled_large(reinterpret_cast<uint16_t*>(bcsrValuesA), bcsrRowPtrA, bcsrColIdxA, metadata,
                            reinterpret_cast<uint16_t*>(sparseMatrixA), reinterpret_cast<uint16_t*>(B),
                            reinterpret_cast<uint16_t*>(C), M, N, K, nonzeroBlocks, blockInfo,
                            relativeBlockIndexMapping, tileInfo);
}
