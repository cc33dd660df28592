// C entry points around the reference's own serial preprocessing code
// (preprocess/src/general/*.cpp), compiled from the reference sources where
// they lie by oracle/Makefile into oracle/_ref/libsmat_ref.so.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle (golden vectors) and as
// the reference CPU baseline (Keeper reordering + CSR::multiply).  Nothing in
// the product links it.
#include "blocking.h"
#include "matrices.h"
#include "utilities.h"

#include <chrono>
#include <cstring>
#include <fstream>
#include <vector>

// Defined in blocking.cpp:431 but not declared in blocking.h.
std::vector<intT> IterativeBlockingKeeper(
    const CSR& cmat, float tau, distFuncGroup distanceFunction,
    intT col_block_size, intT max_row_block_size, bool use_size,
    bool use_pattern, intT& comparison_counter, intT& merge_counter,
    float& average_row_distance, float& average_merge_tau, float& timer_total,
    float& timer_comparisons, float& timer_merges);

namespace {

// Build a reference CSR (preprocess/include/matrices.h:10-86) from arrays.
// The struct has only stream constructors, so start from an empty stream
// (csr.cpp:166-290 then yields rows = 0) and fill the public fields.
struct RefCSR {
  std::ifstream empty;
  CSR* m;
  RefCSR(intT rows, intT cols, const intT* rowptr, const intT* colidx,
         const float* val)
      : m(nullptr) {
    m = new CSR(empty, " ", val == nullptr, el);
    delete[] m->nzcount;
    delete[] m->ja;
    m->rows = rows;
    m->cols = cols;
    m->pattern_only = (val == nullptr);
    m->nzcount = new intT[rows];
    m->ja = new intT*[rows];
    m->ma = val ? new DataT*[rows] : nullptr;
    for (intT i = 0; i < rows; i++) {
      intT n = rowptr[i + 1] - rowptr[i];
      m->nzcount[i] = n;
      m->ja[i] = new intT[n];
      std::memcpy(m->ja[i], colidx + rowptr[i], sizeof(intT) * n);
      if (val) {
        m->ma[i] = new DataT[n];
        std::memcpy(m->ma[i], val + rowptr[i], sizeof(float) * n);
      }
    }
  }
  ~RefCSR() {
    if (m->ma) {
      for (intT i = 0; i < m->rows; i++) delete[] m->ma[i];
      delete[] m->ma;
      m->ma = nullptr;
      m->pattern_only = true;
    }
    delete m;
  }
};

}  // namespace

extern "C" {

float ref_jaccard_distance_group(const intT* a, intT na, intT ga, const intT* b,
                                 intT nb, intT gb, intT block) {
  std::vector<intT> va(a, a + na);
  return JaccardDistanceGroup(va, ga, const_cast<intT*>(b), nb, gb, block);
}

intT ref_merge_rows(const intT* a, intT na, const intT* b, intT nb, intT* out) {
  std::vector<intT> va(a, a + na);
  std::vector<intT> r = merge_rows(va, const_cast<intT*>(b), nb);
  std::copy(r.begin(), r.end(), out);
  return (intT)r.size();
}

// IterativeBlockingKeeper (blocking.cpp:431) on an in-memory pattern CSR.
// counters = [comparison_counter, merge_counter]; returns elapsed seconds.
double ref_keeper(intT rows, intT cols, const intT* rowptr, const intT* colidx,
                  float tau, intT col_block, intT max_rows, intT* grouping,
                  intT* counters) {
  RefCSR c(rows, cols, rowptr, colidx, nullptr);
  intT cmp = 0, mrg = 0;
  float ard = 0, amt = 0, tt = 0, tc = 0, tm = 0;
  auto t0 = std::chrono::steady_clock::now();
  std::vector<intT> g = IterativeBlockingKeeper(*c.m, tau, JaccardDistanceGroup,
                                                col_block, max_rows, false, true,
                                                cmp, mrg, ard, amt, tt, tc, tm);
  auto t1 = std::chrono::steady_clock::now();
  std::copy(g.begin(), g.end(), grouping);
  counters[0] = cmp;
  counters[1] = mrg;
  return std::chrono::duration<double>(t1 - t0).count();
}

// IterativeBlockingPattern, algorithm 0 (blocking.cpp:87).
void ref_algo0(intT rows, intT cols, const intT* rowptr, const intT* colidx,
               float tau, intT block, int use_size, intT* grouping,
               intT* counters) {
  RefCSR c(rows, cols, rowptr, colidx, nullptr);
  intT cmp = 0, mrg = 0;
  float tt = 0;
  std::vector<intT> g = IterativeBlockingPattern(*c.m, tau, JaccardDistanceGroup,
                                                 block, use_size != 0, true, cmp,
                                                 mrg, tt);
  std::copy(g.begin(), g.end(), grouping);
  counters[0] = cmp;
  counters[1] = mrg;
}

void ref_get_permutation(intT n, const intT* grouping, intT* perm) {
  std::vector<intT> g(grouping, grouping + n);
  std::vector<intT> p = get_permutation(g);
  std::copy(p.begin(), p.end(), perm);
}

void ref_fixed_size_grouping(intT n, const intT* grouping, intT row_block,
                             intT* out) {
  std::vector<intT> g(grouping, grouping + n);
  std::vector<intT> r = get_fixed_size_grouping(g, row_block);
  std::copy(r.begin(), r.end(), out);
}

// BlockingEngine::CollectBlockingInfo (blocking.cpp:562).
void ref_collect_blocking_info(intT rows, intT cols, const intT* rowptr,
                               const intT* colidx, const intT* grouping,
                               intT col_block, intT* out, float* avg_height) {
  RefCSR c(rows, cols, rowptr, colidx, nullptr);
  BlockingEngine e;
  e.col_block_size = col_block;
  e.grouping_result.assign(grouping, grouping + rows);
  e.CollectBlockingInfo(*c.m);
  out[0] = e.VBR_nzcount;
  out[1] = e.VBR_nzblocks_count;
  out[2] = e.VBR_longest_row;
  *avg_height = e.VBR_average_height;
}

// CSR::multiply (csr.cpp:47): C (column-major, rows x B_cols) += A * B.
// Returns elapsed seconds of the multiply alone.
double ref_csr_multiply(intT rows, intT cols, const intT* rowptr,
                        const intT* colidx, const float* val, const float* B,
                        intT B_cols, float* C) {
  RefCSR c(rows, cols, rowptr, colidx, val);
  auto t0 = std::chrono::steady_clock::now();
  c.m->multiply(const_cast<float*>(B), B_cols, C);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// CSR::reorder (csr.cpp:97): permuted row order of a pattern CSR.
void ref_reorder_rows(intT rows, intT cols, const intT* rowptr,
                      const intT* colidx, const intT* grouping,
                      intT* out_rowptr, intT* out_colidx) {
  RefCSR c(rows, cols, rowptr, colidx, nullptr);
  std::vector<intT> g(grouping, grouping + rows);
  c.m->reorder(g);
  intT nz = 0;
  out_rowptr[0] = 0;
  for (intT i = 0; i < rows; i++) {
    for (intT k = 0; k < c.m->nzcount[i]; k++) out_colidx[nz++] = c.m->ja[i][k];
    out_rowptr[i + 1] = nz;
  }
}

// CSR::read_from_edgelist (csr.cpp:166-290) on a file, edge-list mode.
// Returns nnz (rowptr/colidx/vals written when they fit in cap), or -1 when
// the reference rejects the file.
intT ref_read_edgelist(const char* path, int pattern_only, intT* rows, intT* cols, intT* rowptr, intT* colidx,
                       float* vals, intT cap_rows, intT cap_nnz) {
  try {
    std::ifstream in(path);
    CSR m(in, " ", pattern_only != 0, el);
    *rows = m.rows;
    *cols = m.cols;
    intT nnz = 0;
    for (intT i = 0; i < m.rows; i++) nnz += m.nzcount[i];
    if (m.rows + 1 <= cap_rows && nnz <= cap_nnz) {
      rowptr[0] = 0;
      intT at = 0;
      for (intT i = 0; i < m.rows; i++) {
        for (intT q = 0; q < m.nzcount[i]; q++, at++) {
          colidx[at] = m.ja[i][q];
          vals[at] = pattern_only ? 1.f : m.ma[i][q];
        }
        rowptr[i + 1] = at;
      }
    }
    return nnz;
  } catch (const std::exception&) {
    return -1;
  }
}

// CSR::read_from_edgelist (csr.cpp:166-290) with its format and symmetrize
// options.  Returns nnz as ref_read_edgelist does, -1 when it throws.
intT ref_read_edgelist_opts(const char* path, int pattern_only, int mat_fmt, int symmetrize, intT* rows, intT* cols,
                            intT* rowptr, intT* colidx, float* vals, intT cap_rows, intT cap_nnz) {
  try {
    std::ifstream empty;
    CSR m(empty, " ", pattern_only != 0, el);  // rows = 0: nothing to free before the re-read
    std::ifstream in(path);
    m.read_from_edgelist(in, " ", pattern_only != 0, mat_fmt ? mtx : el, symmetrize != 0);
    *rows = m.rows;
    *cols = m.cols;
    intT nnz = 0;
    for (intT i = 0; i < m.rows; i++) nnz += m.nzcount[i];
    if (m.rows + 1 <= cap_rows && nnz <= cap_nnz) {
      rowptr[0] = 0;
      intT at = 0;
      for (intT i = 0; i < m.rows; i++) {
        for (intT q = 0; q < m.nzcount[i]; q++, at++) {
          colidx[at] = m.ja[i][q];
          vals[at] = pattern_only ? 1.f : m.ma[i][q];
        }
        rowptr[i + 1] = at;
      }
    }
    return nnz;
  } catch (const std::exception&) {
    return -1;
  }
}

}  // extern "C"
