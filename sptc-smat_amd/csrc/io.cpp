// Host readers: Matrix Market (the reference's mmio_allinone) and edge lists
// (the reference's CSR::read_from_edgelist), into host CSR arrays.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "smat_internal.h"

namespace {

// double -> IEEE binary16, round to nearest even in one step (the reference
// assigns the MTX double to `half`, mmio_highlevel.h:608, a single rounding).
uint16_t double_to_half(double x) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  const uint16_t sign = (uint16_t)((b >> 48) & 0x8000);
  const int e = (int)((b >> 52) & 0x7ff);
  const uint64_t m = b & ((1ull << 52) - 1);
  if (e == 0x7ff) return (uint16_t)(sign | 0x7c00 | (m ? 0x200 : 0));
  if (e == 0) return sign;  // double subnormals are far below the half range
  const int exp = e - 1023;
  if (exp > 15) return (uint16_t)(sign | 0x7c00);
  const uint64_t mant = m | (1ull << 52);
  int shift;
  uint32_t hexp;
  if (exp >= -14) {  // half normal
    shift = 42;
    hexp = (uint32_t)(exp + 15);
  } else {  // half subnormal: units of 2^-24
    shift = 42 + (-14 - exp);
    hexp = 0;
    if (shift >= 54) return sign;  // below half of the smallest subnormal
  }
  uint64_t q = mant >> shift;
  const uint64_t rem = mant & ((1ull << shift) - 1), half = 1ull << (shift - 1);
  if (rem > half || (rem == half && (q & 1))) q++;
  const uint32_t r = hexp ? (hexp << 10) + (uint32_t)(q - 1024) : (uint32_t)q;  // carries roll into the exponent
  if (r >= 0x7c00) return (uint16_t)(sign | 0x7c00);
  return (uint16_t)(sign | r);
}

smat_status alloc_out(smat_csr_host* out, int64_t rows, int64_t nnz) {
  out->rowptr = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (size_t)(rows + 1)));
  out->colidx = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1)));
  out->values_f16 = static_cast<uint16_t*>(std::malloc(sizeof(uint16_t) * (size_t)(nnz > 0 ? nnz : 1)));
  out->values_f32 = static_cast<float*>(std::malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1)));
  if (!out->rowptr || !out->colidx || !out->values_f16 || !out->values_f32) {
    smat_csr_host_free(out);
    return smat::set_error(SMAT_ERR_ALLOC, "host allocation failed (%lld rows, %lld entries)", (long long)rows,
                           (long long)nnz);
  }
  return SMAT_OK;
}

bool lower_eq(const std::string& a, const char* b) {
  if (a.size() != std::strlen(b)) return false;
  for (size_t i = 0; i < a.size(); i++)
    if (std::tolower((unsigned char)a[i]) != b[i]) return false;
  return true;
}

}  // namespace

extern "C" {

void smat_csr_host_free(smat_csr_host* c) {
  if (!c) return;
  std::free(c->rowptr);
  std::free(c->colidx);
  std::free(c->values_f16);
  std::free(c->values_f32);
  c->rowptr = nullptr;
  c->colidx = nullptr;
  c->values_f16 = nullptr;
  c->values_f32 = nullptr;
}

// mmio_allinone (src/cuda_hgemm/src/common/mmio_highlevel.h:523-680):
// coordinate format; pattern entries are 1, complex entries keep the real
// part; rows padded to 16 and columns to k (padSparseMatrix, :514-520);
// symmetric / hermitian files mirror every off-diagonal entry; each row keeps
// file order (mirrored entries interleaved in the order of the entries they
// come from); duplicates are kept.
smat_status smat_read_mtx(const char* path, int32_t k, smat_csr_host* out) {
  SMAT_REQUIRE(path && out && k > 0, "smat_read_mtx: bad argument");
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "r");
  SMAT_REQUIRE(f, "smat_read_mtx: cannot open %s: %s", path, std::strerror(errno));
  char line[4096];
  if (!std::fgets(line, sizeof(line), f)) {
    std::fclose(f);
    return smat::set_error(SMAT_ERR_INVALID, "smat_read_mtx: empty file");
  }
  char banner[64], object[64], format[64], field[64], symmetry[64];
  if (std::sscanf(line, "%63s %63s %63s %63s %63s", banner, object, format, field, symmetry) != 5 ||
      std::string(banner) != "%%MatrixMarket" || !lower_eq(object, "matrix") || !lower_eq(format, "coordinate")) {
    std::fclose(f);
    return smat::set_error(SMAT_ERR_INVALID, "smat_read_mtx: not a Matrix Market coordinate file");
  }
  const bool pattern = lower_eq(field, "pattern"), cplx = lower_eq(field, "complex"),
             integer = lower_eq(field, "integer"), real = lower_eq(field, "real") || lower_eq(field, "double");
  const bool sym = lower_eq(symmetry, "symmetric") || lower_eq(symmetry, "hermitian");
  if (!(pattern || cplx || integer || real)) {
    std::fclose(f);
    return smat::set_error(SMAT_ERR_INVALID, "smat_read_mtx: unknown field %s", field);
  }
  long long m = 0, n = 0, reported = 0;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '%') continue;
    if (std::sscanf(line, "%lld %lld %lld", &m, &n, &reported) == 3) break;
    std::fclose(f);
    return smat::set_error(SMAT_ERR_INVALID, "smat_read_mtx: bad size line");
  }
  if (m <= 0 || n <= 0 || reported < 0 || m >= INT32_MAX || n >= INT32_MAX) {
    std::fclose(f);
    return smat::set_error(SMAT_ERR_INVALID, "smat_read_mtx: bad sizes %lld x %lld, %lld entries", m, n, reported);
  }
  if (sym && m != n) {
    std::fclose(f);
    return smat::set_error(SMAT_ERR_INVALID, "smat_read_mtx: a symmetric file must be square (%lld x %lld)", m, n);
  }
  const int64_t rows = (m % 16) ? m + 16 - m % 16 : m;
  const int64_t cols = (n % k) ? n + k - n % k : n;
  std::vector<int32_t> ri((size_t)reported), ci((size_t)reported);
  std::vector<double> val((size_t)reported);
  for (long long q = 0; q < reported; q++) {
    long long i = 0, j = 0;
    double v = 1.0, im = 0.0;
    int got;
    if (pattern) got = std::fscanf(f, "%lld %lld", &i, &j) - 2;
    else if (cplx) got = std::fscanf(f, "%lld %lld %lg %lg", &i, &j, &v, &im) - 4;
    else if (integer) {
      long long iv = 0;
      got = std::fscanf(f, "%lld %lld %lld", &i, &j, &iv) - 3;
      v = (double)iv;
    } else got = std::fscanf(f, "%lld %lld %lg", &i, &j, &v) - 3;
    if (got != 0 || i < 1 || j < 1 || i > m || j > n) {
      std::fclose(f);
      return smat::set_error(SMAT_ERR_INVALID, "smat_read_mtx: bad entry %lld", q + 1);
    }
    ri[(size_t)q] = (int32_t)(i - 1);
    ci[(size_t)q] = (int32_t)(j - 1);
    val[(size_t)q] = v;
  }
  std::fclose(f);
  std::vector<int64_t> cnt((size_t)rows + 1, 0);
  for (long long q = 0; q < reported; q++) {
    cnt[(size_t)ri[(size_t)q]]++;
    if (sym && ri[(size_t)q] != ci[(size_t)q]) cnt[(size_t)ci[(size_t)q]]++;
  }
  int64_t nnz = 0;
  for (int64_t r = 0; r < rows; r++) nnz += cnt[(size_t)r];
  SMAT_REQUIRE(nnz < INT32_MAX, "smat_read_mtx: %lld entries exceed int32 indices", (long long)nnz);
  SMAT_TRY(alloc_out(out, rows, nnz));
  out->rows = rows;
  out->cols = cols;
  out->nnz = nnz;
  out->symmetric = sym;
  out->rowptr[0] = 0;
  for (int64_t r = 0; r < rows; r++) out->rowptr[r + 1] = out->rowptr[r] + (int32_t)cnt[(size_t)r];
  std::vector<int32_t> fill((size_t)rows, 0);
  auto put = [&](int32_t r, int32_t c, double v) {
    const int64_t at = out->rowptr[r] + fill[(size_t)r]++;
    out->colidx[at] = c;
    out->values_f16[at] = double_to_half(v);
    out->values_f32[at] = (float)v;
  };
  for (long long q = 0; q < reported; q++) {
    const int32_t r = ri[(size_t)q], c = ci[(size_t)q];
    put(r, c, val[(size_t)q]);
    if (sym && r != c) put(c, r, val[(size_t)q]);
  }
  return SMAT_OK;
}

}  // extern "C"

namespace {

// CSR::read_from_edgelist (preprocess/src/general/csr.cpp:166-290):
// leading '#' / '%' lines skipped, then (mat_fmt SMAT_FMT_MTX) one more line,
// the size line; each line "a<d>b[<d>value]" parsed like stoi / stof (leading
// integer or float prefix of each field); in MTX mode a and b are swapped
// (the row is the second field, as the reference does, indices unchanged);
// rows must not descend; rows = last row + 1, cols = largest column + 1;
// columns in file order.  A weighted line with two fields takes its column
// field as the value too (the reference's parser reads the remainder again).
// symmetrize mirrors every entry (i, j) into row j when no entry lies below
// the diagonal (inserted at std::lower_bound of row j, as the reference),
// pattern-only input only; an entry whose column has no row is refused (the
// reference indexes past its row list there).
smat_status read_edgelist_stream(std::istream& in, const char* delimiter, int32_t pattern_only, int32_t mat_fmt,
                                 int32_t symmetrize, smat_csr_host* out) {
  SMAT_REQUIRE(mat_fmt == SMAT_FMT_EL || mat_fmt == SMAT_FMT_MTX, "smat_read_edgelist: unknown format %d", mat_fmt);
  const std::string d = delimiter && *delimiter ? delimiter : " ";
  while (in.peek() == '#' || in.peek() == '%') in.ignore(2048, '\n');
  if (mat_fmt == SMAT_FMT_MTX) in.ignore(2048, '\n');
  std::vector<std::vector<int32_t>> pos;
  std::vector<std::vector<float>> val;
  int64_t max_col = 0;
  bool triangular = true;
  std::string line;
  auto to_int = [](const std::string& s, int64_t& v) {
    char* end = nullptr;
    errno = 0;
    const long long x = std::strtoll(s.c_str(), &end, 10);
    if (end == s.c_str() || errno) return false;
    v = x;
    return true;
  };
  auto to_float = [](const std::string& s, float& v) {
    char* end = nullptr;
    errno = 0;
    const float x = std::strtof(s.c_str(), &end);
    if (end == s.c_str()) return false;
    v = x;
    return true;
  };
  int64_t lineno = 0;
  while (std::getline(in, line)) {
    lineno++;
    size_t p = line.find(d);
    const std::string a = line.substr(0, p);
    std::string rest = p == std::string::npos ? line : line.substr(p + d.size());
    p = rest.find(d);
    const std::string b = rest.substr(0, p);
    int64_t r = 0, c = 0;
    if (!to_int(a, r) || !to_int(b, c))
      return smat::set_error(SMAT_ERR_INVALID, "smat_read_edgelist: bad line %lld", (long long)lineno);
    if (mat_fmt == SMAT_FMT_MTX) std::swap(r, c);
    if (r < 0 || c < 0 || c >= INT32_MAX || r >= INT32_MAX)
      return smat::set_error(SMAT_ERR_INVALID, "smat_read_edgelist: bad line %lld", (long long)lineno);
    if (c < r) triangular = false;
    if (c > max_col) max_col = c;
    float v = 1.f;
    if (!pattern_only) {
      const std::string third = p == std::string::npos ? rest : rest.substr(p + d.size());
      const std::string field = third.substr(0, third.find(d));
      if (!to_float(field, v))
        return smat::set_error(SMAT_ERR_INVALID, "smat_read_edgelist: bad value on line %lld", (long long)lineno);
    }
    if (r < (int64_t)pos.size() - 1)
      return smat::set_error(SMAT_ERR_INVALID, "smat_read_edgelist: rows must be in ascending order (line %lld)",
                             (long long)lineno);
    while ((int64_t)pos.size() <= r) {
      pos.emplace_back();
      val.emplace_back();
    }
    pos[(size_t)r].push_back((int32_t)c);
    val[(size_t)r].push_back(v);
  }
  if (symmetrize && triangular) {
    for (size_t i = 0; i < pos.size(); i++) {
      for (size_t nz = 0; nz < pos[i].size(); nz++) {
        const int32_t j = pos[i][nz];
        if ((size_t)j >= pos.size())
          return smat::set_error(SMAT_ERR_INVALID,
                                 "smat_read_edgelist: symmetrize: column %d has no row (the matrix has %zu rows)", j,
                                 pos.size());
        auto& row = pos[(size_t)j];
        auto it = std::lower_bound(row.begin(), row.end(), (int32_t)i);
        if (it == row.end() || *it != (int32_t)i) {
          SMAT_REQUIRE(pattern_only, "smat_read_edgelist: symmetrize is only implemented for pattern-only input");
          val[(size_t)j].insert(val[(size_t)j].begin() + (it - row.begin()), 1.f);
          row.insert(it, (int32_t)i);
        }
      }
    }
  }
  const int64_t rows = (int64_t)pos.size();
  int64_t nnz = 0;
  for (const auto& r : pos) nnz += (int64_t)r.size();
  SMAT_REQUIRE(nnz < INT32_MAX && rows < INT32_MAX, "smat_read_edgelist: too large");
  SMAT_TRY(alloc_out(out, rows, nnz));
  out->rows = rows;
  out->cols = max_col + 1;  // the reference starts its maximum at 0
  out->nnz = nnz;
  out->rowptr[0] = 0;
  int64_t q = 0;
  for (int64_t r = 0; r < rows; r++) {
    for (size_t k = 0; k < pos[(size_t)r].size(); k++, q++) {
      out->colidx[q] = pos[(size_t)r][k];
      out->values_f32[q] = val[(size_t)r][k];
      out->values_f16[q] = double_to_half((double)val[(size_t)r][k]);
    }
    out->rowptr[r + 1] = (int32_t)q;
  }
  return SMAT_OK;
}

}  // namespace

extern "C" {

smat_status smat_read_edgelist_ex(const char* path, const char* delimiter, int32_t pattern_only, int32_t mat_fmt,
                                  int32_t symmetrize, smat_csr_host* out) {
  SMAT_REQUIRE(path && out, "smat_read_edgelist: bad argument");
  std::memset(out, 0, sizeof(*out));
  std::ifstream in(path);
  SMAT_REQUIRE(in.good(), "smat_read_edgelist: cannot open %s", path);
  return read_edgelist_stream(in, delimiter, pattern_only, mat_fmt, symmetrize, out);
}

smat_status smat_read_edgelist(const char* path, const char* delimiter, int32_t pattern_only, smat_csr_host* out) {
  return smat_read_edgelist_ex(path, delimiter, pattern_only, SMAT_FMT_EL, 0, out);
}

smat_status smat_parse_edgelist(const char* text, int64_t len, const char* delimiter, int32_t pattern_only,
                                int32_t mat_fmt, int32_t symmetrize, smat_csr_host* out) {
  SMAT_REQUIRE(out && (text || len == 0) && len >= 0, "smat_parse_edgelist: bad argument");
  std::memset(out, 0, sizeof(*out));
  std::istringstream in(std::string(text ? text : "", (size_t)len));
  return read_edgelist_stream(in, delimiter, pattern_only, mat_fmt, symmetrize, out);
}

}  // extern "C"
