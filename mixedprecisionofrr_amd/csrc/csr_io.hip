// Matrix input: the binary CSR files of the reference's MatrixReader
// (CPP/io/matrix_reader.hpp:7-45, matrix_reader.cpp:8-85).
//   <prefix>_row_ptr.bin   int32, rows + 1 entries
//   <prefix>_col_ind.bin   int32, nnz entries (0-based)
//   <prefix>_values.bin    float64, nnz entries
//   <prefix>_singular_values.bin  float64 (read_singular_values)
// rows = entries of row_ptr - 1, cols = max column index + 1 (matrix_reader.cpp:
// 23-39).  Host code only: the caller uploads the arrays (int64 row pointers on
// the device, the layout every CSR kernel here reads).  Unlike the reference,
// the structure is validated (row pointers start at 0, never decrease and end
// at the number of stored entries; column indices are non-negative; no file
// ends in a partial element), so a damaged file is an error rather than an
// out-of-bounds gather later.  cols may differ from rows: a rectangular matrix
// reaches only the entry points that take both dimensions.
#include <cstdio>
#include <string>
#include <vector>

#include "internal.hpp"

namespace ofrr {
namespace {

thread_local std::string g_io_error;

int io_fail(int status, const std::string &msg) {
   g_io_error = msg;
   return status;
}

// whole file as elements of T; false when it cannot be opened, is empty or
// ends in a partial element (a truncated file)
template <typename T>
bool read_all(const std::string &path, std::vector<T> &out) {
   out.clear();
   FILE *f = std::fopen(path.c_str(), "rb");
   if (!f) return false;
   std::fseek(f, 0, SEEK_END);
   const long bytes = std::ftell(f);
   std::fseek(f, 0, SEEK_SET);
   if (bytes <= 0 || (size_t)bytes % sizeof(T) != 0) {
      std::fclose(f);
      return false;
   }
   out.resize((size_t)bytes / sizeof(T));
   const size_t got = std::fread(out.data(), sizeof(T), out.size(), f);
   std::fclose(f);
   if (got != out.size()) {
      out.clear();
      return false;
   }
   return !out.empty();
}

struct CsrFiles {
   std::vector<int32_t> rp, ci;
   std::vector<double> va;
   int64_t rows = 0, cols = 0;
};

int load(const char *prefix, CsrFiles &m) {
   if (!prefix) return io_fail(kErrorInvalidArgument, "csr_read: null prefix");
   const std::string p(prefix);
   const std::string frp = p + "_row_ptr.bin", fci = p + "_col_ind.bin", fva = p + "_values.bin";
   // the reference's messages (matrix_reader.cpp:15-33)
   if (!read_all(frp, m.rp)) return io_fail(kErrorInvalidArgument, "Failed to read row pointers from " + frp);
   if (!read_all(fci, m.ci)) return io_fail(kErrorInvalidArgument, "Failed to read column indices from " + fci);
   if (!read_all(fva, m.va)) return io_fail(kErrorInvalidArgument, "Failed to read values from " + fva);
   m.rows = (int64_t)m.rp.size() - 1;
   const int64_t nnz = (int64_t)m.ci.size();
   if (m.rows < 1 || m.rp[0] != 0 || (int64_t)m.rp[m.rows] != nnz)
      return io_fail(kErrorInvalidArgument, "csr_read: row pointers of " + frp + " do not span the column indices");
   if ((int64_t)m.va.size() != nnz)
      return io_fail(kErrorInvalidArgument, "csr_read: " + fva + " and " + fci + " differ in length");
   for (int64_t i = 0; i < m.rows; ++i)
      if (m.rp[i + 1] < m.rp[i])
         return io_fail(kErrorInvalidArgument, "csr_read: row pointers of " + frp + " decrease");
   int64_t cmax = -1;
   for (int32_t c : m.ci) {
      if (c < 0) return io_fail(kErrorInvalidArgument, "csr_read: negative column index in " + fci);
      cmax = std::max<int64_t>(cmax, c);
   }
   m.cols = cmax + 1;
   return kSuccess;
}

}  // namespace
}  // namespace ofrr

using namespace ofrr;

extern "C" {

const char *ofrr_io_last_error(void) { return g_io_error.c_str(); }

int ofrr_csr_file_info(const char *prefix, int64_t *rows, int64_t *cols, int64_t *nnz) {
   CsrFiles m;
   const int st = load(prefix, m);
   if (st) return st;
   if (rows) *rows = m.rows;
   if (cols) *cols = m.cols;
   if (nnz) *nnz = (int64_t)m.ci.size();
   return kSuccess;
}

int ofrr_csr_file_read(const char *prefix, int64_t *row_ptr, int32_t *col_idx, double *vals) {
   if (!row_ptr || !col_idx || !vals) return io_fail(kErrorInvalidArgument, "csr_read: null output");
   CsrFiles m;
   const int st = load(prefix, m);
   if (st) return st;
   for (int64_t i = 0; i <= m.rows; ++i) row_ptr[i] = m.rp[i];
   std::copy(m.ci.begin(), m.ci.end(), col_idx);
   std::copy(m.va.begin(), m.va.end(), vals);
   return kSuccess;
}

int ofrr_singular_values_file(const char *prefix, double *out, int64_t cap, int64_t *count) {
   if (!prefix || !count) return io_fail(kErrorInvalidArgument, "singular_values: null argument");
   if (cap < 0) return io_fail(kErrorInvalidArgument, "singular_values: negative capacity");
   const std::string f = std::string(prefix) + "_singular_values.bin";
   std::vector<double> v;
   if (!read_all(f, v)) return io_fail(kErrorInvalidArgument, "Could not open file " + f);
   *count = (int64_t)v.size();
   if (out) std::copy(v.begin(), v.begin() + std::min<int64_t>(cap, (int64_t)v.size()), out);
   return kSuccess;
}

}  // extern "C"
