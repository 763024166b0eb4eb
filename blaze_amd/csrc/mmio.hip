// Matrix Market reader (host code): the on-disk format for real-world CSR inputs.  Reads the
// "coordinate" form -- fields real / double / integer / pattern, symmetry general / symmetric /
// skew-symmetric -- into the CSR that bzg_csr_plan_create and bzg_*csrmv take: rows in order,
// columns ascending within each row, duplicate entries summed in file order, symmetric and
// skew-symmetric files expanded to both triangles.
#include "common.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};

// Next line that is neither a comment nor blank; false at the end of the file.
bool next_data_line(FILE* f, std::string& line) {
  char buf[4096];
  while (std::fgets(buf, sizeof buf, f)) {
    line = buf;
    // a line longer than the buffer: read its remainder
    while (!line.empty() && line.back() != '\n' && std::fgets(buf, sizeof buf, f)) line += buf;
    size_t p = line.find_first_not_of(" \t\r\n");
    if (p == std::string::npos || line[p] == '%') continue;
    return true;
  }
  return false;
}

}  // namespace

namespace {
bzg_status mm_read_csr(const char* path, bzg_csr_host* out);
}  // namespace

extern "C" {

bzg_status bzg_mm_read_csr(const char* path, bzg_csr_host* out) {
  // no C++ exception leaves the C ABI: allocation failures become a status
  try {
    return mm_read_csr(path, out);
  } catch (const std::bad_alloc&) {
    bzg::set_error("mm_read_csr: out of host memory");
    return BZG_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    bzg::set_error(std::string("mm_read_csr: ") + e.what());
    return BZG_ERR_INVALID_ARGUMENT;
  }
}

}  // extern "C"

namespace {
bzg_status mm_read_csr(const char* path, bzg_csr_host* out) {
  BZG_REQUIRE(path && out, BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: null argument");
  *out = bzg_csr_host{};
  FileCloser fc{std::fopen(path, "r")};
  BZG_REQUIRE(fc.f != nullptr, BZG_ERR_INVALID_ARGUMENT, (std::string("mm_read_csr: cannot open ") + path).c_str());
  char hbuf[1024];
  BZG_REQUIRE(std::fgets(hbuf, sizeof hbuf, fc.f) != nullptr, BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: empty file");
  char banner[64] = {0}, object[64] = {0}, format[64] = {0}, field[64] = {0}, symmetry[64] = {0};
  const int nh = std::sscanf(hbuf, "%63s %63s %63s %63s %63s", banner, object, format, field, symmetry);
  BZG_REQUIRE(nh == 5 && lower(banner) == "%%matrixmarket" && lower(object) == "matrix", BZG_ERR_INVALID_ARGUMENT,
              "mm_read_csr: not a Matrix Market matrix header");
  BZG_REQUIRE(lower(format) == "coordinate", BZG_ERR_UNSUPPORTED,
              "mm_read_csr: only the coordinate (sparse) format is supported");
  const std::string fld = lower(field), sym = lower(symmetry);
  BZG_REQUIRE(fld == "real" || fld == "double" || fld == "integer" || fld == "pattern", BZG_ERR_UNSUPPORTED,
              "mm_read_csr: field must be real, double, integer or pattern");
  BZG_REQUIRE(sym == "general" || sym == "symmetric" || sym == "skew-symmetric", BZG_ERR_UNSUPPORTED,
              "mm_read_csr: symmetry must be general, symmetric or skew-symmetric");
  const bool pattern = fld == "pattern";
  std::string line;
  BZG_REQUIRE(next_data_line(fc.f, line), BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: missing size line");
  long long m = -1, n = -1, entries = -1;
  BZG_REQUIRE(std::sscanf(line.c_str(), "%lld %lld %lld", &m, &n, &entries) == 3 && m >= 0 && n >= 0 &&
                  entries >= 0,
              BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: bad size line");
  BZG_REQUIRE(m < (1LL << 31) && n < (1LL << 31), BZG_ERR_UNSUPPORTED, "mm_read_csr: dimension >= 2^31");
  BZG_REQUIRE(sym == "general" || m == n, BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: symmetric file of a non-square matrix");
  struct E {
    int32_t r, c;
    double v;
  };
  std::vector<E> e;
  // the declared count is not trusted for the reservation (a short file fails below)
  e.reserve((size_t)std::min<long long>(entries, 1LL << 24) * (sym == "general" ? 1 : 2));
  for (long long q = 0; q < entries; ++q) {
    BZG_REQUIRE(next_data_line(fc.f, line), BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: fewer entries than declared");
    long long i = 0, j = 0;
    double v = 1.0;
    const int got = pattern ? std::sscanf(line.c_str(), "%lld %lld", &i, &j)
                            : std::sscanf(line.c_str(), "%lld %lld %lf", &i, &j, &v);
    BZG_REQUIRE(got == (pattern ? 2 : 3), BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: bad entry line");
    BZG_REQUIRE(i >= 1 && i <= m && j >= 1 && j <= n, BZG_ERR_INVALID_ARGUMENT, "mm_read_csr: index out of range");
    BZG_REQUIRE(sym == "general" || i >= j, BZG_ERR_INVALID_ARGUMENT,
                "mm_read_csr: symmetric files store the lower triangle only");
    BZG_REQUIRE(sym != "skew-symmetric" || i != j, BZG_ERR_INVALID_ARGUMENT,
                "mm_read_csr: skew-symmetric files have no diagonal entries");
    e.push_back({(int32_t)(i - 1), (int32_t)(j - 1), v});
    if (sym != "general" && i != j) e.push_back({(int32_t)(j - 1), (int32_t)(i - 1), sym == "symmetric" ? v : -v});
  }
  BZG_REQUIRE(e.size() < (1ULL << 31), BZG_ERR_UNSUPPORTED, "mm_read_csr: more than 2^31 - 1 entries");
  // rows in order, columns ascending, file order among equal (row, column): stable sort, then
  // duplicates are added left to right
  std::stable_sort(e.begin(), e.end(), [](const E& a, const E& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  size_t w = 0;
  for (size_t q = 0; q < e.size(); ++q) {
    if (w > 0 && e[w - 1].r == e[q].r && e[w - 1].c == e[q].c)
      e[w - 1].v = e[w - 1].v + e[q].v;
    else
      e[w++] = e[q];
  }
  e.resize(w);
  auto* rowptr = static_cast<int32_t*>(std::malloc(((size_t)m + 1) * sizeof(int32_t)));
  auto* colidx = static_cast<int32_t*>(std::malloc(std::max<size_t>(w, 1) * sizeof(int32_t)));
  auto* val = static_cast<double*>(std::malloc(std::max<size_t>(w, 1) * sizeof(double)));
  if (!rowptr || !colidx || !val) {
    std::free(rowptr);
    std::free(colidx);
    std::free(val);
    bzg::set_error("mm_read_csr: out of host memory");
    return BZG_ERR_OUT_OF_MEMORY;
  }
  std::fill(rowptr, rowptr + m + 1, 0);
  for (size_t q = 0; q < w; ++q) {
    ++rowptr[e[q].r + 1];
    colidx[q] = e[q].c;
    val[q] = e[q].v;
  }
  for (long long r = 0; r < m; ++r) rowptr[r + 1] += rowptr[r];
  out->n_rows = m;
  out->n_cols = n;
  out->nnz = (int64_t)w;
  out->rowptr = rowptr;
  out->colidx = colidx;
  out->val = val;
  return BZG_OK;
}
}  // namespace

extern "C" {

bzg_status bzg_csr_host_free(bzg_csr_host* m) {
  BZG_REQUIRE(m != nullptr, BZG_ERR_INVALID_ARGUMENT, "csr_host_free: null argument");
  std::free(m->rowptr);
  std::free(m->colidx);
  std::free(m->val);
  *m = bzg_csr_host{};
  return BZG_OK;
}

}  // extern "C"
