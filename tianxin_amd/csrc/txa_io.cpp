// MatrixMarket export and import of the assembled system (host side).
//
// Replaces TpetraLinearObjContainer::writeMatrixMarket
// (lof/Panzer_TpetraLinearObjContainer.hpp:359-367: A through Tpetra's sparse
// MatrixMarket writer, f as "b_vec.mm" and x as "x_vec.mm" through its dense
// writer) and TpetraLinearObjFactory::readVector
// (lof/Panzer_TpetraLinearObjFactory_impl.hpp:743-772: a dense MatrixMarket file
// read into x, dx/dt, d2x/dt2 or f).
//
// Format: the MatrixMarket exchange format, "coordinate real general" for A and
// "array real general" (one column) for vectors.  Indices are global (GID + 1, the
// format is 1-based); rows follow the one-rank row map, i.e. LID order, and the
// entries of a row follow the CRS (sorted columns).  Values are written in the
// shortest decimal form that reads back to the same double, so a file read back
// reproduces the device arrays bit for bit.
#include <charconv>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "txa_internal.hpp"

namespace txa {
namespace {

// Buffered text output to a C stream.
class Out {
 public:
  explicit Out(const char *path) : f_(std::fopen(path, "wb")), buf_(1 << 20) {
    TXA_REQUIRE(f_ != nullptr, TXA_ERR_INVALID, std::string("cannot open ") + path + " for writing");
  }
  ~Out() {
    if (f_) std::fclose(f_);
  }
  void str(const char *s) {
    const size_t n = std::strlen(s);
    reserve(n);
    std::memcpy(buf_.data() + len_, s, n);
    len_ += n;
  }
  void i64(int64_t v) {
    reserve(24);
    len_ = static_cast<size_t>(std::to_chars(buf_.data() + len_, buf_.data() + buf_.size(), v).ptr - buf_.data());
  }
  void f64(double v) {
    reserve(32);
    len_ = static_cast<size_t>(std::to_chars(buf_.data() + len_, buf_.data() + buf_.size(), v).ptr - buf_.data());
  }
  void ch(char c) {
    reserve(1);
    buf_[len_++] = c;
  }
  void close() {
    flush();
    const int rc = std::fclose(f_);
    f_ = nullptr;
    TXA_REQUIRE(rc == 0, TXA_ERR_INVALID, "write failed");
  }

 private:
  void reserve(size_t n) {
    if (len_ + n > buf_.size()) flush();
  }
  void flush() {
    if (len_ && std::fwrite(buf_.data(), 1, len_, f_) != len_) throw Error(TXA_ERR_INVALID, "write failed");
    len_ = 0;
  }
  FILE *f_;
  std::vector<char> buf_;
  size_t len_ = 0;
};

template <typename T>
std::vector<T> to_host(const T *dev, size_t n) {
  std::vector<T> h(n);
  if (n) TXA_HIP(hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}

void write_dense(const char *path, const std::vector<double> &v) {
  Out o(path);
  o.str("%%MatrixMarket matrix array real general\n");
  o.i64(static_cast<int64_t>(v.size()));
  o.str(" 1\n");
  for (double x : v) {
    o.f64(x);
    o.ch('\n');
  }
  o.close();
}

// Whole file into memory.
std::string slurp(const char *path) {
  FILE *f = std::fopen(path, "rb");
  TXA_REQUIRE(f != nullptr, TXA_ERR_INVALID, std::string("cannot open ") + path);
  std::string s;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, n);
  std::fclose(f);
  return s;
}

}  // namespace

void write_matrix_market(Mesh &m, const double *vals, const double *f, const double *x, const char *a_path,
                         const char *b_path, const char *x_path) {
  TXA_REQUIRE(m.has_graph, TXA_ERR_LOGIC, "graph not built");
  const std::vector<int64_t> gid = to_host(m.owned.p, static_cast<size_t>(m.ndofs));
  if (vals && a_path) {
    const std::vector<int64_t> rp = to_host(m.rowptr.p, static_cast<size_t>(m.ndofs) + 1);
    const std::vector<int32_t> col = to_host(m.col.p, static_cast<size_t>(m.nnz));
    const std::vector<double> v = to_host(vals, static_cast<size_t>(m.nnz));
    Out o(a_path);
    o.str("%%MatrixMarket matrix coordinate real general\n");
    o.i64(m.ndofs);
    o.ch(' ');
    o.i64(m.ndofs);
    o.ch(' ');
    o.i64(m.nnz);
    o.ch('\n');
    for (int64_t r = 0; r < m.ndofs; ++r)
      for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
        o.i64(gid[r] + 1);
        o.ch(' ');
        o.i64(gid[col[k]] + 1);
        o.ch(' ');
        o.f64(v[k]);
        o.ch('\n');
      }
    o.close();
  }
  if (f && b_path) write_dense(b_path, to_host(f, static_cast<size_t>(m.ndofs)));
  if (x && x_path) write_dense(x_path, to_host(x, static_cast<size_t>(m.ndofs)));
}

void read_vector_mm(const char *path, double *dst, int64_t n) {
  const std::string s = slurp(path);
  const char *p = s.data(), *end = p + s.size();
  auto line_end = [&](const char *q) {
    const char *e = static_cast<const char *>(std::memchr(q, '\n', static_cast<size_t>(end - q)));
    return e ? e : end;
  };
  // banner, then comment lines
  const char *e = line_end(p);
  const std::string banner(p, e);
  TXA_REQUIRE(banner.rfind("%%MatrixMarket matrix array real general", 0) == 0, TXA_ERR_INVALID,
              std::string(path) + ": not a MatrixMarket array real general file");
  p = e < end ? e + 1 : end;
  while (p < end && *p == '%') {
    e = line_end(p);
    p = e < end ? e + 1 : end;
  }
  auto next_token = [&](const char *&q) {
    while (q < end && (*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r')) ++q;
    const char *b = q;
    while (q < end && !(*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r')) ++q;
    return std::make_pair(b, q);
  };
  int64_t rows = 0, cols = 0;
  auto t = next_token(p);
  TXA_REQUIRE(std::from_chars(t.first, t.second, rows).ec == std::errc(), TXA_ERR_INVALID, "bad size line");
  t = next_token(p);
  TXA_REQUIRE(std::from_chars(t.first, t.second, cols).ec == std::errc(), TXA_ERR_INVALID, "bad size line");
  TXA_REQUIRE(cols == 1, TXA_ERR_INVALID, "a vector file has one column");
  TXA_REQUIRE(rows == n, TXA_ERR_INVALID,
              std::string(path) + ": " + std::to_string(rows) + " entries, expected " + std::to_string(n));
  std::vector<double> h(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    t = next_token(p);
    TXA_REQUIRE(t.first < t.second && std::from_chars(t.first, t.second, h[i]).ec == std::errc(), TXA_ERR_INVALID,
                std::string(path) + ": bad or missing entry " + std::to_string(i + 1));
  }
  if (n) TXA_HIP(hipMemcpy(dst, h.data(), static_cast<size_t>(n) * sizeof(double), hipMemcpyHostToDevice));
}

}  // namespace txa
