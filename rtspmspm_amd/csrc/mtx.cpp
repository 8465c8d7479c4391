// Matrix Market I/O.
//
// Reader: the reference reads "x y val" lines into float triplets, skipping
// '%' lines and silently dropping lines it cannot parse
// (optixSpMSpM.cpp:73-131, Sphere.cpp:26-78, cuSparse/src/util.cpp:9-44).
// This reader accepts the same files (coordinate real general, 1-based, any
// entry order -- the shipped *_small.mtx files are column-sorted) and, in
// addition, integer and pattern fields and symmetric storage; indices stay
// integers (the reference carries them as floats, exact below 2^24);
// duplicate entries are summed; the result is row-sorted CSR.
//
// Writer: ref_compat = 1 follows printResult (optixSpMSpM.cpp:452-494):
// row-major, exact zeros dropped (:479), default ostream precision of 6
// significant digits (:484).  The reference re-writes its header in place
// with nnz = 0 (:465,488-491), which overwrites the start of the first entry
// line; ref_compat = 1 writes the true count instead, and ref_compat = 2
// reproduces the overwrite byte for byte.  ref_compat = 0 writes every
// stored entry, including cancellation zeros, with 9 significant digits, so
// that fp32 values round-trip exactly.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "hostcsr.hpp"

namespace rt {

namespace {
bool istarts(const char* s, const char* p) {
  for (; *p; ++s, ++p)
    if (std::tolower((unsigned char)*s) != *p) return false;
  return true;
}
}  // namespace

void mtx_read(const char* path, HostCsr& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) fail(RT_E_IO, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::string buf(sz > 0 ? sz : 0, '\0');
  if (sz > 0 && std::fread(&buf[0], 1, sz, f) != (size_t)sz) {
    std::fclose(f);
    fail(RT_E_IO, std::string("cannot read ") + path);
  }
  std::fclose(f);
  buf.push_back('\0');
  const char* p = buf.c_str();
  const char* end = p + buf.size() - 1;
  bool pattern = false, symmetric = false, skew = false;
  // Banner (optional) and comments.
  if (istarts(p, "%%matrixmarket")) {
    const char* nl = std::strchr(p, '\n');
    std::string ban(p, nl ? nl - p : end - p);
    for (auto& ch : ban) ch = (char)std::tolower((unsigned char)ch);
    if (ban.find("array") != std::string::npos)
      fail(RT_E_IO, "dense (array) Matrix Market files are not supported");
    if (ban.find("complex") != std::string::npos) fail(RT_E_IO, "complex fields are not supported");
    pattern = ban.find("pattern") != std::string::npos;
    symmetric = ban.find("symmetric") != std::string::npos || ban.find("hermitian") != std::string::npos;
    skew = ban.find("skew-symmetric") != std::string::npos;
  }
  auto next_line = [&](const char* q) {
    const char* nl = std::strchr(q, '\n');
    return nl ? nl + 1 : end;
  };
  while (p < end && (*p == '%' || *p == '\n' || *p == '\r')) p = next_line(p);
  char* q;
  const long long rows = std::strtoll(p, &q, 10);
  const long long cols = std::strtoll(q, &q, 10);
  const long long nnz = std::strtoll(q, &q, 10);
  if (q == p || rows < 0 || cols < 0 || nnz < 0) fail(RT_E_IO, "bad Matrix Market size line");
  p = next_line(q);
  std::vector<i64> ri;
  std::vector<i32> ci;
  std::vector<float> vi;
  // Room for the declared entries, but never more than the file can hold
  // (an entry line has at least 4 bytes): a bogus count cannot make the
  // reader reserve memory the entries will not use.
  const long long room = std::min<long long>(nnz, (long long)(end - p) / 4 + 1) * (symmetric ? 2 : 1);
  ri.reserve(room);
  ci.reserve(room);
  vi.reserve(room);
  while (p < end) {
    if (*p == '%' || *p == '\n' || *p == '\r') {
      p = next_line(p);
      continue;
    }
    char* e1;
    const long long r = std::strtoll(p, &e1, 10);
    char* e2;
    const long long c = std::strtoll(e1, &e2, 10);
    if (e1 == p || e2 == e1) {
      p = next_line(p);  // unparsable line: skipped, as in the reference
      continue;
    }
    float v = 1.0f;
    char* e3 = e2;
    if (!pattern) {
      v = std::strtof(e2, &e3);
      if (e3 == e2) {
        p = next_line(p);
        continue;
      }
    }
    if (r < 1 || r > rows || c < 1 || c > cols) fail(RT_E_IO, "entry index out of range");
    ri.push_back(r - 1);
    ci.push_back((i32)(c - 1));
    vi.push_back(v);
    if (symmetric && r != c) {
      ri.push_back(c - 1);
      ci.push_back((i32)(r - 1));
      vi.push_back(skew ? -v : v);
    }
    p = next_line(e3);
  }
  const i64 n = (i64)ri.size();
  // Counting sort by row, then sort by column inside rows and merge duplicates.
  std::vector<i64> rp(rows + 1, 0);
  for (i64 e = 0; e < n; ++e) ++rp[ri[e] + 1];
  for (i64 i = 0; i < rows; ++i) rp[i + 1] += rp[i];
  std::vector<std::pair<i32, float>> ent(n);
  {
    std::vector<i64> cur(rp.begin(), rp.end() - 1);
    for (i64 e = 0; e < n; ++e) ent[cur[ri[e]]++] = {ci[e], vi[e]};
  }
  std::vector<i64> len(rows, 0);
  for (i64 i = 0; i < rows; ++i) {
    auto b0 = ent.begin() + rp[i], b1 = ent.begin() + rp[i + 1];
    std::stable_sort(b0, b1, [](const std::pair<i32, float>& a, const std::pair<i32, float>& b) {
      return a.first < b.first;
    });
    i64 w = rp[i];
    for (i64 e = rp[i]; e < rp[i + 1]; ++e) {
      if (w > rp[i] && ent[w - 1].first == ent[e].first)
        ent[w - 1].second += ent[e].second;
      else
        ent[w++] = ent[e];
    }
    len[i] = w - rp[i];
  }
  out.alloc(rows, cols, 0);
  out.rp[0] = 0;
  for (i64 i = 0; i < rows; ++i) out.rp[i + 1] = out.rp[i] + len[i];
  out.set_nnz(out.rp[rows]);
  for (i64 i = 0; i < rows; ++i)
    for (i64 t = 0; t < len[i]; ++t) {
      out.col[out.rp[i] + t] = ent[rp[i] + t].first;
      out.val[out.rp[i] + t] = ent[rp[i] + t].second;
    }
}

void mtx_write(const char* path, const rt_csr& C, int ref_compat) {
  RT_CHECK(ref_compat >= 0 && ref_compat <= 2, RT_E_INVALID, "ref_compat must be 0, 1 or 2");
  FILE* f = std::fopen(path, "wb");
  if (!f) fail(RT_E_IO, std::string("cannot open ") + path + " for writing");
  i64 count = 0;
  for (i64 e = 0; e < C.nnz; ++e) count += ref_compat ? (C.val[e] != 0.0f) : 1;
  const char* head = "%%MatrixMarket matrix coordinate real general\n";
  std::fputs(head, f);
  if (ref_compat == 2)  // printResult's first size line: no count (optixSpMSpM.cpp:470)
    std::fprintf(f, "%lld %lld \n", (long long)C.rows, (long long)C.cols);
  else
    std::fprintf(f, "%lld %lld %lld\n", (long long)C.rows, (long long)C.cols, (long long)count);
  const char* fmt = ref_compat ? "%lld %lld %g\n" : "%lld %lld %.9g\n";
  for (i64 i = 0; i < C.rows; ++i)
    for (i64 e = C.row_ptr[i]; e < C.row_ptr[i + 1]; ++e) {
      if (ref_compat && C.val[e] == 0.0f) continue;
      std::fprintf(f, fmt, (long long)(i + 1), (long long)C.col[e] + 1, (double)C.val[e]);
    }
  if (ref_compat == 2) {
    // The in-place rewrite with the never-filled map's size, 0
    // (optixSpMSpM.cpp:487-491).
    if (std::fseek(f, 0, SEEK_SET) != 0) {
      std::fclose(f);
      fail(RT_E_IO, std::string("seek failed: ") + path);
    }
    std::fputs(head, f);
    std::fprintf(f, "%lld %lld 0\n", (long long)C.rows, (long long)C.cols);
  }
  if (std::fclose(f) != 0) fail(RT_E_IO, std::string("write failed: ") + path);
}

}  // namespace rt
