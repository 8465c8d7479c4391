// Host-only exercise of the library's CPU code for the AddressSanitizer +
// UndefinedBehaviorSanitizer build (`make asan`, run by the CPU test suite):
// the seeded generators, Matrix Market reading and writing (every writer
// mode, malformed files), the host CSR checks of the C ABI and the
// tuning-cache parser.  It makes no HIP call and links no GPU code.
//
//   asan_host <scratch directory>     exit status 0 when every check holds
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "hostcsr.hpp"
#include "learned.hpp"

using namespace rt;

namespace {

int g_bad = 0;

void expect(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what.c_str());
    ++g_bad;
  }
}

rt_csr view(const HostCsr& h) {
  rt_csr v{};
  v.rows = h.rows;
  v.cols = h.cols;
  v.nnz = h.nnz;
  v.row_ptr = h.rp;
  v.col = h.col;
  v.val = h.val;
  v.mem = RT_MEM_HOST;
  return v;
}

// True when f throws an rt::Error with the given code.
template <class F>
bool throws(int code, F&& f) {
  try {
    f();
  } catch (const Error& e) {
    return e.code == code;
  }
  return false;
}

void check_valid(const HostCsr& h, const std::string& name) {
  const rt_csr v = view(h);
  expect(!throws(RT_E_INVALID, [&] { validate_host(&v, name.c_str()); }),
         name + ": generated CSR fails the host check");
}

void write_file(const std::string& path, const std::string& text) {
  std::ofstream f(path, std::ios::binary);
  f << text;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "/tmp";
  // ---- generators --------------------------------------------------------
  {
    HostCsr a, b, c, d, e, f, g, P, R, BA, BB;
    gen_uniform_k(1000, 8, 0, a);
    gen_rmat(10, 16, 0.57, 0.19, 0.19, 1, b);
    gen_stencil27(6, 2, c);
    gen_aggregation(6, P, R);
    gen_lognormal_rows(500, 2000, 2.5, 1.0, 1, 200, 3, 0, d);
    gen_fixed_normal(2000, 500, 4, 3, 1, e);
    gen_poisson(1000, 8.0, 4, 0, f);
    std::vector<i64> ar(16), br(16), bc(16);
    gen_batch_poisson(16, 16, 2000, 4.0, 32.0, 4, ar.data(), br.data(), bc.data(), BA, BB);
    check_valid(a, "uniform");
    check_valid(b, "rmat");
    check_valid(c, "stencil");
    check_valid(P, "P");
    check_valid(R, "R");
    check_valid(d, "lognormal");
    check_valid(e, "fixed_normal");
    check_valid(f, "poisson");
    expect(c.nnz == 16 * 16 * 16, "27-point stencil on 6^3 has (3*6-2)^3 entries");
    expect(throws(RT_E_INVALID, [&] { gen_uniform_k(10, 11, 0, g); }),
           "k > n is rejected");
  }
  // ---- Matrix Market -------------------------------------------------------
  {
    HostCsr a;
    gen_uniform_k(300, 5, 7, a);
    const rt_csr v = view(a);
    for (int mode = 0; mode <= 2; ++mode) {
      const std::string p = dir + "/asan_w" + std::to_string(mode) + ".mtx";
      mtx_write(p.c_str(), v, mode);
      if (mode == 0) {
        HostCsr back;
        mtx_read(p.c_str(), back);
        bool same = back.rows == a.rows && back.cols == a.cols && back.nnz == a.nnz;
        for (i64 i = 0; same && i <= a.rows; ++i) same = back.rp[i] == a.rp[i];
        for (i64 q = 0; same && q < a.nnz; ++q)
          same = back.col[q] == a.col[q] && back.val[q] == a.val[q];
        expect(same, "exact-format round trip");
      }
    }
    expect(throws(RT_E_INVALID, [&] { mtx_write((dir + "/x.mtx").c_str(), v, 3); }),
           "writer mode 3 is rejected");
    const char* bad[] = {
        "",                                                               // empty file
        "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",     // dense format
        "%%MatrixMarket matrix coordinate real general\n",                // no size line
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n",  // row out of range
        "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 0 1.0\n",  // column 0
        "%%MatrixMarket matrix coordinate real general\n-3 3 1\n1 1 1.0\n",
    };
    int k = 0;
    for (const char* text : bad) {
      const std::string p = dir + "/asan_bad" + std::to_string(k++) + ".mtx";
      write_file(p, text);
      bool threw = false;
      try {
        HostCsr h;
        mtx_read(p.c_str(), h);
      } catch (const Error&) {
        threw = true;
      }
      expect(threw, "malformed file " + std::to_string(k - 1) + " is rejected");
    }
    // Tolerated as the reference tolerates them (optixSpMSpM.cpp:113-116):
    // fewer entries than declared, an unparsable line (skipped), and a
    // declared count far beyond the file (nothing is reserved for it).
    const char* odd[] = {
        "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1\n2 2 5\n",
        "%%MatrixMarket matrix coordinate real general\n3 3 999999999999999\n2 2 5\n",
    };
    for (const char* text : odd) {
      const std::string p = dir + "/asan_odd" + std::to_string(k++) + ".mtx";
      write_file(p, text);
      HostCsr h;
      mtx_read(p.c_str(), h);
      expect(h.rows == 3 && h.nnz == 1, "tolerated file " + std::to_string(k - 1));
    }
    expect(throws(RT_E_IO, [&] {
             HostCsr h;
             mtx_read((dir + "/absent.mtx").c_str(), h);
           }),
           "a missing file is RT_E_IO");
  }
  // ---- host CSR checks -------------------------------------------------------
  {
    i64 rp[3] = {0, 2, 3};
    i32 col[3] = {0, 2, 1};
    float val[3] = {1, 2, 3};
    rt_csr m{2, 3, 3, rp, col, val, RT_MEM_HOST, 0};
    expect(!throws(RT_E_INVALID, [&] { validate_host(&m, "m"); }), "valid CSR passes");
    col[1] = 3;
    expect(throws(RT_E_INVALID, [&] { validate_host(&m, "m"); }), "column out of range");
    col[1] = 0;
    expect(throws(RT_E_INVALID, [&] { validate_host(&m, "m"); }), "duplicate column");
    col[1] = 2;
    rp[1] = 4;
    expect(throws(RT_E_INVALID, [&] { validate_host(&m, "m"); }), "row_ptr not monotone");
    rp[1] = 2;
    rt_csr n = m;
    n.row_ptr = nullptr;
    expect(throws(RT_E_INVALID, [&] { check_csr_header(&n, "n"); }), "NULL row_ptr");
    n = m;
    n.mem = 7;
    expect(throws(RT_E_INVALID, [&] { check_csr_header(&n, "n"); }), "unknown memory space");
    n = m;
    n.cols = (i64)1 << 40;
    expect(throws(RT_E_LIMIT, [&] { check_csr_header(&n, "n"); }), "columns beyond int32");
  }
  // ---- tuning cache ----------------------------------------------------------
  {
    const std::string p = dir + "/asan_tune.txt";
    write_file(p,
               "v2 k 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 0 31\n"
               "v2 k 17 64 256 1024 2048 17 64 256 1024 2048 0 0 0 0 0 0 31\n"
               "v2 k 16 64\n"
               "garbage\n");
    int valid = 0, rejected = 0;
    tune_cache_check(p, valid, rejected);
    expect(valid == 1 && rejected == 3, "tuning cache: one valid line, three rejected");
  }
  if (g_bad) return 1;
  std::printf("asan host checks: ok\n");
  return 0;
}
