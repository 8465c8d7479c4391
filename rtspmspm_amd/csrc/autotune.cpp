// Autotuner for the row-bin thresholds, the long-row cut and the learned
// method's handling of long rows, with an on-disk cache per matrix class.
//
// The reference has no tuning step: its ray-tracing pipeline is fixed at
// compile time (optixSpMSpM.cpp:174-449).  Here a product's row kernels are
// chosen by thresholds on per-row sizes; rt_autotune times a small set of
// candidate threshold sets on the caller's own operands and records the
// fastest for the operands' class, and rt_spmspm_ex uses the recorded choice
// when rt_opts.use_tuned is set.
//
// Matrix class: rounded log2 of A's rows, B's columns, the mean nnz per row
// of A and of B, whether the call is batched, and the method (P / L).  Cache file: the path in
// RTSPMSPM_TUNE_CACHE, else $HOME/.cache/rtspmspm/tune.txt; one line per
// class: "<class> <sym0..sym4> <num0..num4> <learned_heavy>".
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "learned.hpp"

namespace rt {

namespace {

std::mutex g_mu;
std::map<std::string, Tuning> g_cache;
bool g_loaded = false;

std::string cache_path() {
  if (const char* p = std::getenv("RTSPMSPM_TUNE_CACHE")) return p;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/rtspmspm/tune.txt";
}

int lg(double x) { return x <= 1 ? 0 : (int)std::lround(std::log2(x)); }

void load_locked() {
  if (g_loaded) return;
  g_loaded = true;
  std::ifstream f(cache_path());
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string key;
    Tuning t;
    if (!(is >> key)) continue;
    bool ok = true;
    for (auto& v : t.sym) ok = ok && (bool)(is >> v);
    for (auto& v : t.num) ok = ok && (bool)(is >> v);
    ok = ok && (bool)(is >> t.learned_heavy);
    if (ok) g_cache[key] = t;
  }
}

void save_locked() {
  const std::string path = cache_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) {
    // create parent directories (best effort)
    std::string dir = path.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i)
      if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
  }
  std::ofstream f(path);
  for (auto& kv : g_cache) {
    f << kv.first;
    for (auto v : kv.second.sym) f << ' ' << v;
    for (auto v : kv.second.num) f << ' ' << v;
    f << ' ' << kv.second.learned_heavy << '\n';
  }
}

}  // namespace

std::string matrix_class(const Csr& A, const Csr& B, bool batched) {
  std::ostringstream os;
  os << "r" << lg((double)A.rows) << "c" << lg((double)B.cols) << "a"
     << lg(A.rows ? (double)A.nnz / (double)A.rows : 0) << "b"
     << lg(B.rows ? (double)B.nnz / (double)B.rows : 0) << (batched ? "B" : "S");
  return os.str();
}

std::string tune_key(const Csr& A, const Csr& B, int mode) {
  return matrix_class(A, B, false) + (mode == RT_MODE_LEARNED ? "L" : "P");
}

// Candidate settings: default bins; long-row cut lowered to 2048 / 1024
// (more rows on the bitmap and tile kernels); bins shifted one table up
// (lower load factors); for the learned method each also with long rows
// sized by the predictor.
std::vector<Tuning> tune_candidates(bool learned) {
  std::vector<Tuning> c;
  Tuning t;
  c.push_back(t);
  Tuning cut2k;
  cut2k.sym[4] = cut2k.num[4] = 2048;
  c.push_back(cut2k);
  Tuning cut1k;
  cut1k.sym[3] = cut1k.num[3] = 512;
  cut1k.sym[4] = cut1k.num[4] = 1024;
  c.push_back(cut1k);
  Tuning low;
  const i64 lows[5] = {8, 32, 128, 512, 4096};
  for (int b = 0; b < 5; ++b) low.sym[b] = low.num[b] = lows[b];
  c.push_back(low);
  if (learned) {
    const size_t n = c.size();
    for (size_t i = 0; i < n; ++i) {
      Tuning h = c[i];
      h.learned_heavy = 1;
      c.push_back(h);
    }
  }
  return c;
}

Tuning active_tuning(const Csr& A, const Csr& B, bool use_tuned, int mode) {
  if (!use_tuned) return Tuning{};
  std::lock_guard<std::mutex> g(g_mu);
  load_locked();
  auto it = g_cache.find(tune_key(A, B, mode));
  return it == g_cache.end() ? Tuning{} : it->second;
}

void record_tuning(const std::string& key, const Tuning& t) {
  std::lock_guard<std::mutex> g(g_mu);
  load_locked();
  g_cache[key] = t;
  save_locked();
}

}  // namespace rt
