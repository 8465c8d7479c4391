// Autotuner for the row-bin settings -- bin thresholds (and with them each
// hash table's load factor and the long-row cut), the workgroup shape of
// every hash bin, and the learned method's handling of long rows and choice
// of predicted bins -- with an on-disk cache per matrix class.
//
// The reference has no tuning step: its ray-tracing pipeline is fixed at
// compile time (optixSpMSpM.cpp:174-449; src/CMakeLists.txt:348-376).  Here a
// product's row kernels are chosen by per-row sizes; rt_autotune times
// candidate settings on the caller's own operands (a greedy search:
// threshold sets, then one bin's shape at a time, then the learned options),
// checks each candidate's C against the first candidate's, and records the
// fastest for the operands' class; rt_spmspm_ex uses it when
// rt_opts.use_tuned is set.
//
// Matrix class: rounded log2 of A's rows, B's columns, the mean nnz per row
// of A and of B, whether the call is batched (B) or single (S), and the
// method (P / L).  Cache file: the path in RTSPMSPM_TUNE_CACHE, else
// $HOME/.cache/rtspmspm/tune.txt; one line per class:
//   v2 <class> <sym0..sym4> <num0..num4> <learned_heavy> <shape0..shape4> <learned_bins>
// Lines of another version, malformed lines and settings the kernels cannot
// run (tuning_problem) are dropped when the file is read.
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

const char* const kVersion = "v2";

std::mutex g_mu;
std::map<std::string, Tuning> g_cache;
bool g_loaded = false;
std::string g_path;  // the file g_cache was read from

std::string cache_path() {
  if (const char* p = std::getenv("RTSPMSPM_TUNE_CACHE")) return p;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/rtspmspm/tune.txt";
}

int lg(double x) { return x <= 1 ? 0 : (int)std::lround(std::log2(x)); }

// Parses one cache line; false when it is not a usable v2 line.
bool parse_line(const std::string& line, std::string& key, Tuning& t) {
  std::istringstream is(line);
  std::string ver;
  if (!(is >> ver) || ver != kVersion || !(is >> key)) return false;
  bool ok = true;
  for (auto& v : t.sym) ok = ok && (bool)(is >> v);
  for (auto& v : t.num) ok = ok && (bool)(is >> v);
  ok = ok && (bool)(is >> t.learned_heavy);
  for (auto& v : t.shape) ok = ok && (bool)(is >> v);
  ok = ok && (bool)(is >> t.learned_bins);
  std::string extra;
  if (!ok || (is >> extra)) return false;
  return tuning_problem(t).empty();
}

void load_file(const std::string& path, std::map<std::string, Tuning>* into, int& valid,
               int& rejected) {
  valid = rejected = 0;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos || line[0] == '#') continue;
    std::string key;
    Tuning t;
    if (parse_line(line, key, t)) {
      ++valid;
      if (into) (*into)[key] = t;
    } else {
      ++rejected;
    }
  }
}

// (Re)reads the cache when it has not been read or its path has changed.
void load_locked() {
  const std::string path = cache_path();
  if (g_loaded && path == g_path) return;
  g_loaded = true;
  g_path = path;
  g_cache.clear();
  int v, r;
  load_file(path, &g_cache, v, r);
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
  for (auto& kv : g_cache) f << kVersion << ' ' << kv.first << ' ' << tuning_str(kv.second) << '\n';
}

}  // namespace

std::string tuning_problem(const Tuning& t) {
  for (int b = 0; b < 5; ++b) {
    if (t.sym[b] < 1 || t.num[b] < 1) return "thresholds must be positive";
    if (b > 0 && (t.sym[b] <= t.sym[b - 1] || t.num[b] <= t.num[b - 1]))
      return "thresholds must increase from bin to bin";
    if (t.shape[b] < 0 || t.shape[b] > hash_shapes(b))
      return b == 4 ? "shape of bin 4 must be 0, 1 or 2" : "shape must be 0 or 1";
    const int slots = hash_slots(b, t.shape[b]);
    if (t.sym[b] > slots / 2 || t.num[b] > slots / 2)
      return "bin " + std::to_string(b) + " threshold above half its " +
             std::to_string(slots) + "-slot table";
  }
  if (t.sym[4] > t.num[4]) return "sym[4] must not exceed num[4]";
  if (t.learned_heavy != 0 && t.learned_heavy != 1) return "learned_heavy must be 0 or 1";
  if (t.learned_bins < 0 || t.learned_bins > 31) return "learned_bins must be in [0, 31]";
  return "";
}

std::string tuning_str(const Tuning& t) {
  std::ostringstream os;
  for (auto v : t.sym) os << v << ' ';
  for (auto v : t.num) os << v << ' ';
  os << t.learned_heavy;
  for (auto v : t.shape) os << ' ' << v;
  os << ' ' << t.learned_bins;
  return os.str();
}

std::string tuning_report(const Tuning& t) {
  std::ostringstream os;
  os << "bins (slots/max nnz/lanes):";
  for (int b = 0; b < 5; ++b)
    os << ' ' << hash_slots(b, t.shape[b]) << '/' << t.num[b] << '/' << hash_group(b, t.shape[b]);
  os << "; thresholds sym";
  for (auto v : t.sym) os << ' ' << v;
  os << " num";
  for (auto v : t.num) os << ' ' << v;
  os << " learned_heavy " << t.learned_heavy << " shapes";
  for (auto v : t.shape) os << ' ' << v;
  os << " learned_bins " << t.learned_bins;
  return os.str();
}

std::string matrix_class(const Csr& A, const Csr& B, bool batched) {
  std::ostringstream os;
  os << "r" << lg((double)A.rows) << "c" << lg((double)B.cols) << "a"
     << lg(A.rows ? (double)A.nnz / (double)A.rows : 0) << "b"
     << lg(B.rows ? (double)B.nnz / (double)B.rows : 0) << (batched ? "B" : "S");
  return os.str();
}

std::string tune_key(const Csr& A, const Csr& B, int mode, bool batched) {
  return matrix_class(A, B, batched) + (mode == RT_MODE_LEARNED ? "L" : "P");
}

Tuning autotune_search(bool learned, const std::function<double(const Tuning&)>& eval,
                       std::string& log) {
  Tuning best;
  double best_ms = 0;
  int n = 0;
  std::ostringstream os;
  // A later candidate replaces the best when it is faster by more than 1%
  // (noise).
  auto trial = [&](const Tuning& t, const char* what) {
    if (!tuning_problem(t).empty()) {  // e.g. a 4096 cut with bin 4's half-size table
      os << (n ? ", " : "") << what << ":skipped";
      ++n;
      return;
    }
    const double ms = eval(t);  // < 0: the candidate's C differed (rejected)
    os << (n ? ", " : "") << what << ":" << (ms < 0 ? std::string("rejected") : std::to_string(ms));
    ++n;
    if (ms >= 0 && (best_ms == 0 || ms < best_ms * 0.99)) {
      best = t;
      best_ms = ms;
    }
  };
  // Thresholds: default; long-row cut at 2048 / 1024; load factor 1/4 in
  // bins 0-3 (thresholds at a quarter of the tables), alone and with the
  // 2048 cut.  For the learned method they are searched with every row
  // exact (learned_bins 0, the exact shortcut), so its search starts from
  // the plain method's own best.
  const i64 sets[5][5] = {{16, 64, 256, 1024, 4096},
                          {16, 64, 256, 1024, 2048},
                          {16, 64, 256, 512, 1024},
                          {8, 32, 128, 512, 4096},
                          {8, 32, 128, 512, 2048}};
  const char* names[5] = {"default", "cut2048", "cut1024", "load1/4", "load1/4+cut2048"};
  // The 2048 and 1024 cuts pair with bin 4's 4096-slot table (shape 2:
  // load factor at most 1/2, as in the other bins) unless the base keeps the
  // alternate 8192-slot shape; the load-1/4 sets keep 8192 slots.
  auto threshold_sets = [&](const Tuning& base, const char* tag) {
    for (int c = 0; c < 5; ++c) {
      Tuning t = base;
      for (int b = 0; b < 5; ++b) t.sym[b] = t.num[b] = sets[c][b];
      if (t.shape[4] != 1) t.shape[4] = c == 1 || c == 2 ? 2 : 0;
      const std::string w = std::string(tag) + names[c];
      trial(t, w.c_str());
    }
  };
  Tuning start;
  if (learned) start.learned_bins = 0;
  threshold_sets(start, "");
  RT_CHECK(best_ms > 0, RT_E_INVALID, "autotune: no candidate produced a valid result");
  // Workgroup shape, one bin at a time: every other shape of the bin
  // (candidates that do not fit the thresholds are skipped).
  for (int b = 0; b < 5; ++b) {
    const int cur = best.shape[b];
    for (int alt = 0; alt <= hash_shapes(b); ++alt) {
      if (alt == cur) continue;
      Tuning t = best;
      t.shape[b] = alt;
      const std::string w = "shape" + std::to_string(b) + "=" + std::to_string(alt);
      trial(t, w.c_str());
    }
  }
  if (learned) {
    // Learned options on top: predictions for the rows of every bin, of
    // bins 2-4 or of bins 3-4 (rows with many products per output, where a
    // symbolic pass costs more than the features, the prediction and the
    // copy that replace it); long rows sized by the predictor.  A predicted
    // mask that wins has the threshold sets searched again under it.
    const Tuning exact = best;
    const int masks[3] = {31, 28, 24};
    for (int mk : masks) {
      Tuning u = exact;
      u.learned_bins = mk;
      const std::string w = "learned_bins=" + std::to_string(mk);
      trial(u, w.c_str());
    }
    if (best.learned_bins != 0) threshold_sets(best, "masked ");
    Tuning t = best;
    t.learned_heavy = 1;
    trial(t, "learned_heavy");
  }
  log = "best " + std::to_string(best_ms) + " ms of " + std::to_string(n) +
        " candidates; ms per candidate: " + os.str();
  return best;
}

Tuning active_tuning(const Csr& A, const Csr& B, bool use_tuned, int mode, bool batched) {
  if (!use_tuned) return Tuning{};
  std::lock_guard<std::mutex> g(g_mu);
  load_locked();
  auto it = g_cache.find(tune_key(A, B, mode, batched));
  return it == g_cache.end() ? Tuning{} : it->second;
}

void record_tuning(const std::string& key, const Tuning& t) {
  RT_CHECK(tuning_problem(t).empty(), RT_E_INVALID, "autotune: " + tuning_problem(t));
  std::lock_guard<std::mutex> g(g_mu);
  load_locked();
  g_cache[key] = t;
  save_locked();
}

void tune_cache_check(const std::string& path, int& valid, int& rejected) {
  load_file(path, nullptr, valid, rejected);
}

}  // namespace rt
