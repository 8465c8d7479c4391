// optixSpMSpM -- command-line entry point kept from the reference
// (optixSpMSpM/src/optixSpMSpM/optixSpMSpM.cpp:496-660), now driving the
// MI355X library through its C ABI.
//
//   optixSpMSpM [-m1 A.mtx] [-m2 B.mtx] [-o C.mtx] [-l log] [-h] [--dim=WxH]
//               [--mode ray|plain|learned] [--ablation full|no_atomic|traverse|no_hit|empty]
//               [--ref-compat] [--ref-header]
//
// Same flags, defaults (Matrix/test1.mtx, Matrix/test2.mtx, Matrix/result.mtx,
// optixSpMSpM.cpp:502-504) and log contract: the -l file receives one number,
// the latency in ms formatted by std::to_string (Timing.h:64-100).  The timed
// region, as in the reference (optixSpMSpM.cpp:585-611), starts after the
// inputs are parsed and resident on the device and ends when C is complete
// on the device.  Input paths are used literally, else looked up under
// ./data/ (the reference's sutil::sampleDataFilePath, sutil.cpp:181-211).
// Exit code 0 on success, 1 on any error (optixSpMSpM.cpp:654-658).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/rt_spmspm.h"

namespace {

void usage(const char* argv0) {
  std::cerr << "Usage  : " << argv0 << " [options]\n"
            << "Options: --output | -o <filename>    Specify file for the result matrix\n"
            << "         --mat1 | -m1 <filename>     Specify file for matrix 1\n"
            << "         --mat2 | -m2 <filename>     Specify file for matrix 2\n"
            << "         --log  | -l  <filename>     Specify file for log\n"
            << "         --mode ray|plain|learned    SpGEMM method (default ray)\n"
            << "         --ablation full|no_atomic|traverse|no_hit|empty   (ray method)\n"
            << "         --ref-compat                write C as the reference does\n"
            << "         --ref-header                ... including its in-place header rewrite\n"
            << "         --help | -h                 Print this usage message\n";
  std::exit(1);
}

std::string resolve(const std::string& p) {
  std::ifstream f(p);
  if (f.good()) return p;
  const std::string alt = "data/" + p;
  std::ifstream g(alt);
  if (g.good()) return alt;
  return p;
}

void check(int rc, const char* what) {
  if (rc != RT_OK) throw std::runtime_error(std::string(what) + ": " + rt_last_error());
}

}  // namespace

int main(int argc, char* argv[]) {
  std::string m1 = "Matrix/test1.mtx", m2 = "Matrix/test2.mtx", out = "Matrix/result.mtx";
  std::string log;
  int mode = RT_MODE_RAY, abl = RT_ABL_FULL, ref_compat = 0;
  for (int i = 1; i < argc; ++i) {
    const std::string a(argv[i]);
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) usage(argv[0]);
      return argv[++i];
    };
    if (a == "--help" || a == "-h") usage(argv[0]);
    else if (a == "--output" || a == "-o") out = next();
    else if (a == "--mat1" || a == "-m1") m1 = next();
    else if (a == "--mat2" || a == "-m2") m2 = next();
    else if (a == "--log" || a == "-l") log = next();
    else if (a.rfind("--dim=", 0) == 0) { /* image size of the reference; unused */ }
    else if (a == "--ref-compat") ref_compat = 1;
    else if (a == "--ref-header") ref_compat = 2;
    else if (a == "--mode") {
      const std::string v = next();
      if (v == "ray") mode = RT_MODE_RAY;
      else if (v == "plain") mode = RT_MODE_PLAIN;
      else if (v == "learned") mode = RT_MODE_LEARNED;
      else usage(argv[0]);
    } else if (a == "--ablation") {
      const std::string v = next();
      if (v == "full") abl = RT_ABL_FULL;
      else if (v == "no_atomic") abl = RT_ABL_NO_ATOMIC;
      else if (v == "traverse") abl = RT_ABL_TRAVERSE;
      else if (v == "no_hit") abl = RT_ABL_NO_HIT;
      else if (v == "empty") abl = RT_ABL_EMPTY;
      else usage(argv[0]);
    } else {
      std::cerr << "Unknown option '" << a << "'\n";
      usage(argv[0]);
    }
  }
  rt_csr hA{}, hB{}, dA{}, dB{}, dC{}, hC{};
  int status = 0;
  try {
    static const char* banner[] = {"ATOMIC", "ARCHSUP", "NOMEM", "NOINT", "NOTHING"};
    std::cout << (mode == RT_MODE_RAY ? banner[abl] : (mode == RT_MODE_PLAIN ? "PLAIN" : "LEARNED"))
              << std::endl;
    check(rt_mtx_read(resolve(m2).c_str(), &hB), "reading matrix 2");
    check(rt_mtx_read(resolve(m1).c_str(), &hA), "reading matrix 1");
    check(rt_csr_to_device(&hA, &dA), "uploading matrix 1");
    check(rt_csr_to_device(&hB, &dB), "uploading matrix 2");
    rt_opts o{};
    o.mode = mode;
    o.ablation = abl;
    const auto t0 = std::chrono::high_resolution_clock::now();
    check(rt_spmspm_ex(&dA, &dB, &dC, &o, nullptr), "computing C");
    const auto t1 = std::chrono::high_resolution_clock::now();
    const double ms = std::chrono::duration<double>(t1 - t0).count() * 1000.0;
    check(rt_csr_to_host(&dC, &hC), "downloading C");
    check(rt_mtx_write(out.c_str(), &hC, ref_compat), "writing C");
    std::printf("success!");
    std::fflush(stdout);
    if (!log.empty()) {
      std::ofstream lf(log);
      if (!lf) std::cerr << "Failed to open file for writing." << std::endl;
      lf << std::to_string(ms) << std::flush;
    }
  } catch (const std::exception& e) {
    std::cerr << "Caught exception: " << e.what() << "\n";
    status = 1;
  }
  rt_csr_free(&hA);
  rt_csr_free(&hB);
  rt_csr_free(&dA);
  rt_csr_free(&dB);
  rt_csr_free(&dC);
  rt_csr_free(&hC);
  return status;
}
