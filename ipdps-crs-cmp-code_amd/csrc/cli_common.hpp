// Shared pieces of the drop-in command-line programs cmp, crs and crs-de.
//
// Argument rules follow the reference parser (CMP/OpenMP/src/parser.cpp:52-103):
// "-flag value" pairs only, argc odd and > 1, any unknown flag or a "-h" in
// flag or value position prints the help text and exits with status 1, a
// missing required flag likewise; "-v" defaults to "0".  Logging follows
// CMP/OpenMP/src/log.cpp:37-59: -v 0 FAIL, 1 WARNING, 2 INFO, >= 3 DEBUG;
// INFO to stdout, the rest to stderr, FAIL exits with status 1.
//
// Selection of the search arithmetic and storage, which the reference has no
// flag for, comes from the environment so the flag surface stays identical:
//   SB_MODE  = fast (default) | exact     SB_STORE = f32 (default) | f16 | bf16
//   SB_DEVICE = HIP device index (default 0)
#pragma once

#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "semblance.h"

namespace cli {

class Parser {
 public:
  void add(const std::string& flag, const std::string& help) { args_[flag] = help; }

  void parse(int argc, const char** argv) {
    prog_ = argv[0];
    if (!(argc % 2) || argc == 1) help_exit();
    for (int i = 1; i < argc; i += 2) {
      if (!args_.count(argv[i]) || !std::strcmp(argv[i], "-h") || !std::strcmp(argv[i + 1], "-h")) help_exit();
      raw_.insert({argv[i], argv[i + 1]});  // first occurrence wins, as std::map::insert
    }
  }

  std::string get(const std::string& flag, bool required) const {
    auto a = args_.find(flag);
    auto r = raw_.find(flag);
    if (a != args_.end() && r != raw_.end()) return r->second;
    if (required) help_exit();
    return "0";
  }

  [[noreturn]] void help_exit() const {
    std::printf("Usage: %s [options]\nOptions: \n", prog_.c_str());
    for (auto& kv : args_) std::printf("  %s: %s\n", kv.first.c_str(), kv.second.c_str());
    std::fflush(stdout);
    std::exit(EXIT_FAILURE);
  }

 private:
  std::map<std::string, std::string> args_, raw_;
  std::string prog_ = "prog";
};

enum Level { FAIL = 0, WARNING = 1, INFO = 2, DEBUG = 3 };

inline int& verbosity() {
  static int v = FAIL;
  return v;
}

inline void set_verbosity(int v) { verbosity() = v <= 0 ? FAIL : v == 1 ? WARNING : v == 2 ? INFO : DEBUG; }

inline void log(Level l, const std::string& msg) {
  if (l == FAIL) {
    std::fprintf(stderr, "[FAIL]: %s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
  }
  if (verbosity() < l) return;
  if (l == INFO)
    std::printf("[INFO]: %s\n", msg.c_str());
  else
    std::fprintf(stderr, "[%s]: %s\n", l == WARNING ? "WARNING" : "DEBUG", msg.c_str());
}

inline void check(int rc, const char* what) {
  if (rc != SB_OK) log(FAIL, std::string(what) + ": " + sb_strerror(rc) + ": " + sb_last_error());
}

// std::stof / std::stod / std::stoi with a FAIL exit instead of an uncaught exception.
inline float to_f(const std::string& s) {
  try {
    return std::stof(s);
  } catch (...) {
    log(FAIL, "bad number: " + s);
  }
  return 0;
}
inline double to_d(const std::string& s) {
  try {
    return std::stod(s);
  } catch (...) {
    log(FAIL, "bad number: " + s);
  }
  return 0;
}
inline int to_i(const std::string& s) {
  try {
    return std::stoi(s);
  } catch (...) {
    log(FAIL, "bad integer: " + s);
  }
  return 0;
}

inline int env_mode() {
  const char* m = std::getenv("SB_MODE");
  return (m && !std::strcmp(m, "exact")) ? SB_MODE_EXACT : SB_MODE_FAST;
}
inline int env_store() {
  const char* s = std::getenv("SB_STORE");
  if (s && !std::strcmp(s, "f16")) return SB_STORE_F16;
  if (s && !std::strcmp(s, "bf16")) return SB_STORE_BF16;
  return SB_STORE_F32;
}
inline int env_device(int fallback) {
  const char* d = std::getenv("SB_DEVICE");
  return d ? std::atoi(d) : fallback;
}

// Loads the file and opens the device; the search streams the traces up.
struct Session {
  sb_su_gather* su = nullptr;
  sb_ctx* ctx = nullptr;
  sb_gather* g = nullptr;
  int ncdp = 0, ntr = 0, ns = 0, dt_us = 0, nnbr = 0;
  ~Session() {
    if (g) sb_gather_destroy(g);
    if (ctx) sb_ctx_destroy(ctx);
    if (su) sb_su_free(su);
  }
  // SU input as the reference; files named *.sgy / *.segy are read as SEG-Y.
  void load(const std::string& path, int method, int aph, int apm) {
    std::string low = path;
    for (auto& ch : low) ch = (char)std::tolower((unsigned char)ch);
    auto ends = [&](const char* x) {
      const size_t n = std::strlen(x);
      return low.size() >= n && low.compare(low.size() - n, n, x) == 0;
    };
    const bool segy = ends(".sgy") || ends(".segy");
    check((segy ? sb_segy_load : sb_su_load)(path.c_str(), method, aph, apm, &su), "reading input");
    check(sb_su_info(su, &ncdp, &ntr, &ns, &dt_us, &nnbr), "input");
  }
  void open(int device) { check(sb_ctx_create(device, 0, &ctx), "device"); }
  // Upload overlapped with the search (sb_search_streamed).
  int search(int kind, const void* params, sb_host_result* out) {
    const bool crs = kind != SB_METHOD_CMP;
    return sb_search_streamed(ctx, kind, params, ncdp, sb_su_cdp_offsets(su), ns, dt_us, sb_su_geom(su),
                              sb_su_samples(su), env_store(), crs ? sb_su_nbr_offsets(su) : nullptr,
                              crs ? sb_su_nbr(su) : nullptr, out, nullptr);
  }
  // Sum over CDPs of the traces entering each evaluation (CMP: fold; CRS: supergather).
  double traces_per_eval(bool crs) const {
    const int32_t* off = sb_su_cdp_offsets(su);
    if (!crs) return (double)ntr;
    const int32_t* no = sb_su_nbr_offsets(su);
    const int32_t* nb = sb_su_nbr(su);
    double s = 0;
    for (int i = 0; i < ncdp; i++)
      for (int q = no[i]; q < no[i + 1]; q++) s += off[nb[q] + 1] - off[nb[q]];
    return s;
  }
};

struct HostOut {
  std::vector<int32_t> index;
  std::vector<float> a, b, c, coher, stack;
  sb_host_result view;
  explicit HostOut(size_t n) : index(n), a(n), b(n), c(n), coher(n), stack(n) {
    view = sb_host_result{index.data(), a.data(), b.data(), c.data(), coher.data(), stack.data()};
  }
};

// The throughput INFO line.  By default the OpenMP programs' form,
// "Execution Time: T, <label>: X" (CMP/OpenMP/src/main.cpp:239-241, label
// "Semblances-Trace/s"; CRS/OpenMP/src/main.cpp:326-328, "Mega-Semblances-
// Trace/s" with a giga value).  Run under a name ending in "-cuda" (e.g. a
// cmp-cuda link to cmp) or with SB_INFO=cuda, the CUDA programs' form
// (CMP/CUDA/src/main.cu:316-325, CRS/CUDA/src/main.cu:471-478), whose kernel
// time is the summed device time of the searches.  evals = the reference's
// Semblances-Trace count.
inline bool cuda_info(const char* argv0) {
  const char* e = std::getenv("SB_INFO");
  if (e && !std::strcmp(e, "cuda")) return true;
  const size_t n = std::strlen(argv0);
  return n >= 5 && !std::strcmp(argv0 + n - 5, "-cuda");
}

inline void log_rate(const char* argv0, const Session& s, double t, double evals, const char* omp_label) {
  if (!cuda_info(argv0)) {
    log(INFO, "Execution Time: " + std::to_string(t) + ", " + omp_label + ": " + std::to_string(evals / 1e9 / t));
    return;
  }
  sb_timing tm;
  check(sb_last_timing(s.ctx, &tm), "timing");
  const double k = tm.search_ms / 1e3;
  std::string m = "Total Execution Time: " + std::to_string(t);
  m += ": Giga-Semblances-Trace/s: " + std::to_string(evals / 1e9 / t);
  m += ": Kernel Execution Time: " + std::to_string(k);
  m += ": Kernel Giga-Semblances-Trace/s: " + std::to_string(k > 0 ? evals / 1e9 / k : 0.0);
  log(INFO, m);
}

inline void write(const Session& s, const char* path, const std::vector<float>& v, int rewrite) {
  check(sb_su_write_section(s.su, path, v.data(), rewrite), path);
}

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Creates (truncates) the output files before anything else, as the
// reference does when it constructs its std::ofstream objects.
inline void touch_outputs(const std::vector<const char*>& names) {
  for (const char* n : names) {
    FILE* f = std::fopen(n, "wb");
    if (f) std::fclose(f);
  }
}

}  // namespace cli
