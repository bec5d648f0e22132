// crs-de: the CRS search by differential evolution, with the flag surface the
// reference documents for crs-ocl-de (README.md:88-106).  Outputs as crs.
//
//   -ngen    generations (population size from SB_DE_POP, default 32;
//            F = SB_DE_F (0.5), CR = SB_DE_CR (0.9), seed = SB_DE_SEED (1),
//            a decimal or 0x-hex 64-bit integer)
//   -d       device: 1-based index as in the reference; 0 = the first device
//   -azimuth degrees clockwise from north.  By default the traveltime uses the
//            CRS reference's unsigned midpoint distance |dm| and the value is
//            only checked; with SB_AZIMUTH=project it uses the signed
//            projection of dm onto that direction (SB_OPT_DM_AZIMUTH)
//   -na/-nb/-nc required and parsed as integers, for compatibility: they
//            size the CRS grid search, and differential evolution searches
//            the continuous box instead of a grid
#include <cmath>

#include "cli_common.hpp"

static double env_d(const char* k, double dflt) {
  const char* v = std::getenv(k);
  return v ? std::atof(v) : dflt;
}

// 64-bit seeds: parsed as an unsigned integer (decimal, or hex with 0x), not
// through a double, which would round seeds above 2^53.
static uint64_t env_u64(const char* k, uint64_t dflt) {
  const char* v = std::getenv(k);
  if (!v || !*v) return dflt;
  char* end = nullptr;
  const unsigned long long x = std::strtoull(v, &end, 0);
  if (end == v || *end) cli::log(cli::FAIL, std::string("bad ") + k + ": " + v);
  return (uint64_t)x;
}

int main(int argc, const char** argv) {
  using namespace cli;
  touch_outputs({"crs.a.su", "crs.b.su", "crs.c.su", "crs.coher.su", "crs.stack.su"});
  Parser p;
  p.add("-a0", "A0 constant");
  p.add("-a1", "A1 constant");
  p.add("-aph", "APH constant");
  p.add("-apm", "APM constant");
  p.add("-azimuth", "Azimuth angle");
  p.add("-b0", "B0 constant");
  p.add("-b1", "B1 constant");
  p.add("-c0", "C0 constant");
  p.add("-c1", "C1 constant");
  p.add("-d", "OpenCL device");
  p.add("-i", "Data path");
  p.add("-na", "NA constant");
  p.add("-nb", "NB constant");
  p.add("-nc", "NC constant");
  p.add("-ngen", "Tau constant");
  p.add("-tau", "Tau constant");
  p.add("-v", "Verbosity Level 0-3");
  p.parse(argc, argv);

  sb_de_params prm;
  prm.a0 = to_d(p.get("-a0", true));
  prm.a1 = to_d(p.get("-a1", true));
  prm.b0 = to_d(p.get("-b0", true));
  prm.b1 = to_d(p.get("-b1", true));
  prm.c0 = to_d(p.get("-c0", true));
  prm.c1 = to_d(p.get("-c1", true));
  prm.tau = to_d(p.get("-tau", true));
  for (const char* f : {"-na", "-nb", "-nc"}) (void)to_i(p.get(f, true));
  const double azimuth = to_d(p.get("-azimuth", true));
  const char* az_mode = std::getenv("SB_AZIMUTH");
  const bool project = az_mode && !std::strcmp(az_mode, "project");
  if (az_mode && !project && std::strcmp(az_mode, "distance")) log(FAIL, "SB_AZIMUTH must be project or distance");
  prm.ngen = to_i(p.get("-ngen", true));
  const int aph = to_i(p.get("-aph", true));
  const int apm = to_i(p.get("-apm", true));
  const int d = to_i(p.get("-d", false));
  const std::string path = p.get("-i", true);
  set_verbosity(to_i(p.get("-v", false)));
  prm.pop = (int)env_d("SB_DE_POP", 32);
  prm.F = env_d("SB_DE_F", 0.5);
  prm.CR = env_d("SB_DE_CR", 0.9);
  prm.seed = env_u64("SB_DE_SEED", 1);
  prm.mode = env_mode();

  Session s;
  s.load(path, SB_METHOD_CRS, aph, apm);
  s.open(env_device(d > 0 ? d - 1 : 0));
  if (project) {
    const double deg = std::fmod(std::fmod(azimuth, 360.0) + 360.0, 360.0);
    check(sb_ctx_set_option(s.ctx, SB_OPT_DM_AZIMUTH, (int)std::lround(deg * 1000.0) % 360000), "azimuth");
  }
  const double t0 = now_s();
  HostOut out((size_t)s.ncdp * s.ns);
  check(s.search(SB_GRAPH_CRS_DE, &prm, &out.view), "CRS-DE search");
  const double t = now_s() - t0;
  const double evals = (double)prm.pop * (prm.ngen + 1);
  log_rate(argv[0], s, t, s.traces_per_eval(true) * s.ns * evals, "Mega-Semblances-Trace/s");
  write(s, "crs.a.su", out.a, 1);
  write(s, "crs.b.su", out.b, 1);
  write(s, "crs.c.su", out.c, 1);
  write(s, "crs.coher.su", out.coher, 0);
  write(s, "crs.stack.su", out.stack, 0);
  return EXIT_SUCCESS;
}
