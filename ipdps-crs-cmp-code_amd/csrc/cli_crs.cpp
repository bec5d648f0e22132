// crs: drop-in for the reference CRS programs (crs-omp2 / crs-cuda,
// CRS/OpenMP/src/main.cpp:59-379): same flags, same outputs
// (crs.{a,b,c,coher,stack}.su in the working directory), same INFO line (the
// CUDA program's form when run as crs-cuda or with SB_INFO=cuda).
#include "cli_common.hpp"

int main(int argc, const char** argv) {
  using namespace cli;
  touch_outputs({"crs.a.su", "crs.b.su", "crs.c.su", "crs.coher.su", "crs.stack.su"});
  Parser p;
  p.add("-a0", "A0 constant");
  p.add("-a1", "A1 constant");
  p.add("-na", "NA constant");
  p.add("-b0", "B0 constant");
  p.add("-b1", "B1 constant");
  p.add("-nb", "NB constant");
  p.add("-c0", "C0 constant");
  p.add("-c1", "C1 constant");
  p.add("-nc", "NC constant");
  p.add("-aph", "APH constant");
  p.add("-apm", "APM constant");
  p.add("-tau", "Tau constant");
  p.add("-i", "Data path");
  p.add("-v", "Verbosity Level 0-3");
  p.parse(argc, argv);

  // Reals with stod, counts with stoi (CRS/OpenMP/src/main.cpp:85-98).
  sb_crs_params prm;
  prm.a0 = to_d(p.get("-a0", true));
  prm.a1 = to_d(p.get("-a1", true));
  prm.b0 = to_d(p.get("-b0", true));
  prm.b1 = to_d(p.get("-b1", true));
  prm.c0 = to_d(p.get("-c0", true));
  prm.c1 = to_d(p.get("-c1", true));
  prm.tau = to_d(p.get("-tau", true));
  prm.na = to_i(p.get("-na", true));
  prm.nb = to_i(p.get("-nb", true));
  prm.nc = to_i(p.get("-nc", true));
  const int aph = to_i(p.get("-aph", true));
  const int apm = to_i(p.get("-apm", true));
  const std::string path = p.get("-i", true);
  set_verbosity(to_i(p.get("-v", false)));
  prm.mode = env_mode();

  Session s;
  s.load(path, SB_METHOD_CRS, aph, apm);
  s.open(env_device(0));
  const double t0 = now_s();
  HostOut out((size_t)s.ncdp * s.ns);
  check(s.search(SB_METHOD_CRS, &prm, &out.view), "CRS search");
  const double t = now_s() - t0;
  const double npar = (double)prm.na * prm.nb * prm.nc;
  log_rate(argv[0], s, t, s.traces_per_eval(true) * s.ns * npar, "Mega-Semblances-Trace/s");
  write(s, "crs.a.su", out.a, 1);
  write(s, "crs.b.su", out.b, 1);
  write(s, "crs.c.su", out.c, 1);
  write(s, "crs.coher.su", out.coher, 0);
  write(s, "crs.stack.su", out.stack, 0);
  return EXIT_SUCCESS;
}
