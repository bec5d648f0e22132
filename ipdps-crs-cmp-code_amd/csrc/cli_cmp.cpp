// cmp: drop-in for the reference CMP programs (cmp-omp2 / cmp-cuda,
// CMP/OpenMP/src/main.cpp:51-278): same flags, same outputs
// (cmp.c.su, cmp.coher.su, cmp.stack.su in the working directory), same INFO
// line (the CUDA program's form when run as cmp-cuda or with SB_INFO=cuda).
#include "cli_common.hpp"

int main(int argc, const char** argv) {
  using namespace cli;
  touch_outputs({"cmp.c.su", "cmp.coher.su", "cmp.stack.su"});
  Parser p;
  p.add("-c0", "C0 constant");
  p.add("-c1", "C1 constant");
  p.add("-nc", "NC constant");
  p.add("-aph", "APH constant");
  p.add("-tau", "Tau constant");
  p.add("-i", "Data path");
  p.add("-v", "Verbosity Level 0-3");
  p.parse(argc, argv);

  // Values parsed as the reference does: stof for reals, stoi for counts.
  const float c0 = to_f(p.get("-c0", true));
  const float c1 = to_f(p.get("-c1", true));
  const float tau = to_f(p.get("-tau", true));
  const int nc = to_i(p.get("-nc", true));
  const int aph = to_i(p.get("-aph", true));
  const std::string path = p.get("-i", true);
  set_verbosity(to_i(p.get("-v", false)));

  Session s;
  s.load(path, SB_METHOD_CMP, aph, 0);
  s.open(env_device(0));
  const double t0 = now_s();
  sb_cmp_params prm{c0, c1, nc, tau, env_mode()};
  HostOut out((size_t)s.ncdp * s.ns);
  check(s.search(SB_METHOD_CMP, &prm, &out.view), "CMP search");
  const double t = now_s() - t0;
  log_rate(argv[0], s, t, (double)s.ntr * s.ns * nc, "Semblances-Trace/s");
  write(s, "cmp.c.su", out.c, 1);
  write(s, "cmp.coher.su", out.coher, 0);
  write(s, "cmp.stack.su", out.stack, 0);
  return EXIT_SUCCESS;
}
