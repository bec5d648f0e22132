// Serving loop: placeholder entry points until the resident request loop lands.
#include "sb_internal.hpp"

using namespace sb;

extern "C" int sb_server_create(int, size_t, int, sb_server** out) {
  if (out) *out = nullptr;
  return fail(SB_ERR_UNSUPPORTED, "serving loop not built yet");
}
extern "C" int sb_server_submit(sb_server*, const sb_request*, uint64_t*) {
  return fail(SB_ERR_UNSUPPORTED, "serving loop not built yet");
}
extern "C" int sb_server_wait(sb_server*, uint64_t, double, sb_host_result*, sb_request_stats*) {
  return fail(SB_ERR_UNSUPPORTED, "serving loop not built yet");
}
extern "C" int sb_server_stats(const sb_server*, uint64_t*, uint64_t*, size_t*) {
  return fail(SB_ERR_UNSUPPORTED, "serving loop not built yet");
}
extern "C" void sb_server_destroy(sb_server*) {}
