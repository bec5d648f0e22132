// SU file input/output and CDP binning with the reference's semantics.
//
// Reader:  su_trace::fgettr       CMP/OpenMP/src/su_trace.cpp:37-51
// Filter:  su_gather::su_gather   CMP/OpenMP/src/su_gather.cpp:43-55 (CMP, float)
//                                 CRS/OpenMP/src/su_gather.cpp:214-226 (CRS, double)
// Order:   CMP std::sort by fold descending (su_gather.cpp:68, su_cdp.hpp:53-55)
//          CRS ascending cdp key (std::map order)
// Lists:   CRS neighbour CDPs within apm, CRS/OpenMP/src/su_gather.cpp:241-258
// Writer:  output sections, CMP/OpenMP/src/main.cpp:244-260 (only parameter
//          sections rewrite the header, :249-251)
#include "sb_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>

namespace sb {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace sb

using namespace sb;

namespace {

template <typename T>
T rd(const uint8_t* h, int off) {
  T v;
  std::memcpy(&v, h + off, sizeof(T));
  return v;
}

template <typename T>
void wr(uint8_t* h, int off, T v) {
  std::memcpy(h + off, &v, sizeof(T));
}

sb_trace_geom geom_of(const uint8_t* h) {
  sb_trace_geom g;
  g.cdp = rd<int32_t>(h, kOffCdp);
  g.scalco = rd<int16_t>(h, kOffScalco);
  g.sx = rd<int32_t>(h, kOffSx);
  g.sy = rd<int32_t>(h, kOffSy);
  g.gx = rd<int32_t>(h, kOffGx);
  g.gy = rd<int32_t>(h, kOffGy);
  return g;
}

// su_trace::fscalco, CMP/OpenMP/src/su_trace.cpp:90-95: 0 -> 1, >0 -> s, <0 -> 1/s.
float fscalco(int32_t scalco) {
  if (scalco == 0) return 1.0f;
  if (scalco > 0) return (float)scalco;
  return 1.0f / (float)scalco;
}

// su_trace::halfoffset_x / midpoint_x: float * int, then * 0.5 (a double
// literal), returned as float.
float half_x(const sb_trace_geom& g) { return (float)((double)(fscalco(g.scalco) * (float)(g.gx - g.sx)) * 0.5); }
float half_y(const sb_trace_geom& g) { return (float)((double)(fscalco(g.scalco) * (float)(g.gy - g.sy)) * 0.5); }
float mid_x(const sb_trace_geom& g) { return (float)((double)(fscalco(g.scalco) * (float)(g.gx + g.sx)) * 0.5); }
float mid_y(const sb_trace_geom& g) { return (float)((double)(fscalco(g.scalco) * (float)(g.gy + g.sy)) * 0.5); }

}  // namespace

extern "C" int sb_keep_trace(const sb_trace_geom* g, int method, int aph) {
  if (!g) return 0;
  const int lim = aph * aph;
  if (method == SB_METHOD_CMP) {
    // real = float in CMP: float products and sum, compared with int -> float.
    float hx = half_x(*g), hy = half_y(*g);
    float hh = hx * hx;
    float yy = hy * hy;
    float s = hh + yy;
    return s > (float)lim ? 0 : 1;
  }
  // real = double in CRS.
  double hx = half_x(*g), hy = half_y(*g);
  double s = hx * hx + hy * hy;
  return s > (double)lim ? 0 : 1;
}

extern "C" int sb_crs_neighbors(int ncdp, const sb_trace_geom* first_geom, int apm,
                                int32_t* nbr_off, int32_t* nbr, int* nnbr) {
  if (ncdp < 0 || (ncdp > 0 && !first_geom) || !nnbr) return fail(SB_ERR_INVALID, "sb_crs_neighbors: bad arguments");
  std::vector<float> mx(ncdp), my(ncdp);
  for (int i = 0; i < ncdp; i++) {
    mx[i] = mid_x(first_geom[i]);
    my[i] = mid_y(first_geom[i]);
  }
  const float lim = (float)(apm * apm);
  int64_t total = 0;
  if (nbr_off) nbr_off[0] = 0;
  for (int i = 0; i < ncdp; i++) {
    for (int j = 0; j < ncdp; j++) {
      float dx = mx[j] - mx[i];
      float dy = my[j] - my[i];
      float dxx = dx * dx;
      float dyy = dy * dy;
      if (dxx + dyy <= lim) {
        if (nbr) nbr[total] = j;
        total++;
      }
    }
    if (nbr_off) nbr_off[i + 1] = (int32_t)total;
  }
  if (total > INT32_MAX) return fail(SB_ERR_INVALID, "sb_crs_neighbors: neighbour lists too long");
  *nnbr = (int)total;
  return SB_OK;
}

extern "C" int sb_su_load(const char* path, int method, int aph, int apm, sb_su_gather** out) {
  if (!path || !out || (method != SB_METHOD_CMP && method != SB_METHOD_CRS))
    return fail(SB_ERR_INVALID, "sb_su_load: bad arguments");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(SB_ERR_IO, std::string("cannot open ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);

  struct Cdp {
    std::vector<SuHeader> hdr;
    std::vector<float> data;
  };
  std::map<int, Cdp> by_cdp;
  int ns = -1, dt_first = 0;
  SuHeader h;
  std::vector<float> buf;
  while (true) {
    size_t got = std::fread(h.raw, 1, kSuHeaderBytes, f);
    if (got < (size_t)kSuHeaderBytes) break;  // end of file (a partial header ends the read, as fgettr)
    int tns = rd<uint16_t>(h.raw, kOffNs);
    buf.resize(tns);
    if (std::fread(buf.data(), sizeof(float), tns, f) != (size_t)tns)
      return fail(SB_ERR_IO, "Bad input data");
    sb_trace_geom g = geom_of(h.raw);
    if (!sb_keep_trace(&g, method, aph)) continue;
    if (ns < 0) ns = tns;
    if (tns != ns) return fail(SB_ERR_FORMAT, "traces with different ns in one file");
    Cdp& c = by_cdp[g.cdp];
    c.hdr.push_back(h);
    c.data.insert(c.data.end(), buf.begin(), buf.end());
  }
  if (by_cdp.empty()) return fail(SB_ERR_FORMAT, "no trace passed the aperture test");
  (void)dt_first;

  // CDP order.  CMP: std::sort with the reference comparator (fold descending)
  // over the std::map order; the permutation depends only on comparator
  // outcomes, so sorting (fold, position) records reproduces it.
  struct Rec {
    size_t fold;
    int pos;
  };
  std::vector<Rec> order;
  std::vector<Cdp*> cdps;
  for (auto& kv : by_cdp) {
    order.push_back({kv.second.hdr.size(), (int)cdps.size()});
    cdps.push_back(&kv.second);
  }
  if (method == SB_METHOD_CMP)
    std::sort(order.begin(), order.end(), [](const Rec& a, const Rec& b) { return a.fold > b.fold; });

  std::unique_ptr<sb_su_gather> G(new sb_su_gather);
  G->method = method;
  G->ns = ns;
  size_t ntr = 0;
  for (auto& r : order) ntr += r.fold;
  if (ntr > (size_t)INT32_MAX) return fail(SB_ERR_FORMAT, "too many traces");
  G->cdp_off.resize(order.size() + 1);
  G->geom.resize(ntr);
  G->samples.resize(ntr * (size_t)ns);
  G->first_hdr.resize(order.size());
  size_t t = 0;
  G->cdp_off[0] = 0;
  for (size_t i = 0; i < order.size(); i++) {
    Cdp* c = cdps[order[i].pos];
    G->first_hdr[i] = c->hdr[0];
    for (size_t j = 0; j < c->hdr.size(); j++, t++) G->geom[t] = geom_of(c->hdr[j].raw);
    std::memcpy(G->samples.data() + (t - c->hdr.size()) * ns, c->data.data(), c->data.size() * sizeof(float));
    G->cdp_off[i + 1] = (int32_t)t;
  }
  // dt of the first trace of the first CDP in gather order (su_gather.cpp:82).
  G->dt_us = rd<uint16_t>(G->first_hdr[0].raw, kOffDt);

  if (method == SB_METHOD_CRS) {
    const int ncdp = (int)order.size();
    std::vector<sb_trace_geom> first(ncdp);
    for (int i = 0; i < ncdp; i++) first[i] = geom_of(G->first_hdr[i].raw);
    int nn = 0;
    int rc = sb_crs_neighbors(ncdp, first.data(), apm, nullptr, nullptr, &nn);
    if (rc) return rc;
    G->nbr_off.resize(ncdp + 1);
    G->nbr.resize(nn);
    rc = sb_crs_neighbors(ncdp, first.data(), apm, G->nbr_off.data(), G->nbr.data(), &nn);
    if (rc) return rc;
  }
  *out = G.release();
  return SB_OK;
}

extern "C" void sb_su_free(sb_su_gather* g) { delete g; }

extern "C" int sb_su_info(const sb_su_gather* g, int* ncdp, int* ntr, int* ns, int* dt_us, int* nnbr) {
  if (!g) return fail(SB_ERR_INVALID, "sb_su_info: null gather");
  if (ncdp) *ncdp = (int)g->cdp_off.size() - 1;
  if (ntr) *ntr = (int)g->geom.size();
  if (ns) *ns = g->ns;
  if (dt_us) *dt_us = g->dt_us;
  if (nnbr) *nnbr = (int)g->nbr.size();
  return SB_OK;
}

extern "C" const int32_t* sb_su_cdp_offsets(const sb_su_gather* g) { return g ? g->cdp_off.data() : nullptr; }
extern "C" const float* sb_su_samples(const sb_su_gather* g) { return g ? g->samples.data() : nullptr; }
extern "C" const sb_trace_geom* sb_su_geom(const sb_su_gather* g) { return g ? g->geom.data() : nullptr; }
extern "C" const int32_t* sb_su_nbr_offsets(const sb_su_gather* g) {
  return (g && !g->nbr_off.empty()) ? g->nbr_off.data() : nullptr;
}
extern "C" const int32_t* sb_su_nbr(const sb_su_gather* g) { return (g && !g->nbr_off.empty()) ? g->nbr.data() : nullptr; }

extern "C" int sb_su_write_section(const sb_su_gather* g, const char* path, const float* values,
                                   int rewrite_header) {
  if (!g || !path || !values) return fail(SB_ERR_INVALID, "sb_su_write_section: bad arguments");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(SB_ERR_IO, std::string("cannot create ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  const int ncdp = (int)g->cdp_off.size() - 1;
  for (int i = 0; i < ncdp; i++) {
    SuHeader h = g->first_hdr[i];
    if (rewrite_header) {
    int32_t sx = rd<int32_t>(h.raw, kOffSx), gx = rd<int32_t>(h.raw, kOffGx);
    int32_t sy = rd<int32_t>(h.raw, kOffSy), gy = rd<int32_t>(h.raw, kOffGy);
    int32_t mx = (sx + gx) >> 1, my = (sy + gy) >> 1;
    wr<int32_t>(h.raw, kOffOffset, 0);
    wr<int32_t>(h.raw, kOffSx, mx);
    wr<int32_t>(h.raw, kOffGx, mx);
    wr<int32_t>(h.raw, kOffSy, my);
    wr<int32_t>(h.raw, kOffGy, my);
    }
    if (std::fwrite(h.raw, 1, kSuHeaderBytes, f) != (size_t)kSuHeaderBytes ||
        std::fwrite(values + (size_t)i * g->ns, sizeof(float), g->ns, f) != (size_t)g->ns)
      return fail(SB_ERR_IO, std::string("write failed: ") + path);
  }
  return SB_OK;
}

extern "C" int sb_su_write_traces(const char* path, int ntr, int ns, int dt_us, const sb_trace_geom* geom,
                                  const float* samples) {
  if (!path || ntr < 0 || ns <= 0 || ns > 65535 || dt_us <= 0 || dt_us > 65535 || (ntr && (!geom || !samples)))
    return fail(SB_ERR_INVALID, "sb_su_write_traces: bad arguments");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(SB_ERR_IO, std::string("cannot create ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  int32_t last_cdp = INT32_MIN, cdpt = 0;
  for (int i = 0; i < ntr; i++) {
    SuHeader h;
    std::memset(h.raw, 0, sizeof h.raw);
    const sb_trace_geom& g = geom[i];
    cdpt = (g.cdp == last_cdp) ? cdpt + 1 : 1;
    last_cdp = g.cdp;
    wr<int32_t>(h.raw, kOffTracl, i + 1);
    wr<int32_t>(h.raw, kOffTracr, i + 1);
    wr<int32_t>(h.raw, kOffCdp, g.cdp);
    wr<int32_t>(h.raw, kOffCdpt, cdpt);
    double dx = (double)g.gx - g.sx, dy = (double)g.gy - g.sy;
    wr<int32_t>(h.raw, kOffOffset, (int32_t)std::lround(std::sqrt(dx * dx + dy * dy)));
    wr<int16_t>(h.raw, kOffScalco, (int16_t)g.scalco);
    wr<int32_t>(h.raw, kOffSx, g.sx);
    wr<int32_t>(h.raw, kOffSy, g.sy);
    wr<int32_t>(h.raw, kOffGx, g.gx);
    wr<int32_t>(h.raw, kOffGy, g.gy);
    wr<uint16_t>(h.raw, kOffNs, (uint16_t)ns);
    wr<uint16_t>(h.raw, kOffDt, (uint16_t)dt_us);
    if (std::fwrite(h.raw, 1, kSuHeaderBytes, f) != (size_t)kSuHeaderBytes ||
        std::fwrite(samples + (size_t)i * ns, sizeof(float), ns, f) != (size_t)ns)
      return fail(SB_ERR_IO, std::string("write failed: ") + path);
  }
  return SB_OK;
}

extern "C" const char* sb_last_error(void) { return sb::g_last_error.c_str(); }
