// SU file input/output and CDP binning with the reference's semantics.
//
// Reader:  su_trace::fgettr       CMP/OpenMP/src/su_trace.cpp:37-51
// Filter:  su_gather::su_gather   CMP/OpenMP/src/su_gather.cpp:44-47 (CMP, float)
//                                 CRS/OpenMP/src/su_gather.cpp:44-47 (CRS, double)
// Order:   CMP std::sort by fold descending (su_gather.cpp:68, su_cdp.hpp:53-55)
//          CRS ascending cdp key (std::map order)
// Lists:   CRS neighbour CDPs within apm, CRS/OpenMP/src/su_gather.cpp:68-87 (test :80-81)
// Writer:  output sections, CMP/OpenMP/src/main.cpp:244-260 (only parameter
//          sections rewrite the header, :249-251)
#include "sb_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
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

namespace {

// Source of traces in file order: header in the SU (native-endian SEG-Y
// trace header) layout and the samples as FP32.  Returns 1 for a trace, 0 at
// the end of the data, or a negative status after fail().
using NextTrace = std::function<int(SuHeader&, std::vector<float>&)>;

// The reference's binning (CMP/OpenMP/src/su_gather.cpp:33-106,
// CRS/OpenMP/src/su_gather.cpp:33-90): aph filter, std::map key order, CMP
// fold-descending std::sort, CRS apm neighbour lists; common to SU and SEG-Y.
int build_su_gather(const NextTrace& next, int method, int aph, int apm, sb_su_gather** out) {
  struct Cdp {
    std::vector<SuHeader> hdr;
    std::vector<float> data;
  };
  std::map<int, Cdp> by_cdp;
  int ns = -1;
  SuHeader h;
  std::vector<float> buf;
  while (true) {
    const int got = next(h, buf);
    if (got < 0) return -got;
    if (got == 0) break;
    const int tns = (int)buf.size();
    sb_trace_geom g = geom_of(h.raw);
    if (!sb_keep_trace(&g, method, aph)) continue;
    if (ns < 0) ns = tns;
    if (tns != ns) return fail(SB_ERR_FORMAT, "traces with different ns in one file");
    Cdp& c = by_cdp[g.cdp];
    c.hdr.push_back(h);
    c.data.insert(c.data.end(), buf.begin(), buf.end());
  }
  if (by_cdp.empty()) return fail(SB_ERR_FORMAT, "no trace passed the aperture test");

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

}  // namespace

extern "C" int sb_su_load(const char* path, int method, int aph, int apm, sb_su_gather** out) {
  if (!path || !out || (method != SB_METHOD_CMP && method != SB_METHOD_CRS))
    return fail(SB_ERR_INVALID, "sb_su_load: bad arguments");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(SB_ERR_IO, std::string("cannot open ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  // su_trace fgettr (CMP/OpenMP/src/su_trace.cpp:37-54): a partial header
  // ends the read; a short sample block is "Bad input data".
  auto next = [&](SuHeader& h, std::vector<float>& buf) -> int {
    if (std::fread(h.raw, 1, kSuHeaderBytes, f) < (size_t)kSuHeaderBytes) return 0;
    const int tns = rd<uint16_t>(h.raw, kOffNs);
    buf.resize(tns);
    if (std::fread(buf.data(), sizeof(float), tns, f) != (size_t)tns) return -fail(SB_ERR_IO, "Bad input data");
    return 1;
  };
  return build_su_gather(next, method, aph, apm, out);
}

namespace {

// SEG-Y rev 1 trace header field widths, in order from byte 0 (the same
// fields, at the same offsets, as the SU header).
struct Field {
  int off, size;
};
std::vector<Field> segy_trace_fields() {
  std::vector<Field> f;
  auto run = [&](int off, int n, int size) {
    for (int i = 0; i < n; i++) f.push_back({off + i * size, size});
  };
  run(0, 7, 4);     // tracl .. cdpt
  run(28, 4, 2);    // trid, nvs, nhs, duse
  run(36, 8, 4);    // offset .. gwdep
  run(68, 2, 2);    // scalel, scalco
  run(72, 4, 4);    // sx, sy, gx, gy
  run(88, 46, 2);   // counit .. (bytes 89-180)
  run(180, 5, 4);   // cdp x/y, inline, crossline, shotpoint
  run(200, 2, 2);   // shotpoint scalar, trace value unit
  run(204, 1, 4);   // transduction constant mantissa
  run(208, 7, 2);   // exponent, units, device id, time scalar, source type, energy direction
  run(222, 1, 2);   // energy direction (last)
  run(224, 1, 4);   // source measurement mantissa
  run(228, 2, 2);   // exponent, unit
  run(232, 2, 4);   // unassigned
  return f;
}

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }
uint16_t be16(const uint8_t* p) { return (uint16_t)(((uint16_t)p[0] << 8) | p[1]); }

// IBM System/360 single precision: sign, 7-bit excess-64 base-16 exponent,
// 24-bit fraction 0.F; value = (-1)^s * F * 2^-24 * 16^(e-64).  Exact in
// double, then rounded once to FP32.
float ibm_to_float(uint32_t u) {
  const uint32_t frac = u & 0x00ffffffu;
  if (frac == 0) return (u >> 31) ? -0.0f : 0.0f;
  const int e = (int)((u >> 24) & 0x7f);
  const double v = std::ldexp((double)frac, 4 * (e - 64) - 24);
  return (float)((u >> 31) ? -v : v);
}

}  // namespace

extern "C" int sb_segy_load(const char* path, int method, int aph, int apm, sb_su_gather** out) {
  if (!path || !out || (method != SB_METHOD_CMP && method != SB_METHOD_CRS))
    return fail(SB_ERR_INVALID, "sb_segy_load: bad arguments");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(SB_ERR_IO, std::string("cannot open ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  // 3200-byte textual header, 400-byte binary header (big-endian).
  uint8_t text[3200], bin[400];
  if (std::fread(text, 1, 3200, f) != 3200 || std::fread(bin, 1, 400, f) != 400)
    return fail(SB_ERR_FORMAT, "SEG-Y: file shorter than its 3600-byte headers");
  const int bin_dt = be16(bin + 16), bin_ns = be16(bin + 20), format = be16(bin + 24);
  const int next_text = (int16_t)be16(bin + 304);  // rev 1 extended textual headers (-1: variable)
  if (next_text < 0) return fail(SB_ERR_UNSUPPORTED, "SEG-Y: variable number of extended textual headers");
  if (next_text > 0 && std::fseek(f, 3200L * next_text, SEEK_CUR) != 0)
    return fail(SB_ERR_FORMAT, "SEG-Y: extended textual headers missing");
  int bytes = 0;
  switch (format) {
    case 1: case 2: case 5: bytes = 4; break;  // IBM float, int32, IEEE float
    case 3: bytes = 2; break;                  // int16
    case 8: bytes = 1; break;                  // int8
    default: return fail(SB_ERR_UNSUPPORTED, "SEG-Y: sample format code " + std::to_string(format));
  }
  const std::vector<Field> fields = segy_trace_fields();
  std::vector<uint8_t> raw;
  uint8_t th[kSuHeaderBytes];
  auto next = [&](SuHeader& h, std::vector<float>& buf) -> int {
    const size_t got = std::fread(th, 1, kSuHeaderBytes, f);
    if (got == 0) return 0;
    if (got < (size_t)kSuHeaderBytes) return -fail(SB_ERR_IO, "SEG-Y: truncated trace header");
    for (const Field& fd : fields) {  // big-endian fields -> native layout
      if (fd.size == 4) {
        const uint32_t v = be32(th + fd.off);
        std::memcpy(h.raw + fd.off, &v, 4);
      } else {
        const uint16_t v = be16(th + fd.off);
        std::memcpy(h.raw + fd.off, &v, 2);
      }
    }
    int tns = rd<uint16_t>(h.raw, kOffNs);
    if (tns == 0) {  // per-trace ns and dt default to the binary header's
      tns = bin_ns;
      wr<uint16_t>(h.raw, kOffNs, (uint16_t)tns);
    }
    if (rd<uint16_t>(h.raw, kOffDt) == 0) wr<uint16_t>(h.raw, kOffDt, (uint16_t)bin_dt);
    raw.resize((size_t)tns * bytes);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) return -fail(SB_ERR_IO, "Bad input data");
    buf.resize(tns);
    for (int i = 0; i < tns; i++) {
      const uint8_t* p = raw.data() + (size_t)i * bytes;
      switch (format) {
        case 1: buf[i] = ibm_to_float(be32(p)); break;
        case 2: buf[i] = (float)(int32_t)be32(p); break;
        case 3: buf[i] = (float)(int16_t)be16(p); break;
        case 5: {
          const uint32_t u = be32(p);
          std::memcpy(&buf[i], &u, 4);
          break;
        }
        default: buf[i] = (float)(int8_t)p[0]; break;
      }
    }
    return 1;
  };
  return build_su_gather(next, method, aph, apm, out);
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
