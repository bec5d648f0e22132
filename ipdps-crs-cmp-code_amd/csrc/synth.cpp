// Seeded synthetic gathers for the benchmark configurations.
//
// kind 0 (CMP line): CDP i at x = i * spacing, y = 0; fold uniform in
//   [fold_min, fold_max]; half-offsets either regular (j * h_step) or spread
//   over [0, h_max] with +- h_jitter; events t^2 = t0^2 + C h^2, C = 4 / v^2.
// kind 1 (CRS, 3D-style): midpoints uniform +- mid_half_width in x and y
//   around the CDP centre, half-offsets uniform in [0, h_max] with a uniform
//   azimuth; events t^2 = (t0 + A dm)^2 + t0 B dm^2 + C h^2 with dm the
//   signed inline midpoint displacement from the CDP centre.
// Every event is a Ricker wavelet of peak frequency `freq`.  Gaussian noise
// has sigma = RMS(clean trace) / snr.  Coordinates are stored in cm
// (scalco = -100).  Random streams: SplitMix64 keyed by (seed, cdp, trace).
#include "sb_internal.hpp"

#include <cmath>
#include <cstring>
#include <vector>

using namespace sb;

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

uint64_t mix(uint64_t a, uint64_t b) {
  Rng r(a ^ (b * 0xD6E8FEB86659FD93ull));
  return r.next();
}

struct Event {
  double t0, a, b, c, amp;
};

int ricker_halfwidth_samples(double freq, double dt) {
  // exp(-(pi f t)^2) < 1e-12 beyond |t| = 1.7 / f.
  return (int)std::ceil(1.7 / freq / dt) + 1;
}

void add_ricker(float* tr, int ns, double dt, double t, double amp, double freq, int hw) {
  int c = (int)std::floor(t / dt);
  int lo = c - hw, hi = c + hw;
  if (lo < 0) lo = 0;
  if (hi > ns - 1) hi = ns - 1;
  const double pf2 = M_PI * M_PI * freq * freq;
  for (int n = lo; n <= hi; n++) {
    double tau = n * dt - t;
    double a = pf2 * tau * tau;
    tr[n] += (float)(amp * (1.0 - 2.0 * a) * std::exp(-a));
  }
}

int32_t cm(double meters) { return (int32_t)std::lround(meters * 100.0); }

// Geometry of trace j of CDP i (deterministic; used by both passes).
sb_trace_geom trace_geom(const sb_synth_spec& s, int i, int j, int fold, Rng& r) {
  sb_trace_geom g;
  g.cdp = i + 1;
  g.scalco = -100;
  const double xc = i * s.cdp_spacing;
  if (s.kind == 0) {
    double h;
    if (s.regular_offsets) {
      h = j * s.h_step;
    } else {
      h = fold > 1 ? s.h_max * j / (fold - 1) : 0.0;
    }
    if (s.h_jitter > 0) h += r.uniform(-s.h_jitter, s.h_jitter);
    h = std::fabs(h);
    g.sx = cm(xc - h);
    g.gx = cm(xc + h);
    g.sy = 0;
    g.gy = 0;
  } else {
    double mx = xc + r.uniform(-s.mid_half_width, s.mid_half_width);
    double my = r.uniform(-s.mid_half_width, s.mid_half_width);
    double h = r.uniform(0.0, s.h_max);
    double phi = r.uniform(0.0, 6.283185307179586);
    double hx = h * std::cos(phi), hy = h * std::sin(phi);
    g.sx = cm(mx - hx);
    g.gx = cm(mx + hx);
    g.sy = cm(my - hy);
    g.gy = cm(my + hy);
  }
  return g;
}

int fold_of(const sb_synth_spec& s, int i) {
  Rng r(mix(s.seed, 0x1000000ull + (uint64_t)i));
  int span = s.fold_max - s.fold_min + 1;
  return s.fold_min + (int)(r.next() % (uint64_t)span);
}

}  // namespace

extern "C" int sb_synth(const sb_synth_spec* spec, int* ncdp_out, int* ntr_out, int32_t* cdp_off,
                        sb_trace_geom* geom, float* samples) {
  if (!spec || !ncdp_out || !ntr_out) return fail(SB_ERR_INVALID, "sb_synth: bad arguments");
  const sb_synth_spec& s = *spec;
  if (s.ncdp <= 0 || s.fold_min <= 0 || s.fold_max < s.fold_min || s.ns <= 1 || s.ns > 65535 || s.dt_us <= 0 ||
      s.dt_us > 65535 || s.nevents < 0 || (s.kind != 0 && s.kind != 1))
    return fail(SB_ERR_INVALID, "sb_synth: bad spec");
  const int method = s.kind == 0 ? SB_METHOD_CMP : SB_METHOD_CRS;

  // Pass 1: kept trace count per CDP (geometry only).
  std::vector<int> kept(s.ncdp);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < s.ncdp; i++) {
    int fold = fold_of(s, i), k = 0;
    for (int j = 0; j < fold; j++) {
      Rng r(mix(s.seed, ((uint64_t)i << 24) ^ (uint64_t)j));
      sb_trace_geom g = trace_geom(s, i, j, fold, r);
      if (s.aph <= 0 || sb_keep_trace(&g, method, s.aph)) k++;
    }
    kept[i] = k;
  }
  std::vector<int64_t> off(s.ncdp + 1, 0);
  for (int i = 0; i < s.ncdp; i++) off[i + 1] = off[i] + kept[i];
  if (off[s.ncdp] > INT32_MAX) return fail(SB_ERR_INVALID, "sb_synth: too many traces");
  *ncdp_out = s.ncdp;
  *ntr_out = (int)off[s.ncdp];
  if (!samples && !geom && !cdp_off) return SB_OK;
  if (!samples || !geom || !cdp_off) return fail(SB_ERR_INVALID, "sb_synth: need all output arrays");
  for (int i = 0; i <= s.ncdp; i++) cdp_off[i] = (int32_t)off[i];

  // Events are shared by the whole line.
  std::vector<Event> ev(s.nevents);
  {
    Rng r(mix(s.seed, 0xE7E47ull));
    for (auto& e : ev) {
      if (s.kind == 0) {
        e.t0 = r.uniform(0.2, 1.8);
        double v = r.uniform(1500.0, 4500.0);
        e.c = 4.0 / (v * v);
        e.a = 0;
        e.b = 0;
      } else {
        double T = (s.ns - 1) * s.dt_us * 1e-6;
        e.t0 = r.uniform(0.1 * T, 0.6 * T);
        e.a = r.uniform(-0.7e-3, 0.7e-3);
        e.b = r.uniform(-1e-7, 1e-7);
        e.c = r.uniform(1.975e-7, 1.384e-6);
      }
      e.amp = r.uniform(0.5, 1.0);
    }
  }
  const double dt = s.dt_us * 1e-6;
  const int hw = ricker_halfwidth_samples(s.freq > 0 ? s.freq : 25.0, dt);

#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < s.ncdp; i++) {
    int fold = fold_of(s, i);
    int64_t t = off[i];
    const double xc = i * s.cdp_spacing;
    for (int j = 0; j < fold; j++) {
      Rng r(mix(s.seed, ((uint64_t)i << 24) ^ (uint64_t)j));
      sb_trace_geom g = trace_geom(s, i, j, fold, r);
      if (s.aph > 0 && !sb_keep_trace(&g, method, s.aph)) continue;
      geom[t] = g;
      float* tr = samples + t * (int64_t)s.ns;
      std::memset(tr, 0, sizeof(float) * s.ns);
      double dx = (g.gx - g.sx) * 0.01, dy = (g.gy - g.sy) * 0.01;
      double h2 = 0.25 * (dx * dx + dy * dy);
      double dm = 0.5 * (g.gx + g.sx) * 0.01 - xc;
      for (const Event& e : ev) {
        double tt;
        if (s.kind == 0) {
          tt = e.t0 * e.t0 + e.c * h2;
        } else {
          double q = e.t0 + e.a * dm;
          tt = q * q + e.t0 * e.b * dm * dm + e.c * h2;
        }
        if (tt <= 0) continue;
        add_ricker(tr, s.ns, dt, std::sqrt(tt), e.amp, s.freq > 0 ? s.freq : 25.0, hw);
      }
      if (s.snr > 0) {
        double ss = 0;
        for (int n = 0; n < s.ns; n++) ss += (double)tr[n] * tr[n];
        double sigma = std::sqrt(ss / s.ns) / s.snr;
        if (sigma == 0) sigma = 1.0 / s.snr;
        Rng nr(mix(s.seed ^ 0x5EED5EEDull, ((uint64_t)i << 24) ^ (uint64_t)j));
        for (int n = 0; n < s.ns; n++) tr[n] += (float)(sigma * nr.normal());
      }
      t++;
    }
  }
  return SB_OK;
}
