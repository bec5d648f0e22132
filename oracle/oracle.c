/*
 * oracle.c -- CPU restatement of the reference's semblance search.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with gcc
 * -ffp-contract=off so every float/double operation rounds exactly as the
 * reference's x86-64 SSE code (g++ -O3, no FMA on baseline x86-64).
 *
 * Each routine cites the reference lines it restates.
 */
#define _DEFAULT_SOURCE /* M_PI */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS 1e-13
#define FACTOR 1e6
#define MAXW 64

/* ------------------------------------------------------------------ CMP */
/* CMP/OpenMP/src/main.cpp:68-110 (parameter setup, float), :119-136 (grid,
 * half-offsets), :149-204 (semblance), :207-222 (arg-max), :253 (C output). */
int orc_cmp(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
            double c0d, double c1d, int nc, double taud, int32_t* index, float* pc, float* coher, float* stack,
            float* panel_sem, float* panel_stk) {
  const float c0 = (float)((double)(float)c0d * FACTOR);
  const float c1 = (float)((double)(float)c1d * FACTOR);
  const float itau = (float)taud;
  const float inc = (c1 - c0) * (1.0f / (float)nc);
  float dt = (float)dt_us;
  dt = dt / 1000000.0f;
  const float idt = 1.0f / dt;
  const int tau = ((int)(itau * idt) > 0) ? ((int)(itau * idt)) : 0;
  const float w = (float)((2 * tau) + 1);
  const int wi = 2 * tau + 1;
  if (wi > MAXW) return 1;
  const int ntr = cdp_off[ncdp];

  float* c = (float*)malloc(sizeof(float) * nc);
  float* h = (float*)malloc(sizeof(float) * (ntr > 0 ? ntr : 1));
  for (int i = 0; i < nc; i++) c[i] = c0 + inc * (float)i;
  for (int i = 0; i < ntr; i++) {
    float s = (float)g[i].scalco;
    if (-EPS < s && s < EPS)
      s = 1.0f;
    else if (s < 0)
      s = 1.0f / s;
    float hx = ((float)g[i].gx - (float)g[i].sx) * s;
    float hy = ((float)g[i].gy - (float)g[i].sy) * s;
    h[i] = (float)(0.25 * (double)(hx * hx + hy * hy) / FACTOR);
  }

  for (int cdp = 0; cdp < ncdp; cdp++) {
    const int t_id0 = cdp_off[cdp], t_idf = cdp_off[cdp + 1];
#pragma omp parallel for schedule(static)
    for (int t0 = 0; t0 < ns; t0++) {
      float max_sem = 0.0f, best_stk = 0.0f;
      int max_c = -1;
      for (int c_id = 0; c_id < nc; c_id++) {
        float num[MAXW], ac_sq = 0, ac_lin = 0, den = 0, m = 0;
        int err = 0;
        const float cc = c[c_id];
        float t0s = dt * (float)t0;
        t0s = t0s * t0s;
        for (int j = 0; j < wi; j++) num[j] = 0.0f;
        for (int t_id = t_id0; t_id < t_idf; t_id++) {
          float t = sqrtf(t0s + cc * h[t_id]) * idt;
          /* (int)t with the reference's bounds test; t >= 0 or NaN here. */
          int ok = (t >= (float)tau) && (t < (float)(ns - tau - 1));
          if (ok) {
            int it = (int)t;
            float x = t - (float)it;
            const float* sp = samples + (size_t)t_id * ns + (it - tau);
            float s1 = sp[0], s0;
            for (int j = 0; j < wi; j++) {
              s0 = s1;
              s1 = sp[j + 1];
              float v = (s1 - s0) * x + s0;
              num[j] += v;
              den += v * v;
              ac_lin += v;
            }
            m++;
          } else {
            err++;
          }
        }
        for (int j = 0; j < wi; j++) ac_sq += num[j] * num[j];
        float sem, stk;
        if (den > EPS && m > EPS && w > EPS && err < 2) {
          sem = ac_sq / (den * m);
          stk = ac_lin / (w * m);
        } else {
          sem = -1.0f;
          stk = -1.0f;
        }
        if (panel_sem) panel_sem[((size_t)cdp * ns + t0) * nc + c_id] = sem;
        if (panel_stk) panel_stk[((size_t)cdp * ns + t0) * nc + c_id] = stk;
        if (sem > max_sem) {
          max_sem = sem;
          max_c = c_id;
          best_stk = stk;
        }
      }
      const size_t o = (size_t)cdp * ns + t0;
      index[o] = max_c;
      coher[o] = max_sem;
      stack[o] = max_c > -1 ? best_stk : 0.0f;
      pc[o] = max_c < 0 ? 0.0f : (float)((double)(c0 + inc * (float)max_c) / FACTOR);
    }
  }
  free(c);
  free(h);
  return 0;
}

/* ------------------------------------------------------------------ CRS */
/* Per-trace midpoints and half-offsets, CRS/OpenMP/src/main.cpp:163-177. */
static void crs_geometry(int ntr, const orc_geom* g, double* m0x, double* m0y, double* h0) {
  for (int i = 0; i < ntr; i++) {
    double s = (double)g[i].scalco;
    if (-EPS < s && s < EPS)
      s = 1.0;
    else if (s < 0)
      s = 1.0f / s;
    m0x[i] = ((double)g[i].gx + (double)g[i].sx) * s * 0.5;
    m0y[i] = ((double)g[i].gy + (double)g[i].sy) * s * 0.5;
    double hx = ((double)g[i].gx - (double)g[i].sx) * s;
    double hy = ((double)g[i].gy - (double)g[i].sy) * s;
    h0[i] = 0.25 * (hx * hx + hy * hy) / FACTOR;
  }
}

/* Supergather of one CDP: neighbour CDPs in list order, traces in gather
 * order, distances to the CDP's first trace (CRS/OpenMP/src/main.cpp:179-203,
 * 287-316).  With dir = {sin, cos} of an azimuth (the library's
 * SB_OPT_DM_AZIMUTH extension; the reference has none) m is the signed
 * projection dx sin + dy cos and m2 = m m.  Returns the size. */
static int crs_supergather_dir(int cdp, const int32_t* cdp_off, const int32_t* nbr_off, const int32_t* nbr,
                               const double* m0x, const double* m0y, const double* h0, const double* dir,
                               const double* sgn, int* rows, double* m2, double* m, double* h) {
  const double rx = m0x[cdp_off[cdp]], ry = m0y[cdp_off[cdp]];
  int size = 0;
  for (int q = nbr_off[cdp]; q < nbr_off[cdp + 1]; q++) {
    const int j = nbr[q];
    for (int r = cdp_off[j]; r < cdp_off[j + 1]; r++) {
      double dx = m0x[r] - rx;
      double dy = m0y[r] - ry;
      rows[size] = r;
      if (dir) {
        /* true coordinates: the reference's 1/scalco factor is negative for
           a negative scalco (CRS/OpenMP/src/main.cpp:166-171) */
        const double sr = sgn[cdp_off[cdp]], st = sgn[r];
        dx = m0x[r] * st - rx * sr;
        dy = m0y[r] * st - ry * sr;
        double px = dx * dir[0];
        double py = dy * dir[1];
        double p = px + py;
        m[size] = p;
        m2[size] = p * p;
      } else {
        double mm2 = dx * dx + dy * dy;
        m2[size] = mm2;
        m[size] = sqrt(mm2);
      }
      h[size] = h0[r];
      size++;
    }
  }
  return size;
}

static int crs_supergather(int cdp, const int32_t* cdp_off, const int32_t* nbr_off, const int32_t* nbr,
                           const double* m0x, const double* m0y, const double* h0, int* rows, double* m2,
                           double* m, double* h) {
  return crs_supergather_dir(cdp, cdp_off, nbr_off, nbr, m0x, m0y, h0, NULL, NULL, rows, m2, m, h);
}

static int max_supergather(int ncdp, const int32_t* cdp_off, const int32_t* nbr_off, const int32_t* nbr) {
  int best = 1;
  for (int i = 0; i < ncdp; i++) {
    int s = 0;
    for (int q = nbr_off[i]; q < nbr_off[i + 1]; q++) s += cdp_off[nbr[q] + 1] - cdp_off[nbr[q]];
    if (s > best) best = s;
  }
  return best;
}

/* One (t0, par) of CRS in double, CRS/OpenMP/src/main.cpp:211-265. */
static void crs_eval(double a, double b, double c, double dt, double idt, int tau, int w, int ns, int size,
                     const int* rows, const double* m2, const double* m, const double* h, const float* samples,
                     int t0, double* sem, double* stk) {
  double num[MAXW], ac_sq = 0, ac_lin = 0, den = 0, mm = 0;
  int err = 0;
  const double t0s = dt * t0;
  for (int j = 0; j < w; j++) num[j] = 0.0;
  for (int k = 0; k < size; k++) {
    double t = t0s + a * m[k];
    t = t * t + b * m2[k] + c * h[k];
    t = t < 0.0 ? -1 : sqrt(t);
    const double ti = t * idt;
    if (ti >= (double)tau && ti < (double)(ns - tau - 1)) {
      int it = (int)ti;
      double x = ti - (double)it;
      const float* sp = samples + (size_t)rows[k] * ns + (it - tau);
      double s1 = sp[0], s0;
      for (int j = 0; j < w; j++) {
        s0 = s1;
        s1 = sp[j + 1];
        double v = (s1 - s0) * x + s0;
        num[j] += v;
        den += v * v;
        ac_lin += v;
      }
      mm++;
    } else {
      err++;
    }
  }
  for (int j = 0; j < w; j++) ac_sq += num[j] * num[j];
  if (den > EPS && mm > EPS && w > EPS && err < 2) {
    *sem = ac_sq / (den * mm);
    *stk = ac_lin / (w * mm);
  } else {
    *sem = 0.0;
    *stk = 0.0;
  }
}

int orc_crs(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
            const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, int na, double b0, double b1, int nb,
            double c0, double c1, int nc, double taud, int32_t* index, float* pa, float* pb, float* pc,
            float* coher, float* stack, float* panel_sem, float* panel_stk) {
  return orc_crs_range(ncdp, cdp_off, ns, dt_us, g, samples, nbr_off, nbr, a0, a1, na, b0, b1, nb, c0, c1, nc, taud,
                       0, ncdp, index, pa, pb, pc, coher, stack, panel_sem, panel_stk);
}

int orc_crs_range(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                  const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, int na, double b0, double b1,
                  int nb, double c0, double c1, int nc, double taud, int cdp_begin, int cdp_end, int32_t* index,
                  float* pa, float* pb, float* pc, float* coher, float* stack, float* panel_sem, float* panel_stk) {
  return orc_crs_range_az(ncdp, cdp_off, ns, dt_us, g, samples, nbr_off, nbr, a0, a1, na, b0, b1, nb, c0, c1, nc,
                          taud, -1, cdp_begin, cdp_end, index, pa, pb, pc, coher, stack, panel_sem, panel_stk);
}

int orc_crs_range_az(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                     const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, int na, double b0, double b1,
                     int nb, double c0, double c1, int nc, double taud, int az_mdeg, int cdp_begin, int cdp_end,
                     int32_t* index, float* pa, float* pb, float* pc, float* coher, float* stack, float* panel_sem,
                     float* panel_stk) {
  if (cdp_begin < 0 || cdp_end > ncdp || cdp_begin >= cdp_end) return 1;
  double dirv[2];
  const double* dir = NULL;
  if (az_mdeg >= 0) {
    const double th = (double)az_mdeg / 1000.0 * (M_PI / 180.0);
    dirv[0] = sin(th);
    dirv[1] = cos(th);
    dir = dirv;
  }
  c0 = c0 * FACTOR;
  c1 = c1 * FACTOR;
  const double inc_a = (a1 - a0) * (1.0 / (double)na);
  const double inc_b = (b1 - b0) * (1.0 / (double)nb);
  const double inc_c = (c1 - c0) * (1.0 / (double)nc);
  const int npar = na * nb * nc;
  double dt = (double)dt_us;
  dt = dt / 1000000.0f;
  const double idt = 1.0f / dt;
  const int tau = (int)(taud * idt) > 0 ? (int)(taud * idt) : 0;
  const int w = (2 * tau) + 1;
  if (w > MAXW) return 1;
  const int ntr = cdp_off[ncdp];
  double* m0x = (double*)malloc(sizeof(double) * ntr);
  double* m0y = (double*)malloc(sizeof(double) * ntr);
  double* h0 = (double*)malloc(sizeof(double) * ntr);
  double* sgn = (double*)malloc(sizeof(double) * (ntr > 0 ? ntr : 1));
  crs_geometry(ntr, g, m0x, m0y, h0);
  for (int i = 0; i < ntr; i++) sgn[i] = (g[i].scalco <= -1) ? -1.0 : 1.0; /* the factor 1/scalco is negative */
  const int maxs = max_supergather(ncdp, cdp_off, nbr_off, nbr);
  int* rows = (int*)malloc(sizeof(int) * maxs);
  double* m2 = (double*)malloc(sizeof(double) * maxs);
  double* m = (double*)malloc(sizeof(double) * maxs);
  double* h = (double*)malloc(sizeof(double) * maxs);
  double* pars = (double*)malloc(sizeof(double) * 3 * npar);
  for (int i = 0; i < npar; i++) {
    pars[3 * i] = a0 + inc_a * (i / (nc * nb));
    pars[3 * i + 1] = b0 + inc_b * ((i / nc) % nb);
    pars[3 * i + 2] = c0 + inc_c * (i % nc);
  }
  for (int cdp = cdp_begin; cdp < cdp_end; cdp++) {
    const int size = crs_supergather_dir(cdp, cdp_off, nbr_off, nbr, m0x, m0y, h0, dir, sgn, rows, m2, m, h);
    const size_t row0 = (size_t)(cdp - cdp_begin) * ns;
    double stt00 = 0, sem00 = 0;
    crs_eval(pars[0], pars[1], pars[2], dt, idt, tau, w, ns, size, rows, m2, m, h, samples, 0, &sem00, &stt00);
#pragma omp parallel for schedule(static)
    for (int t0 = 0; t0 < ns; t0++) {
      double max_sem = 0.0, best_stk = 0.0;
      int max_par = -1;
      for (int p = 0; p < npar; p++) {
        double sem, stk;
        crs_eval(pars[3 * p], pars[3 * p + 1], pars[3 * p + 2], dt, idt, tau, w, ns, size, rows, m2, m, h, samples,
                 t0, &sem, &stk);
        if (panel_sem) panel_sem[(row0 + t0) * npar + p] = (float)sem;
        if (panel_stk) panel_stk[(row0 + t0) * npar + p] = (float)stk;
        if (sem > max_sem) {
          max_sem = sem;
          max_par = p;
          best_stk = stk;
        }
      }
      /* No update: absolute index 0 = (t0 = 0, par 0), CRS/OpenMP/src/main.cpp:273,284. */
      if (max_par < 0) {
        max_par = 0;
        best_stk = stt00;
      }
      const size_t o = row0 + t0;
      index[o] = max_par;
      coher[o] = (float)max_sem;
      stack[o] = (float)best_stk;
      pa[o] = (float)(a0 + inc_a * (max_par / (nc * nb)));
      pb[o] = (float)(b0 + inc_b * ((max_par / nc) % nb));
      pc[o] = (float)((c0 + inc_c * (max_par % nc)) / FACTOR);
    }
  }
  free(m0x);
  free(m0y);
  free(h0);
  free(rows);
  free(m2);
  free(m);
  free(h);
  free(pars);
  free(sgn);
  return 0;
}

int orc_crs_eval_f64(int cdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                     const int32_t* nbr_off, const int32_t* nbr, double a, double b, double c, double taud, int t0,
                     double* sem, double* stk) {
  /* Needs geometry of every trace a neighbour list can reach. */
  int ncdp = 0;
  for (int q = nbr_off[cdp]; q < nbr_off[cdp + 1]; q++)
    if (nbr[q] + 1 > ncdp) ncdp = nbr[q] + 1;
  if (cdp + 1 > ncdp) ncdp = cdp + 1;
  const int ntr = cdp_off[ncdp];
  double dt = (double)dt_us / 1e6;
  const double idt = 1.0 / dt;
  const int tau = (int)(taud * idt) > 0 ? (int)(taud * idt) : 0;
  const int w = 2 * tau + 1;
  double* m0x = (double*)malloc(sizeof(double) * ntr);
  double* m0y = (double*)malloc(sizeof(double) * ntr);
  double* h0 = (double*)malloc(sizeof(double) * ntr);
  crs_geometry(ntr, g, m0x, m0y, h0);
  int maxs = 0;
  for (int q = nbr_off[cdp]; q < nbr_off[cdp + 1]; q++) maxs += cdp_off[nbr[q] + 1] - cdp_off[nbr[q]];
  if (maxs < 1) maxs = 1;
  int* rows = (int*)malloc(sizeof(int) * maxs);
  double* m2 = (double*)malloc(sizeof(double) * maxs);
  double* m = (double*)malloc(sizeof(double) * maxs);
  double* h = (double*)malloc(sizeof(double) * maxs);
  const int size = crs_supergather(cdp, cdp_off, nbr_off, nbr, m0x, m0y, h0, rows, m2, m, h);
  crs_eval(a, b, c * FACTOR, dt, idt, tau, w, ns, size, rows, m2, m, h, samples, t0, sem, stk);
  free(m0x);
  free(m0y);
  free(h0);
  free(rows);
  free(m2);
  free(m);
  free(h);
  return 0;
}

/* ------------------------------------------------------------- CMP FP64 */
int orc_cmp_f64(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                double c0d, double c1d, int nc, double taud, int32_t* index, double* coher, double* stack,
                double* panel_sem) {
  const double c0 = c0d * FACTOR, c1 = c1d * FACTOR;
  const double inc = (c1 - c0) / (double)nc;
  const double dt = (double)dt_us / 1e6;
  const double idt = 1.0 / dt;
  /* The window half-width as the float reference computes it. */
  const float dtf = (float)dt_us / 1000000.0f;
  const float idtf = 1.0f / dtf;
  const int tau = ((int)((float)taud * idtf) > 0) ? ((int)((float)taud * idtf)) : 0;
  const int w = 2 * tau + 1;
  if (w > MAXW) return 1;
  const int ntr = cdp_off[ncdp];
  double* h = (double*)malloc(sizeof(double) * (ntr > 0 ? ntr : 1));
  for (int i = 0; i < ntr; i++) {
    double s = (double)g[i].scalco;
    if (s == 0)
      s = 1.0;
    else if (s < 0)
      s = 1.0 / s;
    double hx = ((double)g[i].gx - (double)g[i].sx) * s;
    double hy = ((double)g[i].gy - (double)g[i].sy) * s;
    h[i] = 0.25 * (hx * hx + hy * hy) / FACTOR;
  }
  for (int cdp = 0; cdp < ncdp; cdp++) {
    const int r0 = cdp_off[cdp], r1 = cdp_off[cdp + 1];
#pragma omp parallel for schedule(static)
    for (int t0 = 0; t0 < ns; t0++) {
      double max_sem = 0.0, best_stk = 0.0;
      int max_c = -1;
      for (int ci = 0; ci < nc; ci++) {
        double num[MAXW], den = 0, lin = 0, m = 0, acsq = 0;
        int err = 0;
        const double c = c0 + inc * ci;
        const double t0s = dt * t0;
        for (int j = 0; j < w; j++) num[j] = 0;
        for (int r = r0; r < r1; r++) {
          const double t = sqrt(t0s * t0s + c * h[r]) * idt;
          if (t >= tau && t < ns - tau - 1) {
            const int it = (int)t;
            const double x = t - it;
            const float* sp = samples + (size_t)r * ns + (it - tau);
            for (int j = 0; j < w; j++) {
              const double v = ((double)sp[j + 1] - sp[j]) * x + sp[j];
              num[j] += v;
              den += v * v;
              lin += v;
            }
            m++;
          } else {
            err++;
          }
        }
        for (int j = 0; j < w; j++) acsq += num[j] * num[j];
        double sem = -1, stk = -1;
        if (den > EPS && m > 0 && err < 2) {
          sem = acsq / (den * m);
          stk = lin / (w * m);
        }
        if (panel_sem) panel_sem[((size_t)cdp * ns + t0) * nc + ci] = sem;
        if (sem > max_sem) {
          max_sem = sem;
          max_c = ci;
          best_stk = stk;
        }
      }
      const size_t o = (size_t)cdp * ns + t0;
      index[o] = max_c;
      coher[o] = max_sem;
      stack[o] = max_c > -1 ? best_stk : 0.0;
    }
  }
  free(h);
  return 0;
}

/* --------------------------------------------------------------- CRS-DE */
/* DE/rand/1/bin (Storn & Price 1997) over the box [a0,a1]x[b0,b1]x[c0,c1]
 * per (CDP, t0); fitness = the CRS semblance (CRS/OpenMP/src/main.cpp:218-265
 * in float, the precision of CRS/CUDA/include/utils.hpp:33).  Random draws
 * come from a counter-based hash of (seed, cdp index, t0, generation, member,
 * draw), so the CPU and GPU streams are identical. */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint32_t orc_de_hash(uint64_t seed, uint32_t cdp, uint32_t t0, uint32_t gen, uint32_t member, uint32_t draw) {
  uint64_t k = mix64(seed ^ (0x9E3779B97F4A7C15ull * ((uint64_t)cdp + 1)));
  k = mix64(k ^ (((uint64_t)t0 << 32) | gen));
  k = mix64(k ^ (((uint64_t)member << 32) | draw));
  return (uint32_t)(k >> 32);
}

static float unit(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

/* Index among n - |excl| candidates mapped past the excluded members. */
static int pick(uint32_t h, int n, const int* excl, int nex) {
  int v = (int)(h % (uint32_t)(n - nex));
  int e[3];
  for (int i = 0; i < nex; i++) e[i] = excl[i];
  for (int i = 0; i < nex; i++)
    for (int j = i + 1; j < nex; j++)
      if (e[j] < e[i]) {
        int t = e[i];
        e[i] = e[j];
        e[j] = t;
      }
  for (int i = 0; i < nex; i++)
    if (v >= e[i]) v++;
  return v;
}

/* FP32 fitness in the reference operation order (no contraction). */
static void de_eval(float a, float b, float c, float dt, float idt, int tau, int w, int ns, int size, const int* rows,
                    const float* m2, const float* m, const float* h, const float* samples, int t0, float* sem,
                    float* stk) {
  float num[MAXW], ac_sq = 0, ac_lin = 0, den = 0, mm = 0;
  int err = 0;
  const float t0s = dt * (float)t0;
  for (int j = 0; j < w; j++) num[j] = 0.0f;
  for (int k = 0; k < size; k++) {
    float t = t0s + a * m[k];
    t = t * t + b * m2[k] + c * h[k];
    t = t < 0.0f ? -1.0f : sqrtf(t);
    const float ti = t * idt;
    if (ti >= (float)tau && ti < (float)(ns - tau - 1)) {
      int it = (int)ti;
      float x = ti - (float)it;
      const float* sp = samples + (size_t)rows[k] * ns + (it - tau);
      float s1 = sp[0], s0;
      for (int j = 0; j < w; j++) {
        s0 = s1;
        s1 = sp[j + 1];
        float v = (s1 - s0) * x + s0;
        num[j] += v;
        den += v * v;
        ac_lin += v;
      }
      mm++;
    } else {
      err++;
    }
  }
  for (int j = 0; j < w; j++) ac_sq += num[j] * num[j];
  if (den > EPS && mm > EPS && err < 2) {
    *sem = ac_sq / (den * mm);
    *stk = ac_lin / ((float)w * mm);
  } else {
    *sem = 0.0f;
    *stk = 0.0f;
  }
}

/* One (CDP, t0) population run.  hist (optional) records the trajectory as
 * orc_de_history documents it (oracle.h). */
typedef struct {
  float* pop_x;   /* [(ngen+1)][pop][3] */
  float* pop_f;   /* [(ngen+1)][pop] */
  float* trial_x; /* [ngen][pop][3] */
  float* trial_f; /* [ngen][pop] */
  int32_t* picks; /* [ngen][pop][4] */
} de_hist;

typedef struct {
  const float *lo, *hi;
  float F;
  uint32_t cr_thr;
  uint64_t seed;
  int pop, ngen;
  float dt, idt;
  int tau, w, ns, size;
  const int* rows;
  const float *m2, *m, *h, *samples;
} de_setup;

static void de_point(const de_setup* S, int cdp, int t0, float* best_x, float* best_f, float* best_s,
                     const de_hist* hist) {
  float x[64][3], fit[64], stk[64], tr[64][3], tf[64], ts[64];
  const int pop = S->pop;
  const float* lo = S->lo;
  const float* hi = S->hi;
  const uint64_t seed = S->seed;
#define DE_EVAL(A, B, Cc, F_, S_)                                                                          \
  de_eval((A), (B), (Cc), S->dt, S->idt, S->tau, S->w, S->ns, S->size, S->rows, S->m2, S->m, S->h, S->samples, \
          t0, (F_), (S_))
  for (int i = 0; i < pop; i++) {
    for (int d = 0; d < 3; d++) x[i][d] = lo[d] + unit(orc_de_hash(seed, cdp, t0, 0, i, d)) * (hi[d] - lo[d]);
    DE_EVAL(x[i][0], x[i][1], x[i][2], &fit[i], &stk[i]);
  }
  for (int gen = 1; gen <= S->ngen; gen++) {
    if (hist)
      for (int i = 0; i < pop; i++) {
        for (int d = 0; d < 3; d++) hist->pop_x[((size_t)(gen - 1) * pop + i) * 3 + d] = x[i][d];
        hist->pop_f[(size_t)(gen - 1) * pop + i] = fit[i];
      }
    for (int i = 0; i < pop; i++) {
      int ex[3];
      ex[0] = i;
      const int r1 = pick(orc_de_hash(seed, cdp, t0, gen, i, 0), pop, ex, 1);
      ex[1] = r1;
      const int r2 = pick(orc_de_hash(seed, cdp, t0, gen, i, 1), pop, ex, 2);
      ex[2] = r2;
      const int r3 = pick(orc_de_hash(seed, cdp, t0, gen, i, 2), pop, ex, 3);
      const int jr = (int)(orc_de_hash(seed, cdp, t0, gen, i, 3) % 3u);
      for (int d = 0; d < 3; d++) {
        const uint32_t u = orc_de_hash(seed, cdp, t0, gen, i, 4 + d) >> 8;
        float v;
        if (u < S->cr_thr || d == jr) {
          float diff = x[r2][d] - x[r3][d];
          float fd = S->F * diff;
          v = x[r1][d] + fd;
          v = fminf(fmaxf(v, lo[d]), hi[d]);
        } else {
          v = x[i][d];
        }
        tr[i][d] = v;
      }
      DE_EVAL(tr[i][0], tr[i][1], tr[i][2], &tf[i], &ts[i]);
      if (hist) {
        int32_t* pk = hist->picks + ((size_t)(gen - 1) * pop + i) * 4;
        pk[0] = r1;
        pk[1] = r2;
        pk[2] = r3;
        pk[3] = jr;
        for (int d = 0; d < 3; d++) hist->trial_x[((size_t)(gen - 1) * pop + i) * 3 + d] = tr[i][d];
        hist->trial_f[(size_t)(gen - 1) * pop + i] = tf[i];
      }
    }
    for (int i = 0; i < pop; i++)
      if (tf[i] >= fit[i]) {
        x[i][0] = tr[i][0];
        x[i][1] = tr[i][1];
        x[i][2] = tr[i][2];
        fit[i] = tf[i];
        stk[i] = ts[i];
      }
  }
#undef DE_EVAL
  if (hist)
    for (int i = 0; i < pop; i++) {
      for (int d = 0; d < 3; d++) hist->pop_x[((size_t)S->ngen * pop + i) * 3 + d] = x[i][d];
      hist->pop_f[(size_t)S->ngen * pop + i] = fit[i];
    }
  int best = 0;
  for (int i = 1; i < pop; i++)
    if (fit[i] > fit[best]) best = i;
  best_x[0] = x[best][0];
  best_x[1] = x[best][1];
  best_x[2] = x[best][2];
  *best_f = fit[best];
  *best_s = stk[best];
}

/* Shared set-up of orc_crs_de and orc_de_history: box, constants, geometry. */
typedef struct {
  float lo[3], hi[3];
  double *m0x, *m0y, *h0, *m2d, *md, *hd;
  int* rows;
  float *m2, *m, *h;
  de_setup S;
} de_work;

static int de_open(de_work* W, int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g,
                   const float* samples, const int32_t* nbr_off, const int32_t* nbr, double a0d, double a1d,
                   double b0d, double b1d, double c0d, double c1d, double taud, int pop, int ngen, double Fd,
                   double CRd, uint64_t seed) {
  if (pop < 4 || pop > 64 || ngen < 0) return 1;
  W->lo[0] = (float)a0d;
  W->lo[1] = (float)b0d;
  W->lo[2] = (float)(c0d * FACTOR);
  W->hi[0] = (float)a1d;
  W->hi[1] = (float)b1d;
  W->hi[2] = (float)(c1d * FACTOR);
  de_setup* S = &W->S;
  S->lo = W->lo;
  S->hi = W->hi;
  S->F = (float)Fd;
  S->cr_thr = (uint32_t)(CRd * 16777216.0);
  S->seed = seed;
  S->pop = pop;
  S->ngen = ngen;
  float dt = (float)dt_us;
  dt = dt / 1000000.0f;
  S->dt = dt;
  S->idt = 1.0f / dt;
  /* Window half-width as the CRS reference computes it (double). */
  const double dtd = (double)dt_us / 1000000.0;
  const double idtd = 1.0 / dtd;
  S->tau = (int)(taud * idtd) > 0 ? (int)(taud * idtd) : 0;
  S->w = 2 * S->tau + 1;
  if (S->w > MAXW) return 1;
  S->ns = ns;
  S->samples = samples;
  const int ntr = cdp_off[ncdp];
  W->m0x = (double*)malloc(sizeof(double) * ntr);
  W->m0y = (double*)malloc(sizeof(double) * ntr);
  W->h0 = (double*)malloc(sizeof(double) * ntr);
  crs_geometry(ntr, g, W->m0x, W->m0y, W->h0);
  const int maxs = max_supergather(ncdp, cdp_off, nbr_off, nbr);
  W->rows = (int*)malloc(sizeof(int) * maxs);
  W->m2d = (double*)malloc(sizeof(double) * maxs);
  W->md = (double*)malloc(sizeof(double) * maxs);
  W->hd = (double*)malloc(sizeof(double) * maxs);
  W->m2 = (float*)malloc(sizeof(float) * maxs);
  W->m = (float*)malloc(sizeof(float) * maxs);
  W->h = (float*)malloc(sizeof(float) * maxs);
  S->rows = W->rows;
  S->m2 = W->m2;
  S->m = W->m;
  S->h = W->h;
  return 0;
}

/* Supergather of one CDP into the work arrays (FP64 geometry stored as float,
 * the GPU exact mode's unscaled slots). */
static void de_load_cdp(de_work* W, int cdp, const int32_t* cdp_off, const int32_t* nbr_off, const int32_t* nbr) {
  const int size = crs_supergather(cdp, cdp_off, nbr_off, nbr, W->m0x, W->m0y, W->h0, W->rows, W->m2d, W->md, W->hd);
  for (int k = 0; k < size; k++) {
    W->m2[k] = (float)W->m2d[k];
    W->m[k] = (float)W->md[k];
    W->h[k] = (float)W->hd[k];
  }
  W->S.size = size;
}

static void de_close(de_work* W) {
  free(W->m0x);
  free(W->m0y);
  free(W->h0);
  free(W->rows);
  free(W->m2d);
  free(W->md);
  free(W->hd);
  free(W->m2);
  free(W->m);
  free(W->h);
}

int orc_crs_de(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
               const int32_t* nbr_off, const int32_t* nbr, double a0d, double a1d, double b0d, double b1d, double c0d,
               double c1d, double taud, int pop, int ngen, double Fd, double CRd, uint64_t seed, float* pa,
               float* pb, float* pc, float* coher, float* stack) {
  return orc_crs_de_range(ncdp, cdp_off, ns, dt_us, g, samples, nbr_off, nbr, a0d, a1d, b0d, b1d, c0d, c1d, taud, pop,
                          ngen, Fd, CRd, seed, 0, ncdp, pa, pb, pc, coher, stack);
}

int orc_crs_de_range(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                     const int32_t* nbr_off, const int32_t* nbr, double a0d, double a1d, double b0d, double b1d,
                     double c0d, double c1d, double taud, int pop, int ngen, double Fd, double CRd, uint64_t seed,
                     int cdp_begin, int cdp_end, float* pa, float* pb, float* pc, float* coher, float* stack) {
  if (cdp_begin < 0 || cdp_end > ncdp || cdp_begin >= cdp_end) return 1;
  de_work W;
  if (de_open(&W, ncdp, cdp_off, ns, dt_us, g, samples, nbr_off, nbr, a0d, a1d, b0d, b1d, c0d, c1d, taud, pop, ngen,
              Fd, CRd, seed))
    return 1;
  for (int cdp = cdp_begin; cdp < cdp_end; cdp++) {
    de_load_cdp(&W, cdp, cdp_off, nbr_off, nbr);
#pragma omp parallel for schedule(dynamic, 4)
    for (int t0 = 0; t0 < ns; t0++) {
      float bx[3], bf, bs;
      de_point(&W.S, cdp, t0, bx, &bf, &bs, NULL);
      const size_t o = (size_t)(cdp - cdp_begin) * ns + t0;
      pa[o] = bx[0];
      pb[o] = bx[1];
      pc[o] = (float)((double)bx[2] / FACTOR);
      coher[o] = bf;
      stack[o] = bs;
    }
  }
  de_close(&W);
  return 0;
}

int orc_de_history(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                   const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, double b0, double b1, double c0,
                   double c1, double tau, int pop, int ngen, double F, double CR, uint64_t seed, int cdp, int t0,
                   float* pop_x, float* pop_f, float* trial_x, float* trial_f, int32_t* picks, float* best) {
  if (cdp < 0 || cdp >= ncdp || t0 < 0 || t0 >= ns) return 1;
  de_work W;
  if (de_open(&W, ncdp, cdp_off, ns, dt_us, g, samples, nbr_off, nbr, a0, a1, b0, b1, c0, c1, tau, pop, ngen, F, CR,
              seed))
    return 1;
  de_load_cdp(&W, cdp, cdp_off, nbr_off, nbr);
  de_hist H = {pop_x, pop_f, trial_x, trial_f, picks};
  de_point(&W.S, cdp, t0, best, best + 3, best + 4, &H);
  de_close(&W);
  return 0;
}
