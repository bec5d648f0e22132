/*
 * oracle.h -- CPU restatement of the reference's CMP / CRS semblance search.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libsemblance) never links, calls or falls back to it.
 *
 * Every routine works on a binned, linearized gather (rows [cdp_off[i],
 * cdp_off[i+1]) belong to CDP i; samples [ntr][ns] float) and writes
 * per-(CDP, t0) outputs [ncdp][ns]:
 *   index  : CMP c index (-1 none) / CRS flat (ia*nb+ib)*nc+ic
 *   pa,pb,pc: parameter values as written to the reference's output files
 *   coher, stack
 * panel (optional, may be NULL): every (CDP, t0, param) semblance and stack,
 * [ncdp][ns][npar], for tie analysis on small inputs.
 */
#ifndef SB_ORACLE_H
#define SB_ORACLE_H
#include <stdint.h>

typedef struct {
  int32_t cdp, scalco, sx, sy, gx, gy;
} orc_geom;

/* Reference precision (float) CMP: CMP/OpenMP/src/main.cpp:68-253. */
int orc_cmp(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
            double c0, double c1, int nc, double tau, int32_t* index, float* pc, float* coher, float* stack,
            float* panel_sem, float* panel_stk);

/* Reference precision (double) CRS: CRS/OpenMP/src/main.cpp:85-350. */
int orc_crs(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
            const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, int na, double b0, double b1, int nb,
            double c0, double c1, int nc, double tau, int32_t* index, float* pa, float* pb, float* pc,
            float* coher, float* stack, float* panel_sem, float* panel_stk);

/* orc_crs over CDPs [cdp_begin, cdp_end) only (their supergathers may reach
 * any CDP of the gather); outputs (and panels) have cdp_end - cdp_begin rows. */
int orc_crs_range(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                  const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, int na, double b0, double b1,
                  int nb, double c0, double c1, int nc, double tau, int cdp_begin, int cdp_end, int32_t* index,
                  float* pa, float* pb, float* pc, float* coher, float* stack, float* panel_sem, float* panel_stk);

/* orc_crs_range with az_mdeg >= 0: the midpoint displacement m of every
 * supergather trace is the signed projection of dm onto the azimuth az_mdeg
 * (millidegrees clockwise from +y), m^2 = m m -- the library's
 * SB_OPT_DM_AZIMUTH extension, which has no reference counterpart.  -1: as
 * orc_crs_range. */
int orc_crs_range_az(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                     const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, int na, double b0, double b1,
                     int nb, double c0, double c1, int nc, double tau, int az_mdeg, int cdp_begin, int cdp_end,
                     int32_t* index, float* pa, float* pb, float* pc, float* coher, float* stack, float* panel_sem,
                     float* panel_stk);

/* FP64 oracle for error statements: the CMP search with every quantity in
 * double (geometry from the integer coordinates, grid, traveltime, sums).
 * Outputs as orc_cmp but in double. */
int orc_cmp_f64(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                double c0, double c1, int nc, double tau, int32_t* index, double* coher, double* stack,
                double* panel_sem);

/* FP64 evaluation of one CRS candidate (any a, b, c; c unscaled, s^2/m^2)
 * at (cdp, t0): semblance and stack with the reference's validity rules. */
int orc_crs_eval_f64(int cdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                     const int32_t* nbr_off, const int32_t* nbr, double a, double b, double c, double tau, int t0,
                     double* sem, double* stk);

/* CPU twin of the CRS-DE search (FP32, the GPU exact mode's operation order;
 * the counter-based random stream is shared with the GPU kernel). */
int orc_crs_de(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
               const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, double b0, double b1, double c0,
               double c1, double tau, int pop, int ngen, double F, double CR, uint64_t seed, float* pa, float* pb,
               float* pc, float* coher, float* stack);

/* orc_crs_de over CDPs [cdp_begin, cdp_end) only; outputs have
 * cdp_end - cdp_begin rows. */
int orc_crs_de_range(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                     const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, double b0, double b1,
                     double c0, double c1, double tau, int pop, int ngen, double F, double CR, uint64_t seed,
                     int cdp_begin, int cdp_end, float* pa, float* pb, float* pc, float* coher, float* stack);

/* The same search at one (cdp, t0), with its whole trajectory: pop_x[g]
 * [pop][3] (c scaled by 1e6) and pop_f[g][pop] = the population and its
 * fitness after g generations (g in [0, ngen]; 0 = the initial one); the
 * trials of generation g + 1, made from pop_x[g], in trial_x[g][pop][3] with
 * fitness trial_f[g][pop] and draws picks[g][pop][4] = {r1, r2, r3, jr}
 * (g in [0, ngen)).  best[5] = the reported a, b, c (scaled), fitness and
 * stack. */
int orc_de_history(int ncdp, const int32_t* cdp_off, int ns, int dt_us, const orc_geom* g, const float* samples,
                   const int32_t* nbr_off, const int32_t* nbr, double a0, double a1, double b0, double b1, double c0,
                   double c1, double tau, int pop, int ngen, double F, double CR, uint64_t seed, int cdp, int t0,
                   float* pop_x, float* pop_f, float* trial_x, float* trial_f, int32_t* picks, float* best);

/* Counter-based uniform draw shared with the GPU DE kernel (for tests). */
uint32_t orc_de_hash(uint64_t seed, uint32_t cdp, uint32_t t0, uint32_t gen, uint32_t member, uint32_t draw);

#endif
