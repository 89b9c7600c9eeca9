/* The oracle: a plain-C restatement of the reference's CPU sensor path, for tests only.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load or call this library, and only as the checker (or the timed CPU baseline) - never as the
 * product path, which is the HIP library ballasim_amd/lib/libbsx.so.
 *
 * It restates, in its own code and in the reference's order of operations (fp32 results must be
 * bitwise the reference's fp32 CPU path):
 *   - Philox4x32-10 and the draw layout   src/core/rng/philox.h:45-53, purpose.h:45-53
 *   - exp, log, expm1, sincos (fp32/fp64)  src/core/pmath/explog.h:228-281, trig.h:136-162
 *   - resid_exp, erfcx, dawson             src/core/pmath/special.h:29-211
 *   - the frozen maps                      src/sensor/maps.cpp:26-156
 *   - init                                 src/sensor/pixel.h:792-845
 *   - the instantaneous response           src/sensor/pixel.h:870-907, 1781-1924
 *   - the cascade without noise            src/sensor/pixel.h:909-945, 1186-1232, 1455-1628,
 *                                          substep.h:90-241, stage1.h:48-92, expcubic.h:33-282
 *   - fire and labels                      src/sensor/pixel.h:288-307, 1531-1588
 *   - the CPU backend's pixel blocks       src/sensor/sensor_model.cpp:951-994
 *   - finish_events and the readout        src/sensor/sensor_model.cpp:996-1029, readout.h:22-26
 *   - the in-path noise (frozen diffusion) src/sensor/noise.h, noise_tables.{h,cpp}, pixel.h:394-745,
 *                                          821-845, 1270-1453, 1599-1652 (orc_noise.inc)
 *   - the rate-specified shot noise        (this library's extension; no reference counterpart)
 * How each part is pinned is stated in oracle/README.md.
 */
#ifndef BSX_ORACLE_H_
#define BSX_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#include "../include/bsx/bsx_sensor.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

typedef struct orc_event {
  uint32_t pixel;
  uint8_t polarity;
  uint8_t label;
  uint16_t flags;
  int64_t t_cross;
  float ds, dn, dl, l_first;
} orc_event;

/* The per-pixel state as the reference's fp64 PixelState holds it (subset). */
typedef struct orc_state {
  double z, x1, s2, s3, i_scale, tv, s3_ref, n3_ref, dr_ref;
  int64_t t_ref, blind_until, t_search;
  uint16_t flags;
  float held[8];
  int64_t shot_next;   /* the shot-noise process's next arrival (INT64_MAX: off) */
  uint32_t shot_index; /* its index */
  double nk[3], nk1[3], ln_a1; /* the in-path noise: lattice nodes and the cell's rate */
  int64_t cell;
  uint8_t have_k1;
} orc_state;

typedef struct orc_stats {
  uint64_t substeps, substep_attempts, settled_substeps, leaf_intervals, fallbacks, input_clamps, quarantines,
      pieces, cells, refinements;
} orc_stats;

/* Returns null and fills err (if given) on invalid parameters. num_threads <= 0: all cores. */
orc_model* orc_create(const bsx_params* p, uint32_t width, uint32_t height, uint64_t seed, uint32_t camera,
                      int32_t groups, int32_t num_threads, char* err, size_t err_n);
void orc_destroy(orc_model* m);
/* The frozen maps of n pixels (every output array of n values; hot and hot_rate may be null). */
void orc_draw_maps(const bsx_params* p, uint64_t seed, uint32_t camera, uint32_t n, double* theta_on,
                   double* theta_off, double* leak, double* dark, double* phi0, uint8_t* hot, double* hot_rate);
int orc_init(orc_model* m, int64_t t0_ns, const float* irradiance0);
/* The reference's advance_runs: 0 ok, 1 invalid input (message in err). Events are then sorted. */
int orc_advance_runs(orc_model* m, const uint32_t* ids, size_t n, const int64_t* t, const float* irr,
                     size_t n_samples, const int64_t* row_start, const int32_t* row_count, int64_t t_end_ns,
                     char* err, size_t err_n);
size_t orc_event_count(const orc_model* m);
const orc_event* orc_events(const orc_model* m);
/* t_reported of the sorted events, into out (orc_event_count entries). */
void orc_reported(const orc_model* m, int64_t* out);
void orc_maps(const orc_model* m, double* theta_on, double* theta_off, double* leak, double* dark, double* phi0,
              uint8_t* hot, double* hot_rate);
void orc_states(const orc_model* m, orc_state* out);
void orc_get_stats(const orc_model* m, orc_stats* out);
int64_t orc_time(const orc_model* m);
/* The noise tables of p (frozen diffusion): n_nodes rows of stride values, fp64 and fp32 (either
 * output may be null); 0 ok, 1 bad leaf level. */
int orc_noise_tables(const bsx_params* p, double* out64, float* out32, int32_t* n_nodes, int32_t* stride);
/* normals3 of the noise purposes, fp64 (all 32 bits of each word) and fp32 (either may be null). */
void orc_normals3(uint64_t seed, uint32_t purpose, uint32_t camera, uint32_t pixel, uint64_t index, double* out64,
                  float* out32);

/* The maths, exported for unit tests. */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_expf(float x);
float orc_logf(float x);
float orc_expm1f(float x);
double orc_exp(double x);
double orc_log(double x);
void orc_sincos(double x, double* s, double* c);
void orc_sincosf(float x, float* s, float* c);
float orc_resid_expf(int j, float y);
float orc_erfcxf(float x);
float orc_dawsonf(float x);
float orc_z_step(float z0, float I0, float I1, float h, float k);
/* Over every positive finite float (bit patterns 1 .. 0x7f7fffff): monotonicity violations of
 * orc_logf and the sum mod 2^64 of splitmix64((u << 32) | bits(orc_logf)), on `threads` threads. */
void orc_log_check(int threads, uint64_t* violations, uint64_t* hash);
/* The comparator input g(u) = c0 + c1 u + c2 u^2 + c3 u^3 + C0 R_j(a0 u) + C1 R_j(a1 u) (j = order) and
 * its first firing: g at u, its critical points on (lo, hi) (returns their count, at most 6, into out),
 * and the first integer ns in [max(lo, e0), e1) where g - th_on >= 0 or -g - th_off >= 0 (1 with *on and
 * *t, or 0). */
float orc_xc_eval(const float c[4], const float C[2], const float a[2], int order, float u);
void orc_xc_eval_n(const float c[4], const float C[2], const float a[2], int order, const float* u, size_t n,
                   float* out);
int orc_xc_critical(const float c[4], const float C[2], const float a[2], int order, float lo, float hi,
                    float* out);
int orc_xc_search(const float c[4], const float C[2], const float a[2], int order, int64_t e0, int64_t e1,
                  int64_t lo, float th_on, float th_off, int* on, int64_t* t);

#ifdef __cplusplus
}
#endif

#endif
