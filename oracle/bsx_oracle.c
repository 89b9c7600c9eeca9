/* The oracle (test infrastructure; see bsx_oracle.h): a plain-C restatement of the reference's CPU
 * sensor path. Compiled with -ffp-contract=off and without fast-math, like the reference. */
#include "bsx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "orc_tables.h"

/* ------------------------------------------------------------------------------------------------
 * Philox4x32-10 and the draw layout (src/core/rng/philox.h:45-53, src/core/rng/purpose.h:45-53)
 * ---------------------------------------------------------------------------------------------- */

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static void orc_draw(uint64_t seed, uint32_t purpose, uint32_t camera, uint32_t pixel, uint64_t index,
                     uint32_t out[4]) {
  const uint32_t ctr[4] = {(purpose << 8) | (camera & 0xFFu), pixel, (uint32_t)index, (uint32_t)(index >> 32)};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox(ctr, key, out);
}

static double u_open64(uint32_t u) { return ((double)u + 0.5) * 0x1p-32; }

/* ------------------------------------------------------------------------------------------------
 * pmath (src/core/pmath/explog.h, trig.h, special.h): Cody-Waite + Taylor, atanh series
 * ---------------------------------------------------------------------------------------------- */

static float ldexpf_x(float x, int e) {
  const int e1 = e > 127 ? 127 : (e < -126 ? -126 : e), e2 = e - e1;
  uint32_t b1 = (uint32_t)(e1 + 127) << 23, b2 = (uint32_t)(e2 + 127) << 23;
  float p1, p2;
  memcpy(&p1, &b1, 4);
  memcpy(&p2, &b2, 4);
  return e2 == 0 ? x * p1 : (x * p1) * p2;
}

static double ldexp_x(double x, int e) {
  const int e1 = e > 1023 ? 1023 : (e < -1022 ? -1022 : e), e2 = e - e1;
  uint64_t b1 = (uint64_t)(e1 + 1023) << 52, b2 = (uint64_t)(e2 + 1023) << 52;
  double p1, p2;
  memcpy(&p1, &b1, 8);
  memcpy(&p2, &b2, 8);
  return e2 == 0 ? x * p1 : (x * p1) * p2;
}

static float frexpf_x(float x, int* e) {
  int bias = 0;
  if (x < 0x1p-126f) {
    x = x * 0x1p25f;
    bias = 25;
  }
  uint32_t b;
  memcpy(&b, &x, 4);
  *e = (int)((b >> 23) & 0xff) - 126 - bias;
  b = (b & 0x807fffffu) | (126u << 23);
  float m;
  memcpy(&m, &b, 4);
  return m;
}

static double frexp_x(double x, int* e) {
  int bias = 0;
  if (x < 0x1p-1022) {
    x = x * 0x1p54;
    bias = 54;
  }
  uint64_t b;
  memcpy(&b, &x, 8);
  *e = (int)((b >> 52) & 0x7ff) - 1022 - bias;
  b = (b & 0x800fffffffffffffull) | ((uint64_t)1022 << 52);
  double m;
  memcpy(&m, &b, 8);
  return m;
}

float orc_expf(float x) {
  if (!(x >= -87.33f)) return x != x ? x : 0.0f;
  if (x > 88.72f) return INFINITY;
  const float k = rintf(x * 0x1.715476p+0f);
  const float r = (x - k * 0x1.62e400p-1f) - k * 0x1.7f7d1cp-20f;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return ldexpf_x(p, (int)k);
}

double orc_exp(double x) {
  if (!(x >= -708.39)) return x != x ? x : 0.0;
  if (x > 709.78) return INFINITY;
  const double k = rint(x * 0x1.71547652b82fep+0);
  const double r = (x - k * 0x1.62e42fee00000p-1) - k * 0x1.a39ef35793c76p-33;
  static const double inv_fact[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                      1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,
                                      1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,        0.5,
                                      1.0,                1.0};
  double p = inv_fact[0];
  for (int i = 1; i < 14; ++i) p = p * r + inv_fact[i];
  return ldexp_x(p, (int)k);
}

float orc_logf(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int e = 0;
  float m = frexpf_x(x, &e);
  if (m < 0x1.6a09e6p-1f) {
    m = m + m;
    e -= 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float R = 2.0f / 9.0f;
  R = R * z + 2.0f / 7.0f;
  R = R * z + 2.0f / 5.0f;
  R = R * z + 2.0f / 3.0f;
  R = R * z;
  const float lm = f - s * (f - R);
  const float ef = (float)e;
  return ef * 0x1.62e400p-1f + (lm + ef * 0x1.7f7d1cp-20f);
}

double orc_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  int e = 0;
  double m = frexp_x(x, &e);
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = m + m;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 2.0 / 19.0;
  for (int k = 17; k >= 3; k -= 2) R = R * z + 2.0 / (double)k;
  R = R * z;
  const double lm = f - s * (f - R);
  const double ef = (double)e;
  return ef * 0x1.62e42fee00000p-1 + (lm + ef * 0x1.a39ef35793c76p-33);
}

float orc_expm1f(float x) {
  if (fabsf(x) < 0.35f) {
    float p = 1.0f / 40320.0f;
    p = p * x + 1.0f / 5040.0f;
    p = p * x + 1.0f / 720.0f;
    p = p * x + 1.0f / 120.0f;
    p = p * x + 1.0f / 24.0f;
    p = p * x + 1.0f / 6.0f;
    p = p * x + 0.5f;
    return (p * x) * x + x;
  }
  return orc_expf(x) - 1.0f;
}

void orc_sincos(double x, double* s, double* c) {
  if (!(fabs(x) <= 1.0e6)) {
    *s = *c = NAN;
    return;
  }
  if (x == 0.0) {
    *s = x;
    *c = 1.0;
    return;
  }
  const double q = rint(x * 0x1.45f306dc9c883p-1);
  const double r = q == 0.0 ? x : ((x - q * 0x1.921fb544p+0) - q * 0x1.0b4611a6p-34) - q * 0x1.3198a2e037073p-69;
  const double z = r * r;
  /* sin: odd Taylor terms 1/17! .. 1/3!; cos: even terms 1/18! .. 1/4! (exact reciprocals of n!) */
  double fact[19];
  fact[0] = 1.0;
  for (int i = 1; i <= 18; ++i) fact[i] = fact[i - 1] * (double)i;
  double ps = 1.0 / fact[17];
  ps = ps * z - 1.0 / fact[15];
  ps = ps * z + 1.0 / fact[13];
  ps = ps * z - 1.0 / fact[11];
  ps = ps * z + 1.0 / fact[9];
  ps = ps * z - 1.0 / fact[7];
  ps = ps * z + 1.0 / fact[5];
  ps = ps * z - 1.0 / fact[3];
  const double sr = r + (r * z) * ps;
  double pc = -1.0 / fact[18];
  pc = pc * z + 1.0 / fact[16];
  pc = pc * z - 1.0 / fact[14];
  pc = pc * z + 1.0 / fact[12];
  pc = pc * z - 1.0 / fact[10];
  pc = pc * z + 1.0 / fact[8];
  pc = pc * z - 1.0 / fact[6];
  pc = pc * z + 1.0 / fact[4];
  const double cr = (1.0 - 0.5 * z) + (z * z) * pc;
  switch ((int)q & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

void orc_sincosf(float x, float* s, float* c) {
  if (!(fabsf(x) <= 6000.0f)) {
    *s = *c = NAN;
    return;
  }
  if (x == 0.0f) {
    *s = x;
    *c = 1.0f;
    return;
  }
  const float q = rintf(x * 0x1.45f306p-1f);
  const float r = q == 0.0f ? x : ((x - q * 0x1.922p+0f) - q * -0x1.2aep-18f) - q * -0x1.de973ep-31f;
  const float z = r * r;
  float ps = 1.0f / 362880.0f;
  ps = ps * z - 1.0f / 5040.0f;
  ps = ps * z + 0x1.111110p-7f; /* 1/5! to 8 significant digits, as the reference's fp32 sine */
  ps = ps * z - 1.0f / 6.0f;
  const float sr = r + (r * z) * ps;
  float pc = -1.0f / 3628800.0f;
  pc = pc * z + 1.0f / 40320.0f;
  pc = pc * z - 1.0f / 720.0f;
  pc = pc * z + 1.0f / 24.0f;
  const float cr = (1.0f - 0.5f * z) + (z * z) * pc;
  switch ((int)q & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

float orc_resid_expf(int j, float y) {
  if (j <= 0) return orc_expf(-y);
  float yj = 1.0f, jf = 1.0f;
  for (int i = 1; i <= j; ++i) {
    yj = yj * (-y);
    jf = jf * (float)i;
  }
  if (y > 1.0f) {
    float t = 1.0f, s = 1.0f;
    for (int i = 1; i < j; ++i) {
      t = t * (-y) / (float)i;
      s = s + t;
    }
    return orc_expf(-y) - s;
  }
  float q = 1.0f;
  for (int i = 10; i >= 1; --i) q = 1.0f - y * q / (float)(j + i);
  return yj * (q / jf);
}

float orc_erfcxf(float x) {
  if (!(x > 0.0f && x < INFINITY)) return x > 0.0f ? 0.0f : (x == x ? 1.0f : x);
  const float L = (float)orc_weideman_L24;
  const float lx = L + x;
  const float Z = (L - x) / lx;
  float p = (float)orc_weideman_p24[0];
  for (int i = 1; i < 24; ++i) p = p * Z + (float)orc_weideman_p24[i];
  return 2.0f * p / (lx * lx) + (float)0x1.20dd750429b6dp-1 / lx;
}

float orc_dawsonf(float x) {
  const float ax = fabsf(x);
  if (ax < 1.0f) {
    const float two_x2 = 2.0f * x * x;
    float q = 1.0f;
    for (int n = 12; n >= 1; --n) q = 1.0f - two_x2 * q / (float)(2 * n + 1);
    return x * q;
  }
  if (!(ax <= 1e8f)) return ax == ax ? copysignf(0.5f / ax, x) : x;
  const float h = 0.25f;
  const int n0 = 2 * (int)(0.5f * ax / h + 0.5f);
  const float xp = ax - (float)n0 * h;
  float e1 = orc_expf(2.0f * xp * h);
  const float e2 = e1 * e1;
  float d1 = (float)(n0 + 1), d2 = (float)(n0 - 1), sum = 0.0f;
  for (int i = 0; i < 13; ++i) {
    sum = sum + (float)orc_rybicki[i] * (e1 / d1 + 1.0f / (d2 * e1));
    d1 = d1 + 2.0f;
    d2 = d2 - 2.0f;
    e1 = e1 * e2;
  }
  return copysignf((float)0x1.20dd750429b6dp-1 * orc_expf(-xp * xp) * sum, x);
}

static uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

typedef struct {
  uint32_t u0, u1;
  uint64_t bad, hash;
} logchk;

static void* log_check_worker(void* arg) {
  logchk* c = (logchk*)arg;
  for (uint32_t u = c->u0; u < c->u1; ++u) {
    float x, x1;
    memcpy(&x, &u, 4);
    const float y = orc_logf(x);
    uint32_t b;
    memcpy(&b, &y, 4);
    c->hash += splitmix64(((uint64_t)u << 32) | b);
    if (u + 1 < 0x7f800000u) {
      const uint32_t v = u + 1;
      memcpy(&x1, &v, 4);
      if (orc_logf(x1) < y) ++c->bad;
    }
  }
  return NULL;
}

void orc_log_check(int threads, uint64_t* violations, uint64_t* hash) {
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  logchk c[64];
  const uint32_t n = 0x7f800000u - 1u;
  for (int i = 0; i < threads; ++i) {
    c[i].u0 = 1u + (uint32_t)((uint64_t)n * i / threads);
    c[i].u1 = 1u + (uint32_t)((uint64_t)n * (i + 1) / threads);
    c[i].bad = c[i].hash = 0;
    pthread_create(&th[i], NULL, log_check_worker, &c[i]);
  }
  *violations = 0;
  *hash = 0;
  for (int i = 0; i < threads; ++i) {
    pthread_join(th[i], NULL);
    *violations += c[i].bad;
    *hash += c[i].hash;
  }
}

/* ------------------------------------------------------------------------------------------------
 * The model
 * ---------------------------------------------------------------------------------------------- */

typedef struct {
  float theta_on, theta_off, leak, dark, scale, phi0;
  uint8_t hot;
} pmap32;

struct orc_model {
  bsx_params p;
  uint32_t width, height, camera;
  uint64_t seed;
  int G;
  size_t P;
  int threads;
  double *theta_on, *theta_off, *leak, *dark, *phi0;
  uint8_t* hot;
  double* hot_rate;
  pmap32* m32;
  orc_state* st;
  int64_t t_state;
  int initialized;
  float irr_max;
  int64_t refractory_ns;
  int64_t q_ns, delay_ns, phase_ns;
  orc_event* ev;
  size_t n_ev, cap_ev;
  orc_stats stats;
  /* the noise tables (cascade with noise on): fp32 rows, node count, stride, grid origin and
   * inverse spacing as the fp32 kernel reads them, and the constants the walk uses */
  int noise;
  float* tab32;
  int tab_nodes, tab_stride;
  float tab_ln_min, tab_inv_dln;
  float a3f, margin_z;
};

static void set_err(char* err, size_t n, const char* msg) {
  if (err != NULL && n > 0) {
    strncpy(err, msg, n - 1);
    err[n - 1] = '\0';
  }
}

static double dmaxd(double a, double b) { return a < b ? b : a; }

static double arrhenius(double ea, double T, double T0) {
  return orc_exp(-ea / 8.617333262e-5 * (1.0 / T - 1.0 / T0));
}

/* Four normals of Philox block `block` of the sensor-map purpose by Box-Muller (maps.cpp:26-39). */
static void map_normals4(const orc_model* m, uint32_t px, uint64_t block, double* xi) {
  const double two_pi = 0x1.921fb54442d18p+2;
  uint32_t r[4];
  orc_draw(m->seed, 0x0002, m->camera, px, block, r);
  const double u0 = u_open64(r[0]), u1 = u_open64(r[1]), u2 = u_open64(r[2]), u3 = u_open64(r[3]);
  const double r0 = sqrt(-2.0 * orc_log(u0)), r1 = sqrt(-2.0 * orc_log(u2));
  double s0, c0, s1, c1;
  orc_sincos(two_pi * u1, &s0, &c0);
  orc_sincos(two_pi * u3, &s1, &c1);
  xi[0] = r0 * c0;
  xi[1] = r0 * s0;
  xi[2] = r1 * c1;
  xi[3] = r1 * s1;
}

/* The frozen maps (src/sensor/maps.cpp:21-39, 78-156): block 0 the four normals, block 1 the
 * uniforms (phase, hot, type, rate), block 2 a hot pixel's log-normal rate normal. Hot pixels are
 * leak-type here (the reference's rule without in-path noise, maps.cpp:119-124). */
static void draw_maps(orc_model* m) {
  const bsx_params* p = &m->p;
  const double leak_temp = arrhenius(p->leak_ea_eV, p->temperature_K, p->reference_temperature_K);
  const double dark_temp = arrhenius(p->dark_ea_eV, p->temperature_K, p->reference_temperature_K);
  const double rho = p->theta_onoff_corr, rho_c = sqrt(1.0 - rho * rho);
  const double ln10 = orc_log(10.0);
  const double ln_rmin = orc_log(p->hot_rate_min_hz), ln_rmax = orc_log(p->hot_rate_max_hz);
  for (uint32_t px = 0; px < m->P; ++px) {
    uint32_t u[4];
    double xi[4];
    map_normals4(m, px, 0, xi);
    orc_draw(m->seed, 0x0002, m->camera, px, 1, u);
    const double u_phase = u_open64(u[0]), u_hot = u_open64(u[1]), u_rate = u_open64(u[3]);
    const double xi_off = rho == 0.0 ? xi[1] : rho * xi[0] + rho_c * xi[1];
    const double th_on = dmaxd(p->theta_floor, p->theta_on + p->sigma_theta_on * xi[0]);
    const double th_off = dmaxd(p->theta_floor, p->theta_off + p->sigma_theta_off * xi_off);
    double leak_rate = p->leak_rate_hz * orc_exp(p->leak_sigma_log10 * xi[2] * ln10) * leak_temp;
    uint8_t hot = 0;
    double rate = 0.0;
    if (u_hot < p->hot_fraction) {
      if (p->hot_rate_law == 0) {
        rate = orc_exp(ln_rmin + u_rate * (ln_rmax - ln_rmin));
      } else {
        double z[4];
        map_normals4(m, px, 2, z);
        rate = p->hot_rate_median_hz * orc_exp(p->hot_rate_sigma_ln * z[0]);
        rate = rate < p->hot_rate_min_hz ? p->hot_rate_min_hz : rate > p->hot_rate_max_hz ? p->hot_rate_max_hz : rate;
      }
      hot = 1;
      leak_rate = p->reset_mode == BSX_RESET_PHYSICAL ? rate / (1.0 - rate * p->refractory_s) : rate;
    }
    if (m->hot) m->hot[px] = hot;
    if (m->hot_rate) m->hot_rate[px] = rate;
    m->theta_on[px] = th_on;
    m->theta_off[px] = th_off;
    m->dark[px] = p->dark_current_fA * orc_exp(p->dark_sigma_ln * xi[3]) * dark_temp;
    m->leak[px] = th_on * leak_rate;
    m->phi0[px] = p->initial_phase == BSX_PHASE_UNIFORM     ? u_phase * th_on
                  : p->initial_phase == BSX_PHASE_SYMMETRIC ? -th_off + u_phase * (th_on + th_off)
                                                            : 0.0;
  }
}

void orc_draw_maps(const bsx_params* p, uint64_t seed, uint32_t camera, uint32_t n, double* theta_on,
                   double* theta_off, double* leak, double* dark, double* phi0, uint8_t* hot, double* hot_rate) {
  orc_model m;
  memset(&m, 0, sizeof(m));
  m.p = *p;
  m.seed = seed;
  m.camera = camera;
  m.P = n;
  m.theta_on = theta_on;
  m.theta_off = theta_off;
  m.leak = leak;
  m.dark = dark;
  m.phi0 = phi0;
  m.hot = hot;
  m.hot_rate = hot_rate;
  draw_maps(&m);
}

static void setup_noise(orc_model* m); /* orc_noise.inc */

orc_model* orc_create(const bsx_params* p, uint32_t width, uint32_t height, uint64_t seed, uint32_t camera,
                      int32_t groups, int32_t num_threads, char* err, size_t err_n) {
  if (width == 0 || height == 0 || groups < 1 || groups > 8 || camera > 255) {
    set_err(err, err_n, "orc_create: bad size, groups or camera");
    return NULL;
  }
  const int noise_on = p->noise && p->photoreceptor == BSX_PHOTORECEPTOR_CASCADE;
  if (p->veiling_glare != 0 || p->tau_hp != 0 || (p->shot_rate_hz > 0 && p->photoreceptor == BSX_PHOTORECEPTOR_CASCADE) ||
      (noise_on && (p->noise_diffusion != 0 || p->initial_state != 0 || p->noise_leaf_level < 0 ||
                    p->noise_leaf_level > 10 || (p->hot_fraction != 0 && p->hot_leak_fraction < 1.0)))) {
    set_err(err, err_n,
            "orc_create: glare, the high-pass filter, the shot process on the cascade, transient diffusion, the "
            "stationary start and threshold-type hot pixels are not restated");
    return NULL;
  }
  orc_model* m = (orc_model*)calloc(1, sizeof(orc_model));
  m->p = *p;
  m->width = width;
  m->height = height;
  m->camera = camera;
  m->seed = seed;
  m->G = groups;
  m->P = (size_t)width * height;
  m->threads = num_threads;
  m->theta_on = (double*)malloc(m->P * 8);
  m->theta_off = (double*)malloc(m->P * 8);
  m->leak = (double*)malloc(m->P * 8);
  m->dark = (double*)malloc(m->P * 8);
  m->phi0 = (double*)malloc(m->P * 8);
  m->hot = (uint8_t*)malloc(m->P);
  m->hot_rate = (double*)malloc(m->P * 8);
  m->m32 = (pmap32*)malloc(m->P * sizeof(pmap32));
  m->st = (orc_state*)calloc(m->P, sizeof(orc_state));
  draw_maps(m);
  const float scale = (float)p->illuminance_scale;
  for (size_t i = 0; i < m->P; ++i) {
    m->m32[i].theta_on = (float)m->theta_on[i];
    m->m32[i].theta_off = (float)m->theta_off[i];
    m->m32[i].leak = (float)m->leak[i];
    m->m32[i].dark = (float)m->dark[i];
    m->m32[i].scale = scale * 1.0f;
    m->m32[i].phi0 = (float)m->phi0[i];
    m->m32[i].hot = m->hot[i];
  }
  m->irr_max = (float)(1e9 / (p->illuminance_scale * 1.0));
  m->noise = noise_on;
  if (noise_on) setup_noise(m);
  m->refractory_ns = llround(p->refractory_s * 1e9);
  m->q_ns = llround(p->readout_quantum_s * 1e9);
  m->delay_ns = llround(p->readout_delay_s * 1e9);
  if (m->q_ns > 0) {
    uint32_t r[4];
    orc_draw(seed, 0x0022, camera, 0, 0, r);
    m->phase_ns = (int64_t)(r[0] % (uint64_t)m->q_ns);
  }
  return m;
}

void orc_destroy(orc_model* m) {
  if (m == NULL) return;
  free(m->tab32);
  free(m->theta_on);
  free(m->theta_off);
  free(m->leak);
  free(m->dark);
  free(m->phi0);
  free(m->hot);
  free(m->hot_rate);
  free(m->m32);
  free(m->st);
  free(m->ev);
  free(m);
}

static float clamp_irr(float v, float vmax) { return v >= 0.0f ? (v <= vmax ? v : vmax) : 0.0f; }

/* ------------------------------------------------------------------------------------------------
 * One pixel over one call (src/sensor/pixel.h PixelKernel), fp32
 * ---------------------------------------------------------------------------------------------- */

typedef struct {
  orc_event* v;
  size_t n, cap;
} evvec;

static void ev_push(evvec* e, const orc_event* x) {
  if (e->n == e->cap) {
    e->cap = e->cap ? 2 * e->cap : 256;
    e->v = (orc_event*)realloc(e->v, e->cap * sizeof(orc_event));
  }
  e->v[e->n++] = *x;
}

typedef struct {
  const orc_model* m;
  uint32_t pixel;
  float theta_on, theta_off, leak, dark, scale;
  int hot;
  const int64_t* t;
  const float* irr;
  int n, skip, G;
  int64_t t_state;
  float held[8];
  uint32_t live;
  /* state */
  float z, x1, s2, s3, i_scale, tv, s3_ref, n3_ref, dr_ref;
  int64_t t_ref, blind_until, t_search;
  uint16_t flags;
  int64_t shot_next;
  uint32_t shot_index;
  /* the in-path noise (cascade, noise on): lattice nodes, the cell's rate, the cell and its level-0
   * Hermite and margin; seg is the sample-segment cursor */
  float nk[3], nk1[3], ln_a1, hc0[4], margin0;
  int64_t cell;
  int have_k1;
  int seg;
  evvec* out;
  orc_stats* stats;
} pix;

static int n_aug(const pix* k) { return 1 + k->n - k->skip; }
static int64_t s_time(const pix* k, int i) { return i == 0 ? k->t_state : k->t[i - 1 + k->skip]; }
static float s_value(const pix* k, int i, int g) {
  if (i == 0 && k->skip == 0) return k->held[g];
  return clamp_irr(k->irr[(int64_t)(i - 1 + k->skip) * k->G + g], k->m->irr_max);
}
static int seg_of(const pix* k, int seg, int64_t t) {
  const int last = n_aug(k) - 1;
  while (seg < last && s_time(k, seg + 1) <= t) ++seg;
  return seg;
}
static int64_t seg_end(const pix* k, int seg) { return seg + 1 < n_aug(k) ? s_time(k, seg + 1) : INT64_MAX; }

static float group_value(const pix* k, int seg, int64_t t, int g) {
  if (seg + 1 >= n_aug(k)) return s_value(k, seg, g);
  const int64_t t0 = s_time(k, seg), t1 = s_time(k, seg + 1);
  const float f = (float)(t - t0) / (float)(t1 - t0);
  const float v0 = s_value(k, seg, g), v1 = s_value(k, seg + 1, g);
  return v0 + (v1 - v0) * f;
}

static float photocurrent(const pix* k, int seg, int64_t t) {
  float I = k->dark;
  for (int g = 0; g < k->G; ++g) {
    if (!((k->live >> g) & 1u)) continue;
    I = I + 1.0f * k->scale * group_value(k, seg, t, g);
  }
  return I;
}

static float log_ratio(const pix* k, float I) { return orc_logf(I / k->i_scale); }
static float since_s(int64_t dt) { return (float)((double)dt * 1e-9); }

static float instant_delta(const pix* k, int64_t t, float x) {
  const float ds = x - k->s3_ref;
  const float dn = 0.0f - k->n3_ref;
  const float dl = k->leak * since_s(t - k->t_ref) + k->dr_ref;
  return ds + dn + dl;
}

static void scan_input(pix* k) {
  const uint32_t all = (1u << k->G) - 1u;
  uint32_t live = 0;
  if (k->skip == 0)
    for (int g = 0; g < k->G; ++g)
      if (k->held[g] != 0.0f) live |= 1u << g;
  for (int i = 0; i < k->n && live != all; ++i)
    for (int g = 0; g < k->G; ++g)
      if (clamp_irr(k->irr[(int64_t)i * k->G + g], k->m->irr_max) != 0.0f) live |= 1u << g;
  k->live = live;
}

static void set_reference(pix* k, int64_t t, float s3, float n3, float off) {
  k->t_ref = t;
  k->s3_ref = s3;
  k->n3_ref = n3;
  k->dr_ref = off;
}

static uint8_t split_label(float ds, float dl, float D, float cut) {
  const float a_s = ds / D, a_l = dl / D;
  if (a_s >= cut && a_s >= a_l) return 0;
  if (a_l >= cut && a_l > a_s) return 2;
  return 1;
}

static void fire_ex(pix* k, int64_t t, int on, float s3, float n3, int shot);
static void fire(pix* k, int64_t t, int on, float s3, float n3) { fire_ex(k, t, on, s3, n3, 0); }

/* A shot-noise event (this library's extension): an event of the drawn polarity at the arrival,
 * labelled NOISE and flagged (1 << 13), resetting the pixel as a comparator event does. */
static void fire_ex(pix* k, int64_t t, int on, float s3, float n3, int shot) {
  const float cut = (float)k->m->p.label_share_cut;
  const float ds = s3 - k->s3_ref;
  const float dn = n3 - k->n3_ref;
  const float l = k->leak * since_s(t - k->t_ref);
  const float dl = l + k->dr_ref;
  const float D = ds + dn + dl;
  uint8_t label;
  const int first = (k->flags & (1u << 12)) != 0;
  if (shot) {
    label = 1;
  } else if (first) {
    const float Dl = ds + dn + l;
    label = (on ? Dl > 0.0f : Dl < 0.0f) ? split_label(ds, l, Dl, cut) : 1;
  } else {
    label = split_label(ds, dl, D, cut);
  }
  /* a hot pixel (pixel.h:1563-1569): HOT unless SIGNAL, and flagged (1 << 5) */
  if (k->hot && label != 0) label = 3;
  orc_event e;
  e.pixel = k->pixel;
  e.polarity = on ? 1 : 0;
  e.label = label;
  e.flags = (uint16_t)(k->flags | (shot ? (1u << 13) : 0u) | (k->hot ? (1u << 5) : 0u));
  e.t_cross = t;
  e.ds = ds;
  e.dn = dn;
  e.dl = dl;
  e.l_first = first ? l : 0.0f;
  ev_push(k->out, &e);
  k->flags = 0;
  if (k->m->p.reset_mode == BSX_RESET_IDEAL) {
    set_reference(k, t, s3, n3, D - (on ? k->theta_on : -k->theta_off));
    k->t_search = t + 1 + k->m->refractory_ns;
  } else if (k->m->refractory_ns == 0) {
    set_reference(k, t, s3, n3, -(float)k->m->p.eps_ci);
    k->t_search = t + 1;
  } else {
    k->blind_until = t + k->m->refractory_ns;
    k->t_search = k->blind_until + 1;
  }
}

/* instant_may_fire with constant waveforms (curvature 0, i_lo = dark, i_hi = +inf). */
static int instant_may_fire(const pix* k, int64_t a, int64_t c, float Ia, float Ic) {
  const float eps = 0x1p-23f;
  const float curv = 0.0f, mod = 0.0f, i_lo = k->dark, i_hi = INFINITY;
  float hi = Ia < Ic ? Ic : Ia, lo = Ic < Ia ? Ic : Ia;
  const float h = (float)(c - a) * 1e-9f;
  const float hh = h * h * curv;
  const float margin = 64.0f * eps * (hi + 2.0f * mod);
  if (k->leak == 0.0f && hi - lo >= 0.5f * hh) {
    const float D = instant_delta(k, c, log_ratio(k, Ic));
    return D >= k->theta_on || D <= -k->theta_off;
  }
  if (Ic >= Ia && (curv == 0.0f || hi - lo > hh + 2.0f * margin)) {
    return instant_delta(k, c, log_ratio(k, Ic)) >= k->theta_on;
  }
  float hv = hi + 0.125f * hh;
  hi = (i_hi < hv ? i_hi : hv) + margin;
  float lv = lo - 0.125f * hh;
  lo = (lv < i_lo ? i_lo : lv) - margin;
  const float dl_a = k->leak * since_s(a - k->t_ref) + k->dr_ref;
  const float dl_c = k->leak * since_s(c - k->t_ref) + k->dr_ref;
  const float base = 0.0f - k->n3_ref - k->s3_ref;
  const float x_hi = log_ratio(k, hi);
  if (x_hi + base + dl_c + 16.0f * eps * (1.0f + fabsf(x_hi) + fabsf(k->s3_ref) + fabsf(dl_c)) >= k->theta_on)
    return 1;
  if (!(lo > 0.0f)) return 1;
  const float x_lo = log_ratio(k, lo);
  return x_lo + base + dl_a - 16.0f * eps * (1.0f + fabsf(x_lo) + fabsf(k->s3_ref) + fabsf(dl_a)) <= -k->theta_off;
}

typedef struct {
  int found, on;
  int64_t t;
  float x;
} ihit;

static int instant_fires(const pix* k, int64_t t, ihit* h) {
  const float D = instant_delta(k, t, h->x);
  h->found = D >= k->theta_on || D <= -k->theta_off;
  h->on = D >= k->theta_on;
  h->t = t;
  return h->found;
}

static ihit instant_search(pix* k, int seg, int64_t lo, int64_t hi) {
  typedef struct {
    int64_t a, c;
    float Ia, Ic;
  } node;
  const float I_lo = photocurrent(k, seg, lo);
  ihit out = {0, 0, lo, log_ratio(k, I_lo)};
  if (instant_fires(k, lo, &out)) return out;
  if (hi - lo == 1) return out;
  node stack[32];
  int sp = 0;
  stack[sp++] = (node){lo, hi - 1, I_lo, photocurrent(k, seg, hi - 1)};
  while (sp > 0) {
    const node n = stack[--sp];
    if (n.c - n.a == 1) {
      out.t = n.c;
      out.x = log_ratio(k, n.Ic);
      if (instant_fires(k, n.c, &out)) return out;
      continue;
    }
    ++k->stats->leaf_intervals;
    if (!instant_may_fire(k, n.a, n.c, n.Ia, n.Ic)) continue;
    const int64_t mid = n.a + (n.c - n.a) / 2;
    const float Im = photocurrent(k, seg, mid);
    stack[sp++] = (node){mid, n.c, Im, n.Ic};
    stack[sp++] = (node){n.a, mid, n.Ia, Im};
  }
  out.found = 0;
  return out;
}

/* The rate-specified shot-noise process (this library's extension, no reference counterpart): per
 * pixel a Poisson process of rate r = shot_rate_hz. Arrival k follows arrival k - 1 (or the init
 * time) after ceil(-ln(U) / r * 1e9) ns (at least 1), U the open uniform of word 0 of
 * Philox(seed, 0x101, camera, pixel, k); word 1's low bit is the event's polarity. */
static int64_t shot_gap(const orc_model* m, uint32_t pixel, uint64_t index, int* on) {
  uint32_t w[4];
  orc_draw(m->seed, 0x101, m->camera, pixel, index, w);
  const double e = -orc_log(u_open64(w[0]));
  const double ns = ceil(e / m->p.shot_rate_hz * 1e9);
  if (on != NULL) *on = (int)(w[1] & 1u);
  if (!(ns >= 1.0)) return 1;
  if (ns >= 0x1p62) return (int64_t)1 << 62;
  return (int64_t)ns;
}

/* The arrival at T = shot_next: an event unless T falls in the dead time (T < t_search, which
 * covers the physical blind interval), then the next arrival. */
static void shot_arrival(pix* k, int seg, int64_t T) {
  int on = 0;
  (void)shot_gap(k->m, k->pixel, k->shot_index, &on);
  if (T >= k->t_search) {
    seg = seg_of(k, seg, T);
    fire_ex(k, T, on, log_ratio(k, photocurrent(k, seg, T)), 0.0f, 1);
  }
  k->shot_index += 1;
  k->shot_next = T + shot_gap(k->m, k->pixel, k->shot_index, NULL);
}

static void advance_instant(pix* k, int64_t t_to) {
  int seg = 0;
  int64_t t = k->t_state;
  const float eps_ci = (float)k->m->p.eps_ci;
  while (t < t_to) {
    seg = seg_of(k, seg, t);
    if (k->shot_next == t) {
      shot_arrival(k, seg, t);
      continue;
    }
    int64_t e = t_to;
    const int64_t se = seg_end(k, seg), wb = ((t >> 23) + 1) * ((int64_t)1 << 23);
    if (se < e) e = se;
    if (wb < e) e = wb;
    if (k->shot_next < e) e = k->shot_next;
    ++k->stats->substeps;
    ++k->stats->pieces;
    for (;;) {
      if (k->blind_until != INT64_MIN) {
        if (k->blind_until >= e) break;
        set_reference(k, k->blind_until, log_ratio(k, photocurrent(k, seg, k->blind_until)), 0.0f, -eps_ci);
        k->blind_until = INT64_MIN;
      }
      const int64_t lo = k->t_search > t ? k->t_search : t;
      if (lo >= e) break;
      const ihit hit = instant_search(k, seg, lo, e);
      if (!hit.found) {
        if (e > k->t_search) k->t_search = e;
        break;
      }
      fire(k, hit.t, hit.on, hit.x, 0.0f);
    }
    t = e;
  }
}

/* The cascade without noise, then its in-path noise. */
static void advance_cascade(pix* k, int64_t t_end);
#include "orc_cascade.inc"
#include "orc_noise.inc"

static void load(pix* k, const orc_model* m, size_t p) {
  const orc_state* s = &m->st[p];
  k->z = (float)s->z;
  k->x1 = (float)s->x1;
  k->s2 = (float)s->s2;
  k->s3 = (float)s->s3;
  k->i_scale = (float)s->i_scale;
  k->tv = (float)s->tv;
  k->s3_ref = (float)s->s3_ref;
  k->n3_ref = (float)s->n3_ref;
  k->dr_ref = (float)s->dr_ref;
  k->t_ref = s->t_ref;
  k->blind_until = s->blind_until;
  k->t_search = s->t_search;
  k->flags = s->flags;
  k->shot_next = s->shot_next;
  k->shot_index = s->shot_index;
  for (int g = 0; g < m->G; ++g) k->held[g] = s->held[g];
  for (int i = 0; i < 3; ++i) {
    k->nk[i] = (float)s->nk[i];
    k->nk1[i] = (float)s->nk1[i];
  }
  k->ln_a1 = (float)s->ln_a1;
  k->cell = s->cell;
  k->have_k1 = s->have_k1;
  if (m->noise && k->have_k1) {
    /* a call resumed inside a cell: its level-0 hull as draw_lattice left it */
    float m0[23];
    if (m->p.noise_leaf_level > 0)
      noise_lookup(m, noise_pos(m, k->ln_a1), noise_margin_off(m->p.noise_leaf_level, 0), 23, m0);
    level0(k, m0);
  }
}

static void store(const pix* k, orc_model* m, size_t p) {
  orc_state* s = &m->st[p];
  s->z = k->z;
  s->x1 = k->x1;
  s->s2 = k->s2;
  s->s3 = k->s3;
  s->i_scale = k->i_scale;
  s->tv = k->tv;
  s->s3_ref = k->s3_ref;
  s->n3_ref = k->n3_ref;
  s->dr_ref = k->dr_ref;
  s->t_ref = k->t_ref;
  s->blind_until = k->blind_until;
  s->t_search = k->t_search;
  s->flags = k->flags;
  s->shot_next = k->shot_next;
  s->shot_index = k->shot_index;
  for (int g = 0; g < m->G; ++g) s->held[g] = k->held[g];
  for (int i = 0; i < 3; ++i) {
    s->nk[i] = k->nk[i];
    s->nk1[i] = k->nk1[i];
  }
  s->ln_a1 = k->ln_a1;
  s->cell = k->cell;
  s->have_k1 = (uint8_t)k->have_k1;
}

static void pix_setup(pix* k, const orc_model* m, size_t p) {
  memset(k, 0, sizeof(*k));
  k->m = m;
  k->pixel = (uint32_t)p;
  k->theta_on = m->m32[p].theta_on;
  k->theta_off = m->m32[p].theta_off;
  k->leak = m->m32[p].leak;
  k->dark = m->m32[p].dark;
  k->scale = m->m32[p].scale;
  k->hot = m->m32[p].hot;
  k->G = m->G;
}

int orc_init(orc_model* m, int64_t t0, const float* irr0) {
  for (size_t i = 0; i < m->P * (size_t)m->G; ++i)
    if (!isfinite(irr0[i])) return 1;
  for (size_t p = 0; p < m->P; ++p) {
    pix k;
    pix_setup(&k, m, p);
    k.t_state = t0;
    k.n = 0;
    k.skip = 0;
    k.stats = &m->stats;
    float mean_I = k.dark;
    for (int g = 0; g < m->G; ++g) {
      const float v = clamp_irr(irr0[p * m->G + g], m->irr_max);
      if (!(v == irr0[p * m->G + g])) ++m->stats.input_clamps;
      k.held[g] = v;
      mean_I = mean_I + k.scale * v;
    }
    scan_input(&k);
    k.i_scale = mean_I;
    k.z = 1.0f / mean_I;
    k.x1 = 0.0f;
    k.s2 = k.s3 = 0.0f;
    k.tv = 0.0f;
    const float phi0 = m->m32[p].phi0;
    if (m->p.photoreceptor == BSX_PHOTORECEPTOR_INSTANTANEOUS) {
      set_reference(&k, t0, log_ratio(&k, photocurrent(&k, 0, t0)), 0.0f, phi0);
    } else {
      const float n3_0 = m->noise ? init_noise(&k, t0, mean_I) : 0.0f;
      set_reference(&k, t0, k.s3, n3_0, phi0);
    }
    k.blind_until = INT64_MIN;
    k.t_search = t0 + 1;
    k.flags = 1u << 12;
    k.shot_index = 0;
    k.shot_next = m->p.shot_rate_hz > 0.0 ? t0 + shot_gap(m, (uint32_t)p, 0, NULL) : INT64_MAX;
    store(&k, m, p);
  }
  m->t_state = t0;
  m->initialized = 1;
  m->n_ev = 0;
  return 0;
}


typedef struct {
  orc_model* m;
  const int32_t* row_of;
  const int64_t *t, *row_start;
  const float* irr;
  const int32_t* row_count;
  int64_t t_end;
  atomic_size_t next;
  evvec* outs;
  orc_stats* stats;
} job;

typedef struct {
  job* j;
  int w;
} worker_arg;

static void* worker(void* arg) {
  worker_arg* a = (worker_arg*)arg;
  job* j = a->j;
  orc_model* m = j->m;
  for (;;) {
    const size_t b = atomic_fetch_add(&j->next, 64);
    if (b >= m->P) break;
    const size_t e = b + 64 < m->P ? b + 64 : m->P;
    for (size_t p = b; p < e; ++p) {
      pix k;
      pix_setup(&k, m, p);
      k.t_state = m->t_state;
      const int32_t row = j->row_of[p];
      if (row >= 0) {
        k.t = j->t + j->row_start[row];
        k.irr = j->irr + (size_t)j->row_start[row] * m->G;
        k.n = j->row_count[row];
      }
      k.skip = k.n > 0 && k.t[0] == k.t_state ? 1 : 0;
      load(&k, m, p);
      scan_input(&k);
      k.out = &j->outs[a->w];
      k.stats = &j->stats[a->w];
      if (m->p.photoreceptor == BSX_PHOTORECEPTOR_INSTANTANEOUS) advance_instant(&k, j->t_end);
      else if (m->noise) advance_cascade_noise(&k, j->t_end);
      else advance_cascade(&k, j->t_end);
      if (k.n > 0)
        for (int g = 0; g < m->G; ++g) k.held[g] = s_value(&k, n_aug(&k) - 1, g);
      store(&k, m, p);
    }
  }
  return NULL;
}

static int cmp_ev(const void* a_, const void* b_) {
  const orc_event *a = (const orc_event*)a_, *b = (const orc_event*)b_;
  if (a->t_cross != b->t_cross) return a->t_cross < b->t_cross ? -1 : 1;
  return a->pixel < b->pixel ? -1 : (a->pixel > b->pixel ? 1 : 0);
}

int orc_advance_runs(orc_model* m, const uint32_t* ids, size_t n, const int64_t* t, const float* irr,
                     size_t n_samples, const int64_t* row_start, const int32_t* row_count, int64_t t_end,
                     char* err, size_t err_n) {
  if (!m->initialized) {
    set_err(err, err_n, "advance: not initialized");
    return 2;
  }
  if (t_end < m->t_state) {
    set_err(err, err_n, "advance: t_end_ns is before the model time");
    return 1;
  }
  int32_t* row_of = (int32_t*)malloc(m->P * sizeof(int32_t));
  for (size_t i = 0; i < m->P; ++i) row_of[i] = -1;
  uint64_t clamps = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t p = ids[i];
    const char* bad = NULL;
    if (p >= m->P) bad = "advance: pixel id out of range";
    else if (row_of[p] >= 0) bad = "advance: duplicate pixel id";
    else {
      row_of[p] = (int32_t)i;
      const int64_t s = row_start[i];
      const int32_t c = row_count[i];
      if (c < 0 || s < 0 || (size_t)s + (size_t)c > n_samples) bad = "advance: sample run out of range";
      for (int32_t jj = 0; bad == NULL && jj < c; ++jj)
        if (jj == 0 ? t[s] < m->t_state : t[s + jj] <= t[s + jj - 1])
          bad = "advance: sample times must be strictly increasing and start at or after the model time";
      if (bad == NULL && c > 0 && t[s + c - 1] > t_end) bad = "advance: sample after t_end_ns";
      for (size_t jj = 0; bad == NULL && jj < (size_t)c * m->G; ++jj) {
        const float v = irr[(size_t)s * m->G + jj];
        if (!isfinite(v)) bad = "advance: non-finite irradiance";
        else if (clamp_irr(v, m->irr_max) != v) ++clamps;
      }
    }
    if (bad != NULL) {
      free(row_of);
      set_err(err, err_n, bad);
      return 1;
    }
  }
  m->stats.input_clamps += clamps;
  int W = m->threads;
  if (W <= 0) {
    W = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (W <= 0) W = 1;
  }
  job j;
  j.m = m;
  j.row_of = row_of;
  j.t = t;
  j.row_start = row_start;
  j.irr = irr;
  j.row_count = row_count;
  j.t_end = t_end;
  atomic_init(&j.next, 0);
  j.outs = (evvec*)calloc((size_t)W, sizeof(evvec));
  j.stats = (orc_stats*)calloc((size_t)W, sizeof(orc_stats));
  pthread_t* th = (pthread_t*)malloc((size_t)W * sizeof(pthread_t));
  worker_arg* args = (worker_arg*)malloc((size_t)W * sizeof(worker_arg));
  for (int w = 0; w < W; ++w) {
    args[w].j = &j;
    args[w].w = w;
    if (W == 1) worker(&args[w]);
    else pthread_create(&th[w], NULL, worker, &args[w]);
  }
  if (W > 1)
    for (int w = 0; w < W; ++w) pthread_join(th[w], NULL);
  size_t total = 0;
  for (int w = 0; w < W; ++w) total += j.outs[w].n;
  if (total > m->cap_ev) {
    m->cap_ev = total;
    m->ev = (orc_event*)realloc(m->ev, total * sizeof(orc_event) + 1);
  }
  m->n_ev = 0;
  for (int w = 0; w < W; ++w) {
    if (j.outs[w].n) memcpy(m->ev + m->n_ev, j.outs[w].v, j.outs[w].n * sizeof(orc_event));
    m->n_ev += j.outs[w].n;
    free(j.outs[w].v);
    const orc_stats* s = &j.stats[w];
    m->stats.substeps += s->substeps;
    m->stats.substep_attempts += s->substep_attempts;
    m->stats.settled_substeps += s->settled_substeps;
    m->stats.leaf_intervals += s->leaf_intervals;
    m->stats.fallbacks += s->fallbacks;
    m->stats.quarantines += s->quarantines;
    m->stats.pieces += s->pieces;
    m->stats.cells += s->cells;
    m->stats.refinements += s->refinements;
  }
  qsort(m->ev, m->n_ev, sizeof(orc_event), cmp_ev);
  free(j.outs);
  free(j.stats);
  free(th);
  free(args);
  free(row_of);
  m->t_state = t_end;
  return 0;
}

size_t orc_event_count(const orc_model* m) { return m->n_ev; }
const orc_event* orc_events(const orc_model* m) { return m->ev; }

void orc_reported(const orc_model* m, int64_t* out) {
  for (size_t i = 0; i < m->n_ev; ++i) {
    int64_t t = m->ev[i].t_cross + m->delay_ns;
    if (m->q_ns > 0) {
      const int64_t d = t - m->phase_ns;
      int64_t q = d / m->q_ns;
      if (d % m->q_ns < 0) --q;
      t = m->phase_ns + m->q_ns * q;
    }
    out[i] = t + m->p.clock_offset_ns + llround(m->p.clock_drift_ppm * 1e-6 * (double)t);
  }
}

void orc_maps(const orc_model* m, double* a, double* b, double* c, double* d, double* e, uint8_t* hot,
              double* hot_rate) {
  for (size_t i = 0; i < m->P; ++i) {
    if (hot) hot[i] = m->hot[i];
    if (hot_rate) hot_rate[i] = m->hot_rate[i];
    if (a) a[i] = m->theta_on[i];
    if (b) b[i] = m->theta_off[i];
    if (c) c[i] = m->leak[i];
    if (d) d[i] = m->dark[i];
    if (e) e[i] = m->phi0[i];
  }
}

void orc_states(const orc_model* m, orc_state* out) { memcpy(out, m->st, m->P * sizeof(orc_state)); }
void orc_get_stats(const orc_model* m, orc_stats* out) { *out = m->stats; }
int64_t orc_time(const orc_model* m) { return m->t_state; }
