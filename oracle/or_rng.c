/*
 * or_rng.c -- random number generators of the oracle.
 *
 *  - MT19937 exactly as source/RNG/mtprng.f90 computes it, including its
 *    64-bit storage quirk: mag01(1) = -1727483681_INT64 is sign-extended, so
 *    state words carry all-ones upper halves that leak into the first
 *    tempering shift (mtprng.f90:193-237).  Kept on purpose: the stream must
 *    be identical to the reference's to reproduce its golden values.
 *  - Marsaglia-Tsang ziggurat r4_nor / r4_nor_setup
 *    (source/RNG/randomnumbers.f90:330-509), including the single precision
 *    log() in the tail (randomnumbers.f90:398-399).
 *  - Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy
 *    as 1, 2, 3") and the Box-Muller transform used by the HIP kernels for
 *    their counter-based noise; restated here so the device stream can be
 *    checked on the CPU.
 */
#include "uasd_oracle.h"
#include <math.h>
#include <string.h>

const double OR_GAMA    = 1.760859644e11;
const double OR_K_BOLT  = 1.38064852e-23;
const double OR_MUB     = 9.274009994e-24;
const double OR_MRY     = 2.179872325e-21;
const double OR_PI      = 3.141592653589793;
/* dbl_tolerance=1e-14 is a default-real literal in parameters.f90:12 */
const double OR_DBL_TOLERANCE = (double)1e-14f;

/* ---------------------------------------------------------------- MT19937 */
#define MT_N 624
#define MT_M 397

void or_mt_init(or_mt *s, int32_t seed)                 /* mtprng.f90:118-139 */
{
  s->mt[0] = (uint64_t)(int64_t)seed;
  for (int i = 1; i < MT_N; ++i) {
    uint64_t p = s->mt[i - 1];
    s->mt[i] = 4294967295ULL & (1812433253ULL * (p ^ (p >> 30)) + (uint64_t)i);
  }
  s->mti = MT_N;
}

uint64_t or_mt_rand64(or_mt *s)                          /* mtprng.f90:185-238 */
{
  static const uint64_t mag01[2] = {0ULL, (uint64_t)(int64_t)-1727483681LL};
  const uint64_t UPPER = 2147483648ULL, LOWER = 2147483647ULL;
  const uint64_t TB = (uint64_t)(int64_t)-1658038656LL;
  const uint64_t TC = (uint64_t)(int64_t)-272236544LL;
  uint64_t r;
  if (s->mti >= MT_N) {
    int kk;
    for (kk = 0; kk <= MT_N - MT_M - 1; ++kk) {
      r = (s->mt[kk] & UPPER) | (s->mt[kk + 1] & LOWER);
      s->mt[kk] = s->mt[kk + MT_M] ^ (r >> 1) ^ mag01[r & 1ULL];
    }
    for (kk = MT_N - MT_M; kk <= MT_N - 2; ++kk) {
      r = (s->mt[kk] & UPPER) | (s->mt[kk + 1] & LOWER);
      s->mt[kk] = s->mt[kk + (MT_M - MT_N)] ^ (r >> 1) ^ mag01[r & 1ULL];
    }
    r = (s->mt[MT_N - 1] & UPPER) | (s->mt[0] & LOWER);
    s->mt[MT_N - 1] = s->mt[MT_M - 1] ^ (r >> 1) ^ mag01[r & 1ULL];
    s->mti = 0;
  }
  r = s->mt[s->mti];
  s->mti += 1;
  r = r ^ (r >> 11);
  r = 4294967295ULL & (r ^ ((r << 7) & TB));
  r = 4294967295ULL & (r ^ ((r << 15) & TC));
  r = r ^ (r >> 18);
  return r;
}

int32_t or_mt_rand(or_mt *s)                             /* mtprng.f90:241-262 */
{
  uint64_t x = or_mt_rand64(s);
  return (int32_t)(uint32_t)x;
}

double or_mt_real1(or_mt *s)                             /* mtprng.f90:286-301 */
{
  const double factor = 1.0 / 4294967295.0;
  return (double)or_mt_rand64(s) * factor;
}

double or_mt_real2(or_mt *s)                             /* mtprng.f90:304-319 */
{
  const double factor = 1.0 / 4294967296.0;
  return (double)or_mt_rand64(s) * factor;
}

/* --------------------------------------------------------------- ziggurat */
void or_zig_setup(or_zig *z, int32_t seed)        /* randomnumbers.f90:443-509 */
{
  const double m1 = 2147483648.0;
  const double vn = 9.91256303526217e-03;
  double dn = 3.442619855899, tn = 3.442619855899, q;
  or_mt_init(&z->st, seed);
  q = vn / exp(-0.5 * dn * dn);
  z->kn[0] = (int32_t)((dn / q) * m1);
  z->kn[1] = 0;
  z->wn[0] = q / m1;
  z->wn[127] = dn / m1;
  z->fn[0] = 1.0;
  z->fn[127] = exp(-0.5 * dn * dn);
  for (int i = 127; i >= 2; --i) {           /* Fortran indices i, i+1 */
    dn = sqrt(-2.0 * log(vn / dn + exp(-0.5 * dn * dn)));
    z->kn[i] = (int32_t)((dn / tn) * m1);    /* kn(i+1) */
    tn = dn;
    z->fn[i - 1] = exp(-0.5 * dn * dn);      /* fn(i)   */
    z->wn[i - 1] = dn / m1;                  /* wn(i)   */
  }
}

static inline int32_t f_abs32(int32_t h)  /* Fortran abs() wraps at -2^31 */
{
  return (h == INT32_MIN) ? INT32_MIN : (h < 0 ? -h : h);
}

double or_zig_nor(or_zig *z)                      /* randomnumbers.f90:330-438 */
{
  const double r = 3.442620;
  int32_t hz = or_mt_rand(&z->st);
  int32_t iz = hz & 127;
  double value, x, y;
  if (f_abs32(hz) < z->kn[iz]) {
    value = (double)hz * z->wn[iz];
  } else {
    for (;;) {
      if (iz == 0) {
        for (;;) {
          x = -0.2904764 * (double)logf((float)or_mt_real1(&z->st));
          y = -(double)logf((float)or_mt_real1(&z->st));
          if (x * x <= y + y) break;
        }
        value = (hz <= 0) ? (-r - x) : (r + x);
        break;
      }
      x = (double)hz * z->wn[iz];
      if (z->fn[iz] + or_mt_real1(&z->st) * (z->fn[iz - 1] - z->fn[iz]) <
          exp(-0.5 * x * x)) {
        value = x;
        break;
      }
      hz = or_mt_rand(&z->st);
      iz = hz & 127;
      if (f_abs32(hz) < z->kn[iz]) {
        value = (double)hz * z->wn[iz];
        break;
      }
    }
  }
  return value;
}

/* ------------------------------------------------------------ Philox4x32 */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double or_u32_to_open01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

static void counter_words(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                          uint32_t ens, uint32_t out[4])
{
  uint32_t ctr[4] = {site, ens, (uint32_t)step, (tag & 0xffu) | ((uint32_t)(step >> 32) << 8)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32_10(ctr, key, out);
}

void or_counter_words(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site, uint32_t ens,
                      uint32_t out[4])
{
  counter_words(seed, tag, step, site, ens, out);
}

/* Uniform on (0, 1) from the top 24 bits of a word, exact in single
 * precision (the FP32 MC engine's uniforms). */
float or_u24(uint32_t w) { return ((float)(w >> 8) + 0.5f) * 5.9604644775390625e-08f; }

/* Three N(0,1) deviates per (site, ensemble, step): Box-Muller on the four
 * Philox words, (u0,u1) -> z0,z1 and (u2,u3) -> z2. */
void or_noise3_f64(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                   uint32_t ens, double out[3])
{
  const double two_pi = 6.283185307179586;
  uint32_t w[4];
  counter_words(seed, tag, step, site, ens, w);
  double u0 = or_u32_to_open01(w[0]), u1 = or_u32_to_open01(w[1]);
  double u2 = or_u32_to_open01(w[2]), u3 = or_u32_to_open01(w[3]);
  double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
  double t0 = two_pi * u1, t1 = two_pi * u3;
  out[0] = r0 * cos(t0);
  out[1] = r0 * sin(t0);
  out[2] = r1 * cos(t1);
}

/* Thermal-field deviates of the LLG steps (both precisions of the HIP
 * kernels): Box-Muller in single precision with polynomial ln, sin and cos,
 * as the kernels compute it, operation for operation (IEEE adds, multiplies,
 * fmaf, sqrtf, rintf; no contraction), so the deviates agree bit for bit.
 * (The reference's ziggurat deviates are double precision, r4_nor,
 * randomnumbers.f90:379,389; the FP64 engine draws or_thermal3_f64.)
 * Polynomials fitted by least squares on
 * Chebyshev nodes (Lawson reweighting):
 *   ln(u):  u = m 2^e, m in [sqrt(1/2), sqrt(2)], x = m - 1,
 *           ln(1+x) = x + x^2 (-1/2 + x P(x))
 *   angle:  2 pi u = (pi/2) v, v in (0, 4), q = rint(v), f = v - q,
 *           sin((pi/2) f) = f S(f^2), cos((pi/2) f) = C(f^2). */
static float bm_ln(float u)
{
  uint32_t bits;
  memcpy(&bits, &u, 4);
  int ex = (int)(bits >> 23) - 127;
  uint32_t mb = (bits & 0x7fffffu) | 0x3f800000u;
  float m;
  memcpy(&m, &mb, 4);
  if (m > 1.41421354f) {
    m = m * 0.5f;
    ex += 1;
  }
  const float x = m - 1.0f;
  float p = -0.0830537900328636f;
  p = fmaf(p, x, 0.127988979220390f);
  p = fmaf(p, x, -0.129434198141098f);
  p = fmaf(p, x, 0.141902372241020f);
  p = fmaf(p, x, -0.166434884071350f);
  p = fmaf(p, x, 0.200018793344498f);
  p = fmaf(p, x, -0.250002801418304f);
  p = fmaf(p, x, 0.333333253860474f);
  const float t = fmaf(p, x, -0.5f);
  const float x2 = x * x;
  const float lnm = fmaf(t, x2, x);
  return fmaf((float)ex, 0.693147182464599609375f, lnm);
}

static void bm_sincos_q(float v, float *s, float *c)
{
  const float qf = rintf(v);
  const float f = v - qf;
  const int q = (int)qf & 3;
  const float f2 = f * f;
  float sp = 0.000158195078256540f;
  sp = fmaf(sp, f2, -0.00468126079067588f);
  sp = fmaf(sp, f2, 0.0796925798058510f);
  sp = fmaf(sp, f2, -0.645964086055756f);
  sp = fmaf(sp, f2, 1.57079637050629f);
  const float sn = sp * f;
  float cs = 0.000903362990356982f;
  cs = fmaf(cs, f2, -0.0208599362522364f);
  cs = fmaf(cs, f2, 0.253669172525406f);
  cs = fmaf(cs, f2, -1.23370051383972f);
  cs = fmaf(cs, f2, 1.0f);
  *s = q == 0 ? sn : q == 1 ? cs : q == 2 ? -sn : -cs;
  *c = q == 0 ? cs : q == 1 ? -sn : q == 2 ? -cs : sn;
}

float or_bm_ln(float u) { return bm_ln(u); }
void or_bm_sincos(float v, float out[2]) { bm_sincos_q(v, &out[0], &out[1]); }

void or_thermal3(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site, uint32_t ens,
                 float out[3])
{
  uint32_t w[4];
  counter_words(seed, tag, step, site, ens, w);
  const float s32 = 2.3283064365386963e-10f, s22 = 2.384185791015625e-07f;
  const float u0 = ((float)w[0] + 0.5f) * s32, u2 = ((float)w[2] + 0.5f) * s32;
  const float v1 = (float)((w[1] >> 9) * 2u + 1u) * s22, v3 = (float)((w[3] >> 9) * 2u + 1u) * s22;
  const float r0 = sqrtf(-2.0f * bm_ln(u0)), r1 = sqrtf(-2.0f * bm_ln(u2));
  float s0, c0, s1, c1;
  bm_sincos_q(v1, &s0, &c0);
  bm_sincos_q(v3, &s1, &c1);
  out[0] = r0 * c0;
  out[1] = r0 * s0;
  out[2] = r1 * c1;
}

/* Thermal-field deviates of the FP64 engine: Box-Muller in double precision
 * on 52-bit uniforms u = (x + 1/2) 2^-52, x the top 52 bits of a pair of
 * Philox words -- block 0x11 gives the radii (words 0,1 and 2,3), block 0x12
 * the angles, 2 pi u.  The kernels take the angles by sincospi / cospi;
 * libm's cos / sin of 2 pi u here agree with them to a few ulps. */
static double u52(uint32_t hi, uint32_t lo)
{
  const uint64_t x = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
  return ((double)x + 0.5) * 2.220446049250313080847e-16;
}

void or_thermal3_f64(uint64_t seed, uint64_t step, uint32_t site, uint32_t ens, double out[3])
{
  uint32_t w[4], x[4];
  counter_words(seed, 0x11u, step, site, ens, w);
  counter_words(seed, 0x12u, step, site, ens, x);
  const double r0 = sqrt(-2.0 * log(u52(w[0], w[1]))), r1 = sqrt(-2.0 * log(u52(w[2], w[3])));
  const double t0 = 2.0 * OR_PI * u52(x[0], x[1]), t1 = 2.0 * OR_PI * u52(x[2], x[3]);
  out[0] = r0 * cos(t0);
  out[1] = r0 * sin(t0);
  out[2] = r1 * cos(t1);
}

void or_uniform4(uint64_t seed, uint32_t tag, uint64_t step, uint32_t site,
                 uint32_t ens, double out[4])
{
  uint32_t w[4];
  counter_words(seed, tag, step, site, ens, w);
  for (int i = 0; i < 4; ++i) out[i] = or_u32_to_open01(w[i]);
}

void or_mt_fill_real2(or_mt *s, double *out, long n)
{
  for (long i = 0; i < n; ++i) out[i] = or_mt_real2(s);
}

void or_zig_fill(or_zig *z, double *out, long n)       /* fill_rngarray 548-565 */
{
  for (long i = 0; i < n; ++i) out[i] = or_zig_nor(z);
}
