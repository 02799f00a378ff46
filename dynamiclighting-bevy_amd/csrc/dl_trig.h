// f32 sin/cos with the exact results of the host's libm.
//
// The reference computes f32::sin / f32::cos, which Rust std sends to glibc sinf/cosf; glibc's
// results are not correctly rounded (about 1% of inputs in [-0.8, 8] differ from the nearest
// float).  Kernels therefore evaluate sin/cos in double precision (argument reduction by a
// two-part pi/2 split and the fdlibm polynomial kernels, fixed operation order, no FMA) and
// round once to f32; the host compares that value with its own sinf/cosf for every float in the
// domain the bake uses and stores the inputs where they differ.  Every such input lies close to
// a rounding midpoint of the double value (glibc is off by less than 0.07 ulp), so a kernel only
// looks its input up in a bitmap of the exceptions (with per-32-input ranks into their results)
// when the double value is within the host-measured band around a midpoint.
#pragma once

#include "dl_math.h"

namespace dl {

// fdlibm reduction constants: pi/2 split into a 33-bit head and the tails.
static constexpr double kInvPio2 = 6.36619772367581382433e-01;
static constexpr double kPio2_1 = 1.57079632673412561417e+00;
static constexpr double kPio2_1t = 6.07710050650619224932e-11;
// fdlibm __kernel_sin / __kernel_cos coefficients.
static constexpr double kS1 = -1.66666666666666324348e-01, kS2 = 8.33333333332248946124e-03,
                        kS3 = -1.98412698298579493134e-04, kS4 = 2.75573137070700676789e-06,
                        kS5 = -2.50507602534068634195e-08, kS6 = 1.58969099521155010221e-10;
static constexpr double kC1 = 4.16666666666666019037e-02, kC2 = -1.38888888888741095749e-03,
                        kC3 = 2.48015872894767294178e-05, kC4 = -2.75573143513906633035e-07,
                        kC5 = 2.08757232129817482790e-09, kC6 = -1.13596475577881948265e-11;

DLI double poly_sin(double x) {
  double z = x * x;
  double v = z * x;
  double r = kS2 + z * (kS3 + z * (kS4 + z * (kS5 + z * kS6)));
  return x + v * (kS1 + z * r);
}
DLI double poly_cos(double x) {
  double z = x * x;
  double r = z * (kC1 + z * (kC2 + z * (kC3 + z * (kC4 + z * (kC5 + z * kC6)))));
  return 1.0 - (0.5 * z - z * r);
}
// Argument reduction: x = fn * pi/2 + r with |r| <= pi/4 (for the domain's |x| <= 8).
DLI double trig_reduce(float xf, int* quadrant) {
  const double x = (double)xf;
  const double fn = floor(x * kInvPio2 + 0.5);
  *quadrant = (int)fn;
  return (x - fn * kPio2_1) - fn * kPio2_1t;
}
// sin (which = 0) or cos (which = 1) of the reduced argument, before rounding: both kernels are
// evaluated and the quadrant selects, so lanes do not diverge.
DLI double trig_select(double r, int quadrant, int which) {
  const double ps = poly_sin(r), pc = poly_cos(r);
  const int q = (quadrant + which) & 3;
  const double v = (q & 1) ? pc : ps;
  return (q & 2) ? -v : v;
}
DLI double trig_value(float xf, int which) {
  int q;
  const double r = trig_reduce(xf, &q);
  return trig_select(r, q, which);
}
// which = 0: sin, 1: cos.  Rounded once to f32.
DLI float trig_core(float xf, int which) { return (float)trig_value(xf, which); }

// Device view of one function's exception list: per run of 32 consecutive domain inputs one
// (bitmap of the exceptions in the run, exceptions in all earlier runs) pair, and the host libm
// results of the exceptions in domain order.
struct TrigTable {
  const uint2* runs;
  const uint32_t* vals;
  // bit r: run r of `runs` holds an exception (1.1 MB, small enough to stay in cache, so most
  // inputs near a midpoint are cleared without reading their run)
  const uint32_t* any;
};

struct TrigTables {
  TrigTable sin, cos;
  // Inputs whose double value lies farther than band * (midpoint gap) from both rounding
  // midpoints of its f32 rounding are never exceptions (checked on the host for every float of
  // the domain); only the others are searched.  Domain: bits [kDomPosLo, kDomPosHi] then
  // [kDomNegLo, kDomNegHi].
  double band;
};
static constexpr uint32_t kDomPosLo = 0x37800000u, kDomPosHi = 0x41000000u;  // [2^-16, 8]
static constexpr uint32_t kDomNegLo = 0xB7800000u, kDomNegHi = 0xBF4CCCCDu;  // [-0.8, -2^-16]
static constexpr uint32_t kDomPosCount = kDomPosHi - kDomPosLo + 1;
static constexpr uint32_t kDomCount = kDomPosCount + (kDomNegHi - kDomNegLo + 1);
static constexpr uint32_t kTrigRuns = (kDomCount + 31) / 32;
DLI bool trig_in_domain(uint32_t bits) {
  return (bits >= kDomPosLo && bits <= kDomPosHi) || (bits >= kDomNegLo && bits <= kDomNegHi);
}
// Position of an in-domain input in the domain (positive range first).
DLI uint32_t trig_domain_index(uint32_t bits) {
  return bits <= kDomPosHi ? bits - kDomPosLo : kDomPosCount + (bits - kDomNegLo);
}

DLI uint32_t f2u(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}
DLI float u2f(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// An in-domain input's libm result when it is an exception, else `computed`: one load tells
// whether it is (and its rank among the exceptions), a second one fetches the result.
DLI float trig_lookup(const TrigTable& t, float x, float computed) {
  const uint32_t di = trig_domain_index(f2u(x));
  if (!((t.any[di >> 10] >> ((di >> 5) & 31u)) & 1u)) return computed;
  const uint2 run = t.runs[di >> 5];
  const uint32_t bit = di & 31u;
  if (!((run.x >> bit) & 1u)) return computed;
  return u2f(t.vals[run.y + (uint32_t)__builtin_popcount(run.x & ((1u << bit) - 1u))]);
}
// Whether the double value v (rounded to f) lies within band * gap of a rounding midpoint,
// where gap is the distance between the midpoints below and above f.
DLI bool trig_near_midpoint(double v, float f, double band) {
  const uint32_t b = f2u(f);
  if ((b & 0x7fffffffu) == 0u || (b & 0x7f800000u) == 0x7f800000u) return true;
  const float away = u2f(b + 1u), toward = u2f(b - 1u);  // larger and smaller magnitude
  const double m1 = ((double)f + (double)away) * 0.5, m0 = ((double)f + (double)toward) * 0.5;
  const double lo = m0 < m1 ? m0 : m1, hi = m0 < m1 ? m1 : m0;
  const double a = v - lo, c = hi - v;
  return (a < c ? a : c) <= band * (hi - lo);
}
DLI float trig_finish(const TrigTables& t, const TrigTable& tab, float x, double v) {
  const float f = (float)v;
  if (!trig_in_domain(f2u(x)) || !trig_near_midpoint(v, f, t.band)) return f;
  return trig_lookup(tab, x, f);
}
DLI float lsinf(const TrigTables& t, float x) { return trig_finish(t, t.sin, x, trig_value(x, 0)); }
DLI float lcosf(const TrigTables& t, float x) { return trig_finish(t, t.cos, x, trig_value(x, 1)); }
// sin and cos of one argument with one reduction.
DLI void lsincosf(const TrigTables& t, float x, float* s, float* c) {
  int q;
  const double r = trig_reduce(x, &q);
  *s = trig_finish(t, t.sin, x, trig_select(r, q, 0));
  *c = trig_finish(t, t.cos, x, trig_select(r, q, 1));
}

}  // namespace dl
