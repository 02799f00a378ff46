// f32 sin/cos with the exact results of the host's libm.
//
// The reference computes f32::sin / f32::cos, which Rust std sends to glibc sinf/cosf; glibc's
// results are not correctly rounded (about 1% of inputs in [-0.8, 8] differ from the nearest
// float).  Kernels therefore evaluate sin/cos in double precision (argument reduction by a
// two-part pi/2 split and the fdlibm polynomial kernels in fused multiply-adds, a fixed
// operation sequence that the host and the device evaluate identically) and round once to f32;
// the host compares that value with its own sinf/cosf for every float in the domain the bake uses
// and stores the inputs where they differ.  Every such input lies close to a rounding midpoint of
// the double value (glibc is off by less than 0.07 ulp), so a kernel only looks its input up in a
// bitmap of the exceptions (with per-32-input ranks into their results) when the double value's
// bits below f32 precision lie within the host-measured distance of the midpoint pattern.
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

// x + x^3 (S1 + x^2 (S2 + ... + x^2 S6)) and 1 - x^2 / 2 + x^4 (C1 + ... + x^2 C5), by Horner's
// rule in fused multiply-adds (std::fma on the host, v_fma_f64 on the device: both exact).
DLI double poly_sin(double x) {
  const double z = x * x;
  const double v = z * x;
  const double r = fma(z, fma(z, fma(z, fma(z, kS6, kS5), kS4), kS3), kS2);
  return fma(v, fma(z, r, kS1), x);
}
DLI double poly_cos(double x) {
  const double z = x * x;
  const double r = fma(z, fma(z, fma(z, fma(z, fma(z, kC6, kC5), kC4), kC3), kC2), kC1);
  return fma(z * z, r, fma(-0.5, z, 1.0));
}
// Argument reduction: x = fn * pi/2 + r with |r| <= pi/4 (for the domain's |x| <= 8).
DLI double trig_reduce(float xf, int* quadrant) {
  const double x = (double)xf;
  const double fn = floor(fma(x, kInvPio2, 0.5));
  *quadrant = (int)fn;
  return fma(-fn, kPio2_1t, fma(-fn, kPio2_1, x));
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
  // Inputs whose double value has its 29 bits below f32 precision farther than `tol` from the
  // midpoint pattern 2^28 are never exceptions (checked on the host for every float of the
  // domain); only the others are searched.  Domain: bits [kDomPosLo, kDomPosHi] then
  // [kDomNegLo, kDomNegHi].
  uint32_t tol;
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
// Distance of the double v from the nearest rounding midpoint of its f32 rounding, in units of
// 2^-29 f32 ulp: the 29 mantissa bits below f32 precision against the midpoint pattern 2^28.  Only
// meaningful when v rounds to a normal float of its own binade; values below 2^-125 in magnitude
// (f32 denormal range and its edge) report 0, so they always count as near.
DLI uint32_t trig_midpoint_distance(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  if (((uint32_t)(b >> 52) & 0x7ffu) < 1023u - 125u) return 0u;
  const uint32_t low = (uint32_t)b & 0x1fffffffu;
  return low >= 0x10000000u ? low - 0x10000000u : 0x10000000u - low;
}
DLI bool trig_near_midpoint(double v, uint32_t tol) { return trig_midpoint_distance(v) <= tol; }
DLI float trig_finish(const TrigTables& t, const TrigTable& tab, float x, double v) {
  const float f = (float)v;
  if (!trig_in_domain(f2u(x)) || !trig_near_midpoint(v, t.tol)) return f;
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
