// f32 sin/cos with the exact results of the host's libm.
//
// The reference computes f32::sin / f32::cos, which Rust std sends to glibc sinf/cosf; glibc's
// results are not correctly rounded (about 1-2% of inputs in [-0.8, 8] differ from the nearest
// float).  Kernels therefore evaluate sin/cos in double precision (argument reduction by a
// three-part pi/2 split and the fdlibm polynomial kernels, fixed operation order, no FMA) and
// round once to f32; the host compares that value with its own sinf/cosf for every float in the
// domain the bake uses and stores the inputs where they differ.  A kernel looks its input up in
// that sorted exception list (a bucket index on the top 16 bits plus a binary search).
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
// which = 0: sin, 1: cos.  Rounded once to f32.
DLI float trig_core(float xf, int which) {
  double x = (double)xf;
  double fn = floor(x * kInvPio2 + 0.5);
  double r = (x - fn * kPio2_1) - fn * kPio2_1t;
  int q = ((int)fn + which) & 3;
  double v;
  switch (q) {
    case 0: v = poly_sin(r); break;
    case 1: v = poly_cos(r); break;
    case 2: v = -poly_sin(r); break;
    default: v = -poly_cos(r); break;
  }
  return (float)v;
}

// Device view of one function's exception list.
struct TrigTable {
  const uint32_t* bucket;  // 65537 entries: first exception with (bits >> 16) >= b
  const uint32_t* keys;    // input bit patterns, ascending
  const uint32_t* vals;    // host libm result bits
};
struct TrigTables {
  TrigTable sin, cos;
  // One bit per float of the table domain, set when sin or cos of it is an exception; inputs
  // whose bit is clear skip the search.  Domain: bits [kDomPosLo, kDomPosHi] then
  // [kDomNegLo, kDomNegHi].
  const uint32_t* flags;
};
static constexpr uint32_t kDomPosLo = 0x37800000u, kDomPosHi = 0x41000000u;  // [2^-16, 8]
static constexpr uint32_t kDomNegLo = 0xB7800000u, kDomNegHi = 0xBF4CCCCDu;  // [-0.8, -2^-16]
static constexpr uint32_t kDomPosCount = kDomPosHi - kDomPosLo + 1;
static constexpr uint32_t kDomCount = kDomPosCount + (kDomNegHi - kDomNegLo + 1);
// Index of a float in the domain, or ~0u outside it.
DLI uint32_t trig_domain_index(uint32_t bits) {
  if (bits >= kDomPosLo && bits <= kDomPosHi) return bits - kDomPosLo;
  if (bits >= kDomNegLo && bits <= kDomNegHi) return kDomPosCount + (bits - kDomNegLo);
  return ~0u;
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

DLI float trig_lookup(const TrigTable& t, float x, float computed) {
  uint32_t bits = f2u(x);
  uint32_t b = bits >> 16;
  uint32_t lo = t.bucket[b], hi = t.bucket[b + 1];
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint32_t k = t.keys[mid];
    if (k == bits) return u2f(t.vals[mid]);
    if (k < bits) lo = mid + 1;
    else hi = mid;
  }
  return computed;
}
// Exceptions exist only inside the domain (outside it the tables are empty).
DLI bool trig_flagged(const TrigTables& t, float x) {
  uint32_t i = trig_domain_index(f2u(x));
  return i != ~0u && ((t.flags[i >> 5] >> (i & 31u)) & 1u);
}
DLI float lsinf(const TrigTables& t, float x) {
  float c = trig_core(x, 0);
  return trig_flagged(t, x) ? trig_lookup(t.sin, x, c) : c;
}
DLI float lcosf(const TrigTables& t, float x) {
  float c = trig_core(x, 1);
  return trig_flagged(t, x) ? trig_lookup(t.cos, x, c) : c;
}

}  // namespace dl
