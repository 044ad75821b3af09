// dd.hpp -- double-double arithmetic (hi + lo), host and device.
//
// The reference's dd_add / dd_sub / dd_mul / dd_div / dd_sqrt
// (src/dd_func.f90:5-250; inline forms src/ddadd.h, ddsub.h, ddmul.h,
// dddiv.h): Knuth's TwoSum for the hi parts, the lo parts added left to
// right, and Dekker's product with the 2^27+1 split -- no fused
// multiply-add, which is why the library is compiled with
// -ffp-contract=off: every operation below rounds exactly as written.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace parcel {

struct dd {
  double hi, lo;
};

__host__ __device__ __forceinline__ dd mkdd(double h, double l) {
  dd r;
  r.hi = h;
  r.lo = l;
  return r;
}

__host__ __device__ __forceinline__ dd dd_add(dd a, dd b) {
  const double p1 = a.hi + b.hi;
  const double p2 = p1 - a.hi;
  double t1l = (a.hi - (p1 - p2)) + (b.hi - p2);
  const double t1h = p1;
  t1l = t1l + a.lo + b.lo;
  const double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  return mkdd(s1, t1l);
}

__host__ __device__ __forceinline__ dd dd_sub(dd a, dd b) {
  const double p1 = a.hi - b.hi;
  const double p2 = p1 - a.hi;
  double t1l = (a.hi - (p1 - p2)) - (b.hi + p2);
  const double t1h = p1;
  t1l = t1l + a.lo - b.lo;
  const double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  return mkdd(s1, t1l);
}

__host__ __device__ __forceinline__ dd dd_mul(dd a, dd b) {
  const double t1h = a.hi * b.hi;
  double u1 = 134217729.0 * a.hi;
  const double p1 = u1 - (u1 - a.hi);
  const double p2 = a.hi - p1;
  u1 = 134217729.0 * b.hi;
  const double r1 = u1 - (u1 - b.hi);
  const double r2 = b.hi - r1;
  double t1l = ((p1 * r1 - t1h) + p1 * r2 + p2 * r1) + p2 * r2;
  t1l = t1l + (a.hi * b.lo + a.lo * b.hi);
  const double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  return mkdd(s1, t1l);
}

__host__ __device__ __forceinline__ dd dd_div(dd a, dd b) {
  const double q1 = a.hi / b.hi;
  double t1h = b.hi * q1;
  double u1 = 134217729.0 * b.hi;
  const double p1 = u1 - (u1 - b.hi);
  const double p2 = b.hi - p1;
  u1 = 134217729.0 * q1;
  double r1 = u1 - (u1 - q1);
  const double r2 = q1 - r1;
  double t1l = ((p1 * r1 - t1h) + p1 * r2 + p2 * r1) + p2 * r2;
  t1l = t1l + (b.lo * q1);
  double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  t1h = s1;
  s1 = a.hi - t1h;
  r1 = s1 - a.hi;
  double s2 = (a.hi - (s1 - r1)) - (t1h + r1);
  s2 = s2 - t1l;
  s2 = s2 + a.lo;
  const double q2 = (s1 + s2) / b.hi;
  const double h = q1 + q2;
  return mkdd(h, q2 - (h - q1));
}

// dd_sqrt (dd_func.f90:146-250); a zero argument is not returned early in
// the reference (the computation continues and yields NaN).
__host__ __device__ __forceinline__ dd dd_sqrt(dd a) {
  if (a.hi < 0.0 || (a.hi == 0.0 && a.lo < 0.0)) return mkdd(0.0, 0.0);
  const double x = 1.0 / sqrt(a.hi);
  const double ax = a.hi * x;
  const dd y = dd_mul(mkdd(x, 0.0), mkdd(0.5, 0.0));
  const dd z = dd_mul(mkdd(ax, 0.0), mkdd(ax, 0.0));
  const dd t2 = dd_sub(a, z);
  const dd t1 = dd_mul(t2, y);
  return dd_add(t1, mkdd(ax, 0.0));
}

}  // namespace parcel
