// dp_device.h -- device-side system semantics of the expansion path.
//
// Every function here is the gfx950 counterpart of a dynobench::Model_robot
// method or nigh metric the reference calls on the hot path.  They are
// compiled with -ffp-contract=off and written as explicit sequences of
// IEEE operations so that FP64 results are bit-identical to any other
// implementation that rounds the same expression in the same order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dynoplan_amd.h"

#define DP_PI 3.141592653589793
#define DP_TWO_PI 6.283185307179586
#define DP_HALF_PI 1.5707963267948966

// Metric and geometry helpers are compiled for the host too: the host search
// loop evaluates the same expressions, rounding for rounding.
#define DP_HD __host__ __device__ __forceinline__

#define DP_ASIN_TERMS 30
#define DP_TRIG_TERMS 10

// System constants visible to every kernel.  Uploaded once per context.
struct dp_dev_sys {
  int32_t system, nx, nu, pos_dim;
  double dt;
  double lo[DP_MAX_NX], hi[DP_MAX_NX];  // bounds incl. world position bounds
  double w[4];
  double max_vel, max_w, max_acc, max_wacc;
  double body_hl, body_hw, trailer_hl, trailer_hw, hitch;
  double radius;
  double quad[8];
  double asin_c[DP_ASIN_TERMS];
  double cos_c[DP_TRIG_TERMS];
  double sinc_c[DP_TRIG_TERMS];
  float near_eps;
};

DP_HD double dp_so2(double a, double b) {
  double d = fabs(a - b);
  return d > DP_PI ? DP_TWO_PI - d : d;
}

DP_HD double dp_l2_2(double ax, double ay, double bx, double by) {
  double dx = ax - bx, dy = ay - by;
  double s = dx * dx;
  s = s + dy * dy;
  return sqrt(s);
}

DP_HD double dp_l2_3(const double* a, const double* b) {
  double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  double s = d0 * d0;
  s = s + d1 * d1;
  s = s + d2 * d2;
  return sqrt(s);
}

DP_HD double dp_asin_series(const dp_dev_sys& S, double t) {
  double z = t * t;
  double p = S.asin_c[DP_ASIN_TERMS - 1];
#pragma unroll
  for (int n = DP_ASIN_TERMS - 2; n >= 0; --n) p = p * z + S.asin_c[n];
  return t * p;
}

// acos on [0,1] through asin series whose argument stays <= 1/2.
DP_HD double dp_acos01(const dp_dev_sys& S, double x) {
  if (x >= 1.0) return 0.0;
  if (x >= 0.5) {
    double t = sqrt((1.0 - x) * 0.5);
    return 2.0 * dp_asin_series(S, t);
  }
  return DP_HALF_PI - dp_asin_series(S, x);
}

DP_HD double dp_quat_angle(const dp_dev_sys& S, const double* a, const double* b) {
  double dot = a[0] * b[0];
  dot = dot + a[1] * b[1];
  dot = dot + a[2] * b[2];
  dot = dot + a[3] * b[3];
  dot = fabs(dot);
  if (dot > 1.0) dot = 1.0;
  return dp_acos01(S, dot);
}

// robot->distance (nigh_custom_spaces.hpp:391-573 key spaces).
DP_HD double dp_distance(const dp_dev_sys& S, const double* a, const double* b) {
  switch (S.system) {
    case DP_SYS_UNICYCLE1:
      return S.w[0] * dp_l2_2(a[0], a[1], b[0], b[1]) + S.w[1] * dp_so2(a[2], b[2]);
    case DP_SYS_UNICYCLE2: {
      double d = S.w[0] * dp_l2_2(a[0], a[1], b[0], b[1]) + S.w[1] * dp_so2(a[2], b[2]);
      d = d + S.w[2] * fabs(a[3] - b[3]);
      d = d + S.w[3] * fabs(a[4] - b[4]);
      return d;
    }
    case DP_SYS_CAR_TRAILER: {
      double d = S.w[0] * dp_l2_2(a[0], a[1], b[0], b[1]) + S.w[1] * dp_so2(a[2], b[2]);
      d = d + S.w[2] * dp_so2(a[3], b[3]);
      return d;
    }
    default: {  // quadrotor
      double d = S.w[0] * dp_l2_3(a, b) + S.w[1] * dp_quat_angle(S, a + 3, b + 3);
      d = d + S.w[2] * dp_l2_3(a + 7, b + 7);
      d = d + S.w[3] * dp_l2_3(a + 10, b + 10);
      return d;
    }
  }
}

// dp_distance(S, a, b) <= r, decided with early exits.  The metric is a sum
// of non-negative terms added left to right; IEEE addition is monotone in
// each operand, so a partial sum (or the sum with a term left out) above r
// proves the full sum is above r.  When no exit applies the full sum is
// formed in dp_distance's exact operation order, so the decision is always
// identical to evaluating dp_distance.
DP_HD bool dp_within(const dp_dev_sys& S, const double* a, const double* b, double r) {
  if (S.system != DP_SYS_QUADROTOR) return dp_distance(S, a, b) <= r;
  const double A = S.w[0] * dp_l2_3(a, b);
  if (A > r) return false;
  const double C = S.w[2] * dp_l2_3(a + 7, b + 7);
  const double D = S.w[3] * dp_l2_3(a + 10, b + 10);
  if ((A + C) + D > r) return false;
  double d = A + S.w[1] * dp_quat_angle(S, a + 3, b + 3);
  d = d + C;
  d = d + D;
  return d <= r;
}

// The first term of dp_distance, the weighted position distance, by the same
// expression: when it alone exceeds r so does the distance (the other terms
// are non-negative and IEEE addition is monotone).
DP_HD double dp_pos_term(const dp_dev_sys& S, const double* a, const double* b) {
  if (S.system == DP_SYS_QUADROTOR) return S.w[0] * dp_l2_3(a, b);
  return S.w[0] * dp_l2_2(a[0], a[1], b[0], b[1]);
}

// dp_within(S, x, y, r) for y stored by component (y_k = col[k * stride + i]):
// the position is read first, the other components only when the position
// term does not already exceed r.
template <int NX>
__device__ __forceinline__ bool dp_within_col(const dp_dev_sys& S, const double* x, const double* col,
                                              int64_t stride, int64_t i, double r) {
  constexpr int PD = NX == 13 ? 3 : 2;
  double y[NX];
#pragma unroll
  for (int k = 0; k < PD; ++k) y[k] = col[(int64_t)k * stride + i];
  if (dp_pos_term(S, x, y) > r) return false;
#pragma unroll
  for (int k = PD; k < NX; ++k) y[k] = col[(int64_t)k * stride + i];
  return dp_within(S, x, y, r);
}

// A bound T on the squared position difference such that s > T, with s the
// sum of squares as dp_l2_2 (planar) or dp_l2_3 (quadrotor) forms it,
// implies w0 * sqrt(s) > r after rounding: the relative margin 1e-9 exceeds
// the few ulps that sqrt and the product can lose.
DP_HD double dp_pos_bound(const dp_dev_sys& S, double r) {
  if (!(S.w[0] > 0.0)) return INFINITY;
  const double t = r / S.w[0];
  return t * t * (1.0 + 1e-9);
}

// dp_within(S, x, y, r) for planar systems, deciding most pairs without the
// FP64 square root.  Rejections first (exact, see dp_within_col_pf); then the
// distance with the position term from an FP32 square root: its relative
// error (conversion plus sqrtf, under 3e-7) keeps the estimate within
// 3.1e-7 * r of the FP64 distance whenever s <= T, so an estimate outside
// r * (1 -+ 1e-6) decides the comparison; inside that band the FP64
// expression decides.  The answer is always dp_within's.
template <int NX>
__device__ __forceinline__ bool dp_within_planar_fast(const dp_dev_sys& S, const double* x, const double* y, double r,
                                                      double T) {
  static_assert(NX != 13, "planar systems only");
  const double dx = x[0] - y[0], dy = x[1] - y[1];
  double s = dx * dx;
  s = s + dy * dy;
  if (s > T) return false;
  const double th = S.w[1] * dp_so2(x[2], y[2]);
  if (th > r) return false;
  double rest = th;
  if constexpr (NX >= 5) {
    if (S.system == DP_SYS_UNICYCLE2) {
      rest = rest + S.w[2] * fabs(x[3] - y[3]);
      rest = rest + S.w[3] * fabs(x[4] - y[4]);
    }
  }
  if constexpr (NX >= 4) {
    if (S.system == DP_SYS_CAR_TRAILER) rest = rest + S.w[2] * dp_so2(x[3], y[3]);
  }
  const double est = S.w[0] * (double)sqrtf((float)s) + rest;
  if (est < r * (1.0 - 1e-6)) return true;
  if (est > r * (1.0 + 1e-6)) return false;
  return dp_within(S, x, y, r);
}

// dp_within(S, x, y, r) for the quadrotor, with a rejection that skips the
// FP64 square roots for most candidates: the position, velocity and rate
// terms estimated with FP32 square roots of the FP64 squared sums (as
// dp_l2_3 forms them).  Conversions, the weights in FP32, sqrtf and two FP32
// additions keep the estimate within a relative 1e-6 of the FP64 sum
// A + C + D of dp_within (all terms non-negative), so an estimate above
// r * (1 + 1e-5) proves (A + C) + D > r, where dp_within returns false.
// Otherwise dp_within decides.  The answer is always dp_within's.
__device__ __forceinline__ bool dp_within_quad_fast(const dp_dev_sys& S, const double* x, const double* y, double r) {
  auto sq3 = [](const double* a, const double* b) {
    const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    double s = d0 * d0;
    s = s + d1 * d1;
    s = s + d2 * d2;
    return s;
  };
  const float est = (float)S.w[0] * sqrtf((float)sq3(x, y)) + (float)S.w[2] * sqrtf((float)sq3(x + 7, y + 7)) +
                    (float)S.w[3] * sqrtf((float)sq3(x + 10, y + 10));
  if (est > (float)(r * (1.0 + 1e-5))) return false;
  return dp_within(S, x, y, r);
}

// dp_within_col with exact rejections ahead of it: the squared position
// difference above T (dp_pos_bound), and for planar systems the heading
// term w1 * so2 alone above r (dp_distance adds non-negative terms to it, and
// IEEE addition is monotone); then the FP32-estimate paths above.  None
// changes the answer; they skip the square roots for most candidates of a
// grid scan.
template <int NX>
__device__ __forceinline__ bool dp_within_col_pf(const dp_dev_sys& S, const double* x, const double* col,
                                                 int64_t stride, int64_t i, double r, double T) {
  if constexpr (NX != 13) {
    const double dx = x[0] - col[i], dy = x[1] - col[stride + i];
    double s = dx * dx;
    s = s + dy * dy;
    if (s > T) return false;
    if (S.w[1] * dp_so2(x[2], col[2 * stride + i]) > r) return false;
    double y[NX];
#pragma unroll
    for (int k = 0; k < NX; ++k) y[k] = col[(int64_t)k * stride + i];
    return dp_within_planar_fast<NX>(S, x, y, r, T);
  } else {
    // the squared position difference first (exact against T, dp_pos_bound),
    // the other components only when it passes
    double y[13];
#pragma unroll
    for (int k = 0; k < 3; ++k) y[k] = col[(int64_t)k * stride + i];
    const double d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2];
    double s = d0 * d0;
    s = s + d1 * d1;
    s = s + d2 * d2;
    if (s > T) return false;
#pragma unroll
    for (int k = 3; k < 13; ++k) y[k] = col[(int64_t)k * stride + i];
    return dp_within_quad_fast(S, x, y, r);
  }
}

DP_HD double dp_max(double a, double b) { return b > a ? b : a; }

// robot->lower_bound_time.
DP_HD double dp_lower_bound_time(const dp_dev_sys& S, const double* a, const double* b) {
  switch (S.system) {
    case DP_SYS_UNICYCLE1:
    case DP_SYS_CAR_TRAILER:
      return dp_max(dp_l2_2(a[0], a[1], b[0], b[1]) / S.max_vel, dp_so2(a[2], b[2]) / S.max_w);
    case DP_SYS_UNICYCLE2: {
      double t = dp_max(dp_l2_2(a[0], a[1], b[0], b[1]) / S.max_vel, dp_so2(a[2], b[2]) / S.max_w);
      t = dp_max(t, fabs(a[3] - b[3]) / S.max_acc);
      t = dp_max(t, fabs(a[4] - b[4]) / S.max_wacc);
      return t;
    }
    default:
      return dp_max(dp_l2_3(a, b) / S.max_vel, dp_l2_3(a + 7, b + 7) / S.max_acc);
  }
}

DP_HD bool dp_is_offset_dim(const dp_dev_sys& S, int c) {
  if (S.system == DP_SYS_QUADROTOR) return c <= 2 || (c >= 7 && c <= 9);
  return c <= 1;
}

// Separating-axis test, oriented box vs axis-aligned box; touching collides.
DP_HD bool dp_obb_aabb(double cx, double cy, double c, double s, double hl, double hw,
                                            double ox, double oy, double hx, double hy) {
  double tx = ox - cx;
  double ty = oy - cy;
  double ac = fabs(c);
  double as = fabs(s);
  if (fabs(tx) > hx + (hl * ac + hw * as)) return false;
  if (fabs(ty) > hy + (hl * as + hw * ac)) return false;
  double pu = tx * c + ty * s;
  if (fabs(pu) > hl + (hx * ac + hy * as)) return false;
  double pv = ty * c - tx * s;
  if (fabs(pv) > hw + (hx * as + hy * ac)) return false;
  return true;
}

DP_HD bool dp_sphere_box(double px, double py, double pz, double r, const double* b) {
  double dx = fabs(px - b[0]) - b[3];
  double dy = fabs(py - b[1]) - b[4];
  double dz = fabs(pz - b[2]) - b[5];
  dx = dx > 0.0 ? dx : 0.0;
  dy = dy > 0.0 ? dy : 0.0;
  dz = dz > 0.0 ? dz : 0.0;
  double d2 = dx * dx + dy * dy;
  d2 = d2 + dz * dz;
  return d2 <= r * r;
}

DP_HD bool dp_sphere_sphere(double px, double py, double pz, double r, const double* o) {
  double dx = px - o[0], dy = py - o[1], dz = pz - o[2];
  double d2 = dx * dx + dy * dy;
  d2 = d2 + dz * dz;
  double rr = r + o[3];
  return d2 <= rr * rr;
}

// Wave-level helpers (wave64).
// Quadrotor step (explicit Euler for position and velocity, body-frame
// exp-map quaternion update, X-config torques; u = 1 is per-rotor hover) in
// FP64, one fixed operation order.
DP_HD void dp_quad_cos_sinc(const dp_dev_sys& S, double z, double* c, double* sinc) {
  double pc = S.cos_c[DP_TRIG_TERMS - 1];
  double ps = S.sinc_c[DP_TRIG_TERMS - 1];
#pragma unroll
  for (int n = DP_TRIG_TERMS - 2; n >= 0; --n) {
    pc = pc * z + S.cos_c[n];
    ps = ps * z + S.sinc_c[n];
  }
  *c = pc;
  *sinc = ps;
}

// Attitude and angular-rate part of the step, and the velocity increment
// dt * (thrust acceleration) it implies.  It reads only the quaternion and
// the angular rate (s[3..6], s[10..12]) and the rotors, so for a primitive
// rolled out from any translated start it is the same sequence: it is
// evaluated once per primitive and shared by every pair.
// Writes o[3..6], o[10..12] and dv[3] (o[7 + c] = s[7 + c] + dv[c]).
DP_HD void dp_quad_rot_step64(const dp_dev_sys& S, const double* s, const double* u, double* o, double* dv) {
  const double m = S.quad[0], g = S.quad[1], arm = S.quad[2], t2t = S.quad[3];
  const double Jx = S.quad[4], Jy = S.quad[5], Jz = S.quad[6];
  const double dt = S.dt;
  const double qx = s[3], qy = s[4], qz = s[5], qw = s[6];
  const double wx = s[10], wy = s[11], wz = s[12];
  double g4 = g * 0.25;
  double usum = ((u[0] + u[1]) + u[2]) + u[3];
  double f = g4 * usum;
  double F = m * g4;
  double tx = (arm * F) * (((u[2] + u[3]) - u[0]) - u[1]);
  double ty = (arm * F) * (((u[1] + u[2]) - u[0]) - u[3]);
  double tz = (t2t * F) * (((u[1] + u[3]) - u[0]) - u[2]);
  double rzx = 2.0 * (qx * qz + qw * qy);
  double rzy = 2.0 * (qy * qz - qw * qx);
  double rzz = 1.0 - 2.0 * (qx * qx + qy * qy);
  double ax = f * rzx;
  double ay = f * rzy;
  double az = f * rzz - g;
  double jwx = Jx * wx, jwy = Jy * wy, jwz = Jz * wz;
  double cx = wy * jwz - wz * jwy;
  double cy = wz * jwx - wx * jwz;
  double cz = wx * jwy - wy * jwx;
  double dwx = (tx - cx) / Jx;
  double dwy = (ty - cy) / Jy;
  double dwz = (tz - cz) / Jz;
  dv[0] = dt * ax;
  dv[1] = dt * ay;
  dv[2] = dt * az;
  o[10] = wx + dt * dwx;
  o[11] = wy + dt * dwy;
  o[12] = wz + dt * dwz;
  double h = 0.5 * dt;
  double z = (h * h) * ((wx * wx + wy * wy) + wz * wz);
  double ec, esinc;
  dp_quad_cos_sinc(S, z, &ec, &esinc);
  double k = h * esinc;
  double ex = k * wx, ey = k * wy, ez = k * wz, ew = ec;
  double nw = ((qw * ew - qx * ex) - qy * ey) - qz * ez;
  double nx = ((qw * ex + qx * ew) + qy * ez) - qz * ey;
  double ny = ((qw * ey - qx * ez) + qy * ew) + qz * ex;
  double nz = ((qw * ez + qx * ey) - qy * ex) + qz * ew;
  double n2 = ((nx * nx + ny * ny) + nz * nz) + nw * nw;
  double inv = 1.0 / sqrt(n2);
  o[3] = nx * inv;
  o[4] = ny * inv;
  o[5] = nz * inv;
  o[6] = nw * inv;
}

// Translation part of the step: explicit Euler with the old velocity, then
// the velocity increment of dp_quad_rot_step64.
DP_HD void dp_quad_trans_step64(double dt, const double* s, const double* dv, double* o) {
  o[0] = s[0] + dt * s[7];
  o[1] = s[1] + dt * s[8];
  o[2] = s[2] + dt * s[9];
  o[7] = s[7] + dv[0];
  o[8] = s[8] + dv[1];
  o[9] = s[9] + dv[2];
}

DP_HD void dp_quad_step64(const dp_dev_sys& S, const double* s, const double* u, double* o) {
  double dv[3];
  dp_quad_rot_step64(S, s, u, o, dv);
  dp_quad_trans_step64(S.dt, s, dv, o);
}

__device__ __forceinline__ uint64_t dp_ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int dp_lane() { return __lane_id(); }
