/*
 * oracle.c -- CPU FP64 restatement of Dynoplan's db-A* node-expansion path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path and
 * the timed CPU baseline.  Built with -O2 -ffp-contract=off, no fast-math,
 * so every expression rounds exactly as written.
 *
 * Reference lines followed:
 *   expansion loop          src/dbastar/dbastar.cpp:806-919
 *   check_lazy_trajectory   src/dbastar/dbastar.cpp:391-490
 *   check_goal              src/dbastar/dbastar.cpp:492-512
 *   Expander::expand_lazy   include/dynoplan/dbastar/dbastar.hpp:270-312
 *   metric spaces           include/dynoplan/nigh_custom_spaces.hpp:30-32,45-49,111-115,122-125,391-573
 *   nearestR                include/dynoplan/nigh_custom_spaces.hpp:236-247
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.141592653589793
#define OR_TWO_PI 6.283185307179586
#define OR_HALF_PI 1.5707963267948966

/* ------------------------------------------------------------------ */
/* series used by the quadrotor metric and dynamics                    */
/* ------------------------------------------------------------------ */

#define OR_ASIN_TERMS 30
#define OR_TRIG_TERMS 10
static double g_asin_c[OR_ASIN_TERMS];
static double g_cos_c[OR_TRIG_TERMS];
static double g_sinc_c[OR_TRIG_TERMS];
static int g_series_ready = 0;

static void series_init(void) {
  if (g_series_ready) return;
  /* asin(t) = sum_n a_n t^(2n+1) / (2n+1),  a_n = (2n)! / (4^n (n!)^2) */
  double a = 1.0;
  g_asin_c[0] = 1.0;
  for (int n = 1; n < OR_ASIN_TERMS; ++n) {
    a = a * (double)(2 * n - 1) / (double)(2 * n);
    g_asin_c[n] = a / (double)(2 * n + 1);
  }
  /* cos(x) = sum (-1)^n z^n/(2n)!,  sin(x)/x = sum (-1)^n z^n/(2n+1)!,  z = x^2 */
  g_cos_c[0] = 1.0;
  g_sinc_c[0] = 1.0;
  for (int n = 1; n < OR_TRIG_TERMS; ++n) {
    g_cos_c[n] = -g_cos_c[n - 1] / (double)((2 * n - 1) * (2 * n));
    g_sinc_c[n] = -g_sinc_c[n - 1] / (double)((2 * n) * (2 * n + 1));
  }
  g_series_ready = 1;
}

static double asin_series(double t) {
  double z = t * t;
  double p = g_asin_c[OR_ASIN_TERMS - 1];
  for (int n = OR_ASIN_TERMS - 2; n >= 0; --n) p = p * z + g_asin_c[n];
  return t * p;
}

/* acos on [0,1]: x >= 1/2 uses acos(x) = 2 asin(sqrt((1-x)/2)); otherwise
 * acos(x) = pi/2 - asin(x).  Both series arguments stay <= 1/2. */
double oracle_acos01(double x) {
  series_init();
  if (x >= 1.0) return 0.0;
  if (x >= 0.5) {
    double t = sqrt((1.0 - x) * 0.5);
    return 2.0 * asin_series(t);
  }
  return OR_HALF_PI - asin_series(x);
}

void oracle_cos_sinc(double z, double* c, double* sinc) {
  series_init();
  double pc = g_cos_c[OR_TRIG_TERMS - 1];
  double ps = g_sinc_c[OR_TRIG_TERMS - 1];
  for (int n = OR_TRIG_TERMS - 2; n >= 0; --n) {
    pc = pc * z + g_cos_c[n];
    ps = ps * z + g_sinc_c[n];
  }
  *c = pc;
  *sinc = ps;
}

/* ------------------------------------------------------------------ */
/* metric (nigh CartesianSpace of ScaledSpaces; sums in component order) */
/* ------------------------------------------------------------------ */

static double so2(double a, double b) {
  double d = fabs(a - b);
  return d > OR_PI ? OR_TWO_PI - d : d;
}

static double l2n(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    double d = a[i] - b[i];
    s = s + d * d;
  }
  return sqrt(s);
}

static double quat_angle(const double* a, const double* b) {
  double dot = a[0] * b[0];
  dot = dot + a[1] * b[1];
  dot = dot + a[2] * b[2];
  dot = dot + a[3] * b[3];
  dot = fabs(dot);
  if (dot > 1.0) dot = 1.0;
  return oracle_acos01(dot);
}

/* robot->distance, i.e. the nigh key distance of nigh_factory2
 * (nigh_custom_spaces.hpp:391-573): unicycle1 (:399-411), unicycle2
 * (:428-441), quad3d (:529-544), car1 (:547-556). */
double oracle_distance(const dp_system_desc* sys, const double* a, const double* b) {
  const double* w = sys->dist_w;
  switch (sys->system) {
    case DP_SYS_UNICYCLE1:
      return w[0] * l2n(a, b, 2) + w[1] * so2(a[2], b[2]);
    case DP_SYS_UNICYCLE2: {
      double d = w[0] * l2n(a, b, 2) + w[1] * so2(a[2], b[2]);
      d = d + w[2] * fabs(a[3] - b[3]);
      d = d + w[3] * fabs(a[4] - b[4]);
      return d;
    }
    case DP_SYS_CAR_TRAILER: {
      double d = w[0] * l2n(a, b, 2) + w[1] * so2(a[2], b[2]);
      d = d + w[2] * so2(a[3], b[3]);
      return d;
    }
    case DP_SYS_QUADROTOR: {
      double d = w[0] * l2n(a, b, 3) + w[1] * quat_angle(a + 3, b + 3);
      d = d + w[2] * l2n(a + 7, b + 7, 3);
      d = d + w[3] * l2n(a + 10, b + 10, 3);
      return d;
    }
  }
  return NAN;
}

static double dmax(double a, double b) { return b > a ? b : a; }

/* robot->lower_bound_time (dynobench, absent: parity unpinned). */
double oracle_lower_bound_time(const dp_system_desc* sys, const double* a, const double* b) {
  switch (sys->system) {
    case DP_SYS_UNICYCLE1:
      return dmax(l2n(a, b, 2) / sys->max_vel, so2(a[2], b[2]) / sys->max_angular_vel);
    case DP_SYS_UNICYCLE2: {
      double t = dmax(l2n(a, b, 2) / sys->max_vel, so2(a[2], b[2]) / sys->max_angular_vel);
      t = dmax(t, fabs(a[3] - b[3]) / sys->max_acc);
      t = dmax(t, fabs(a[4] - b[4]) / sys->max_angular_acc);
      return t;
    }
    case DP_SYS_CAR_TRAILER:
      return dmax(l2n(a, b, 2) / sys->max_vel, so2(a[2], b[2]) / sys->max_angular_vel);
    case DP_SYS_QUADROTOR:
      return dmax(l2n(a, b, 3) / sys->max_vel, l2n(a + 7, b + 7, 3) / sys->max_acc);
  }
  return NAN;
}

/* ------------------------------------------------------------------ */
/* bounds and collision                                                */
/* ------------------------------------------------------------------ */

/* is_state_valid: every component within [x_lb, x_ub]; position rows use
 * the environment bounds (robot_factory(..., p_lb, p_ub)). */
int oracle_state_valid(const dp_system_desc* sys, const dp_world* world, const double* s) {
  for (int c = 0; c < sys->nx; ++c) {
    double lo = c < sys->pos_dim ? world->lo[c] : sys->x_lb[c];
    double hi = c < sys->pos_dim ? world->hi[c] : sys->x_ub[c];
    if (!(s[c] >= lo && s[c] <= hi)) return 0;
  }
  return 1;
}

/* Separating-axis test, oriented box (centre cx,cy; axis (c,s); half
 * lengths hl along the axis, hw across) against an axis-aligned box (centre
 * ox,oy; half hx,hy).  Touching counts as collision. */
int oracle_obb_aabb(double cx, double cy, double c, double s, double hl, double hw,
                    double ox, double oy, double hx, double hy) {
  double tx = ox - cx;
  double ty = oy - cy;
  double ac = fabs(c);
  double as = fabs(s);
  if (fabs(tx) > hx + (hl * ac + hw * as)) return 0;
  if (fabs(ty) > hy + (hl * as + hw * ac)) return 0;
  double pu = tx * c + ty * s;
  if (fabs(pu) > hl + (hx * ac + hy * as)) return 0;
  double pv = ty * c - tx * s;
  if (fabs(pv) > hw + (hx * as + hy * ac)) return 0;
  return 1;
}

/* Footprint boxes of a planar state: (cx, cy, c, s, hl, hw) each. */
static int planar_boxes(const dp_system_desc* sys, const double* s, double* bx) {
  double c0 = cos(s[2]), s0 = sin(s[2]);
  bx[0] = s[0];
  bx[1] = s[1];
  bx[2] = c0;
  bx[3] = s0;
  bx[4] = sys->body_half[0];
  bx[5] = sys->body_half[1];
  if (sys->system != DP_SYS_CAR_TRAILER) return 1;
  double c1 = cos(s[3]), s1 = sin(s[3]);
  bx[6] = s[0] + -(sys->hitch * c1);
  bx[7] = s[1] + -(sys->hitch * s1);
  bx[8] = c1;
  bx[9] = s1;
  bx[10] = sys->trailer_half[0];
  bx[11] = sys->trailer_half[1];
  return 2;
}

static int sphere_box(const double* p, double r, const double* b) {
  double dx = fabs(p[0] - b[0]) - b[3];
  double dy = fabs(p[1] - b[1]) - b[4];
  double dz = fabs(p[2] - b[2]) - b[5];
  dx = dx > 0.0 ? dx : 0.0;
  dy = dy > 0.0 ? dy : 0.0;
  dz = dz > 0.0 ? dz : 0.0;
  double d2 = dx * dx + dy * dy;
  d2 = d2 + dz * dz;
  return d2 <= r * r;
}

static int sphere_sphere(const double* p, double r, const double* o) {
  double dx = p[0] - o[0], dy = p[1] - o[1], dz = p[2] - o[2];
  double d2 = dx * dx + dy * dy;
  d2 = d2 + dz * dz;
  double rr = r + o[3];
  return d2 <= rr * rr;
}

static int collides_subset(const dp_system_desc* sys, const dp_world* world, const double* s,
                           const int32_t* cand_b, int32_t ncb, const int32_t* cand_s, int32_t ncs) {
  if (sys->system == DP_SYS_QUADROTOR) {
    double r = sys->body_half[0];
    for (int32_t t = 0; t < ncb; ++t) {
      int32_t o = cand_b ? cand_b[t] : t;
      if (sphere_box(s, r, world->boxes + 6 * o)) return 1;
    }
    for (int32_t t = 0; t < ncs; ++t) {
      int32_t o = cand_s ? cand_s[t] : t;
      if (sphere_sphere(s, r, world->spheres + 4 * o)) return 1;
    }
    return 0;
  }
  double bx[12];
  int nb = planar_boxes(sys, s, bx);
  for (int b = 0; b < nb; ++b) {
    const double* q = bx + 6 * b;
    for (int32_t t = 0; t < ncb; ++t) {
      int32_t o = cand_b ? cand_b[t] : t;
      const double* ob = world->boxes + 6 * o;
      if (oracle_obb_aabb(q[0], q[1], q[2], q[3], q[4], q[5], ob[0], ob[1], ob[3], ob[4])) return 1;
    }
  }
  return 0;
}

int oracle_state_collides(const dp_system_desc* sys, const dp_world* world, const double* s) {
  return collides_subset(sys, world, s, NULL, world->n_boxes, NULL, world->n_spheres);
}

/* ------------------------------------------------------------------ */
/* dynamics and transform_primitive                                     */
/* ------------------------------------------------------------------ */

void oracle_unicycle1_step(const dp_system_desc* sys, const double* s, const double* u, double* o) {
  double dt = sys->dt;
  o[0] = s[0] + dt * (u[0] * cos(s[2]));
  o[1] = s[1] + dt * (u[0] * sin(s[2]));
  double th = s[2] + dt * u[1];
  if (th > OR_PI) th = th - OR_TWO_PI;
  if (th <= -OR_PI) th = th + OR_TWO_PI;
  o[2] = th;
}

/* Quadrotor, explicit Euler with body-frame quaternion exponential:
 *   p+ = p + dt v,   v+ = v + dt (R(q) e3 f + (0,0,-g)),
 *   w+ = w + dt J^-1 (tau - w x J w),   q+ = normalize(q (x) exp(w dt/2)),
 * f = (g/4) sum(u) (u = 1 per rotor hovers), rotor force F = m g/4 per unit
 * u, X-configuration torques.  State (p, qx qy qz qw, v, w). */
void oracle_quad_step(const dp_system_desc* sys, const double* s, const double* u, double* o) {
  const double m = sys->quad[0], g = sys->quad[1], arm = sys->quad[2], t2t = sys->quad[3];
  const double Jx = sys->quad[4], Jy = sys->quad[5], Jz = sys->quad[6];
  const double dt = sys->dt;
  const double qx = s[3], qy = s[4], qz = s[5], qw = s[6];
  const double vx = s[7], vy = s[8], vz = s[9];
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

  o[0] = s[0] + dt * vx;
  o[1] = s[1] + dt * vy;
  o[2] = s[2] + dt * vz;
  o[7] = vx + dt * ax;
  o[8] = vy + dt * ay;
  o[9] = vz + dt * az;
  o[10] = wx + dt * dwx;
  o[11] = wy + dt * dwy;
  o[12] = wz + dt * dwz;

  double h = 0.5 * dt;
  double z = (h * h) * ((wx * wx + wy * wy) + wz * wz);
  double ec, esinc;
  oracle_cos_sinc(z, &ec, &esinc);
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

static int is_offset_dim(const dp_system_desc* sys, int c) {
  if (sys->system == DP_SYS_QUADROTOR) return c <= 2 || (c >= 7 && c <= 9);
  return c <= 1;
}

/* Model_robot::canonical_state: the offset components zeroed. */
static void canonical_state(const dp_system_desc* sys, const double* x, double* c) {
  for (int i = 0; i < sys->nx; ++i) c[i] = is_offset_dim(sys, i) ? 0.0 : x[i];
}

int32_t oracle_transform_primitive(const dp_system_desc* sys, const dp_prims* prims, int32_t j,
                                   const double* node, double* out) {
  const int nx = sys->nx;
  const int32_t s0 = prims->offsets[j];
  const int32_t L1 = prims->offsets[j + 1] - s0;
  const double* st = prims->states + (int64_t)s0 * nx;
  if (sys->system != DP_SYS_QUADROTOR) {
    for (int32_t k = 0; k < L1; ++k)
      for (int c = 0; c < nx; ++c)
        out[k * nx + c] = is_offset_dim(sys, c) ? st[k * nx + c] + node[c] : st[k * nx + c];
    return L1;
  }
  for (int c = 0; c < nx; ++c) out[c] = is_offset_dim(sys, c) ? st[c] + node[c] : st[c];
  const double* act = prims->actions + (int64_t)(s0 - j) * prims->nu;
  for (int32_t k = 0; k + 1 < L1; ++k)
    oracle_quad_step(sys, out + k * nx, act + k * prims->nu, out + (k + 1) * nx);
  return L1;
}

/* check_goal index rule, dbastar.cpp:499-500:
 *   size_t index_to_check = float(nn + 1) / (num_check_goal + 1) * size; */
int32_t oracle_goal_index(int32_t nn, int32_t num_check_goal, int32_t size) {
  float r = (float)(nn + 1) / (float)(num_check_goal + 1);
  return (int32_t)(r * (float)size);
}

/* ------------------------------------------------------------------ */
/* nearest-neighbour sets (T_n)                                         */
/* ------------------------------------------------------------------ */

typedef struct {
  int nx, pos_dim;
  int64_t n, cap;
  double* pts;
  /* grid mode */
  int grid;
  double cell;
  int64_t nbuckets;
  int64_t* head;
  int64_t* next;
} nnset;

static int64_t cell_coord(double x, double cell) { return (int64_t)floor(x / cell); }

static uint64_t cell_hash(int64_t ix, int64_t iy, int64_t iz) {
  uint64_t h = (uint64_t)ix * 0x9E3779B97F4A7C15ull;
  h ^= (uint64_t)iy * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
  h ^= (uint64_t)iz * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
  return h;
}

static void nn_init(nnset* s, int nx, int pos_dim, int grid, double cell, int64_t cap) {
  memset(s, 0, sizeof(*s));
  s->nx = nx;
  s->pos_dim = pos_dim;
  s->cap = cap > 16 ? cap : 16;
  s->pts = (double*)malloc(sizeof(double) * s->cap * nx);
  s->grid = grid;
  s->cell = cell;
  if (grid) {
    s->nbuckets = 1;
    while (s->nbuckets < 2 * s->cap) s->nbuckets <<= 1;
    s->head = (int64_t*)malloc(sizeof(int64_t) * s->nbuckets);
    for (int64_t b = 0; b < s->nbuckets; ++b) s->head[b] = -1;
    s->next = (int64_t*)malloc(sizeof(int64_t) * s->cap);
  }
}

static void nn_free(nnset* s) {
  free(s->pts);
  free(s->head);
  free(s->next);
}

static void nn_add(nnset* s, const double* x) {
  if (s->n == s->cap) {
    s->cap *= 2;
    s->pts = (double*)realloc(s->pts, sizeof(double) * s->cap * s->nx);
    if (s->grid) {
      s->next = (int64_t*)realloc(s->next, sizeof(int64_t) * s->cap);
      /* rehash into a larger table */
      free(s->head);
      s->nbuckets <<= 1;
      s->head = (int64_t*)malloc(sizeof(int64_t) * s->nbuckets);
      for (int64_t b = 0; b < s->nbuckets; ++b) s->head[b] = -1;
      for (int64_t i = 0; i < s->n; ++i) {
        const double* p = s->pts + i * s->nx;
        int64_t iz = s->pos_dim > 2 ? cell_coord(p[2], s->cell) : 0;
        uint64_t b = cell_hash(cell_coord(p[0], s->cell), cell_coord(p[1], s->cell), iz) & (s->nbuckets - 1);
        s->next[i] = s->head[b];
        s->head[b] = i;
      }
    }
  }
  memcpy(s->pts + s->n * s->nx, x, sizeof(double) * s->nx);
  if (s->grid) {
    int64_t iz = s->pos_dim > 2 ? cell_coord(x[2], s->cell) : 0;
    uint64_t b = cell_hash(cell_coord(x[0], s->cell), cell_coord(x[1], s->cell), iz) & (s->nbuckets - 1);
    s->next[s->n] = s->head[b];
    s->head[b] = s->n;
  }
  s->n++;
}

/* nearestR(x, r) != empty */
static int nn_any_within(const dp_system_desc* sys, const nnset* s, const double* x, double r) {
  if (!s->grid) {
    for (int64_t i = 0; i < s->n; ++i)
      if (oracle_distance(sys, x, s->pts + i * s->nx) <= r) return 1;
    return 0;
  }
  int64_t cx = cell_coord(x[0], s->cell), cy = cell_coord(x[1], s->cell);
  int64_t cz = s->pos_dim > 2 ? cell_coord(x[2], s->cell) : 0;
  int zr = s->pos_dim > 2 ? 1 : 0;
  for (int64_t dx = -1; dx <= 1; ++dx)
    for (int64_t dy = -1; dy <= 1; ++dy)
      for (int64_t dz = -zr; dz <= zr; ++dz) {
        uint64_t b = cell_hash(cx + dx, cy + dy, cz + dz) & (s->nbuckets - 1);
        for (int64_t i = s->head[b]; i >= 0; i = s->next[i])
          if (oracle_distance(sys, x, s->pts + i * s->nx) <= r) return 1;
      }
  return 0;
}

/* ------------------------------------------------------------------ */
/* the batch                                                             */
/* ------------------------------------------------------------------ */

static void swept_candidates(const dp_system_desc* sys, const dp_world* world, const double* traj,
                             int32_t L1, int32_t* cb, int32_t* ncb, int32_t* cs, int32_t* ncs) {
  /* Conservative cull: obstacles whose AABB misses the trajectory's swept
   * footprint AABB (padded by 1e-9 m) cannot pass the SAT x/y axes. */
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const int nx = sys->nx;
  for (int32_t k = 0; k < L1; ++k) {
    const double* s = traj + k * nx;
    if (sys->system == DP_SYS_QUADROTOR) {
      double r = sys->body_half[0];
      for (int a = 0; a < 3; ++a) {
        if (s[a] - r < lo[a]) lo[a] = s[a] - r;
        if (s[a] + r > hi[a]) hi[a] = s[a] + r;
      }
    } else {
      double bx[12];
      int nb = planar_boxes(sys, s, bx);
      for (int b = 0; b < nb; ++b) {
        const double* q = bx + 6 * b;
        double ac = fabs(q[2]), as = fabs(q[3]);
        double ex = q[4] * ac + q[5] * as, ey = q[4] * as + q[5] * ac;
        if (q[0] - ex < lo[0]) lo[0] = q[0] - ex;
        if (q[0] + ex > hi[0]) hi[0] = q[0] + ex;
        if (q[1] - ey < lo[1]) lo[1] = q[1] - ey;
        if (q[1] + ey > hi[1]) hi[1] = q[1] + ey;
      }
      lo[2] = -INFINITY;
      hi[2] = INFINITY;
    }
  }
  const double pad = 1e-9;
  int32_t nb = 0, ns = 0;
  for (int32_t o = 0; o < world->n_boxes; ++o) {
    const double* b = world->boxes + 6 * o;
    int dims = sys->system == DP_SYS_QUADROTOR ? 3 : 2;
    int hit = 1;
    for (int a = 0; a < dims; ++a)
      if (b[a] + b[3 + a] < lo[a] - pad || b[a] - b[3 + a] > hi[a] + pad) hit = 0;
    if (hit) cb[nb++] = o;
  }
  for (int32_t o = 0; o < world->n_spheres; ++o) {
    const double* b = world->spheres + 4 * o;
    int hit = 1;
    for (int a = 0; a < 3; ++a)
      if (b[a] + b[3] < lo[a] - pad || b[a] - b[3] > hi[a] + pad) hit = 0;
    if (hit) cs[ns++] = o;
  }
  *ncb = nb;
  *ncs = ns;
}

/* The break of dbastar.cpp:916-918, taken after each FREE pair:
 * num_expansion_best_node (size_t, the NOVEL pairs so far) >=
 * limit_branching_factor (int, converted to size_t).  A limit of 0 breaks
 * after the first FREE pair whatever its resolution; a negative limit
 * converts to a huge value and never breaks. */
static int cap_reached(const dp_expand_params* par, int32_t novel_in_node) {
  return par->limit_branching_factor >= 0 && novel_in_node >= par->limit_branching_factor;
}

/* T_n->nearestR(x, radius) (nigh_custom_spaces.hpp:236-247), the rewiring
 * query of dbastar.cpp:883-912, by brute force: for each query, every
 * visited index v with distance <= radius, ascending (CSR).  Returns the
 * total number of hits; out_idx receives at most idx_cap of them. */
int64_t oracle_radius(const dp_system_desc* sys, const double* visited, int64_t nv, const double* queries,
                      int64_t n, double radius, int64_t* out_offsets, int64_t* out_idx, int64_t idx_cap) {
  const int nx = sys->nx;
  int64_t w = 0;
  out_offsets[0] = 0;
  for (int64_t q = 0; q < n; ++q) {
    for (int64_t v = 0; v < nv; ++v)
      if (oracle_distance(sys, queries + q * nx, visited + v * nx) <= radius) {
        if (out_idx && w < idx_cap) out_idx[w] = v;
        ++w;
      }
    out_offsets[q + 1] = w;
  }
  return w;
}

/* tdbastar's constraint test of one pair (tdbastar.cpp:417-443): the state
 * index a constraint at `time` selects in a placed primitive of L1 states
 * expanded from a node with gScore g (passed as float, tdbastar.cpp:317),
 * or -1 when it selects none.
 *   float time_offset = constraint.time - best_node_gScore;
 *   int time_index = std::lround(time_offset / robot.ref_dt);
 * The last state counts only when the pair reaches the goal. */
int32_t oracle_constraint_index(double time, double g, double dt, int32_t L1, int32_t reaches_goal) {
  const float time_offset = (float)(time - (double)(float)g);
  const int32_t ti = (int32_t)lround((double)time_offset / dt);
  int32_t k = -1;
  if (reaches_goal && ti >= L1 - 1) k = L1 - 1;
  if (ti >= 0 && ti < L1 - 1) k = ti;
  return k;
}

/* The primitive placed at the node.  Forward: transform_primitive with the
 * node's offset (dbastar.hpp:174-177).  Backward (tdbastar's reverse search,
 * tdbastar.hpp:193-196): with offset - last position, so that the placed
 * primitive ends at the node; planar systems (the reference assumes a 2-D
 * offset there). */
static int32_t place_primitive(const dp_system_desc* sys, const dp_prims* prims, int32_t j, const double* node,
                               int backward, double* out) {
  if (!backward) return oracle_transform_primitive(sys, prims, j, node, out);
  const double* last = prims->states + (int64_t)(prims->offsets[j + 1] - 1) * sys->nx;
  double o[DP_MAX_NX];
  for (int c = 0; c < sys->nx; ++c) o[c] = is_offset_dim(sys, c) ? node[c] - last[c] : node[c];
  return oracle_transform_primitive(sys, prims, j, o, out);
}

/* Applicability key of primitive j: its first state, or for the reverse
 * search its last state (nigh_factory_t, nigh_custom_spaces.hpp:576-586). */
static const double* app_key(const dp_system_desc* sys, const dp_prims* prims, int32_t j, int backward) {
  return prims->states + (int64_t)(backward ? prims->offsets[j + 1] - 1 : prims->offsets[j]) * sys->nx;
}

/* First violated constraint of a valid pair (tdbastar.cpp:408-443), in list
 * order: returns the state index it checked, -1 if none is violated. */
static int32_t constraint_violation(const dp_system_desc* sys, const oracle_ext* ext, const double* traj, int32_t L1,
                                    double g, const dp_expand_params* par) {
  const int nx = sys->nx;
  const double delta = (double)par->delta;
  const double* rg_state = traj + (int64_t)(ext->backward ? L1 - 1 : 0) * nx;
  const int32_t reaches_goal = oracle_distance(sys, rg_state, par->goal) <= delta;
  for (int32_t c = 0; c < ext->n_constraints; ++c) {
    const int32_t k = oracle_constraint_index(ext->con_times[c], g, sys->dt, L1, reaches_goal);
    if (k >= 0 && oracle_distance(sys, traj + (int64_t)k * nx, ext->con_states + (int64_t)c * nx) <= delta)
      return k;
  }
  return -1;
}

int64_t oracle_expand(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions,
                      const dp_world* world, const double* visited, int64_t nv,
                      const double* nodes, const double* g_parent, int32_t N,
                      const dp_expand_params* par, int32_t nn_mode, int32_t col_mode,
                      dp_pairs_out* out, dp_counts* counts) {
  return oracle_expand_ext(sys, prims, max_motions, world, visited, nv, nodes, g_parent, N, par, nn_mode, col_mode,
                           NULL, out, counts);
}

int64_t oracle_expand_ext(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions,
                          const dp_world* world, const double* visited, int64_t nv,
                          const double* nodes, const double* g_parent, int32_t N,
                          const dp_expand_params* par, int32_t nn_mode, int32_t col_mode,
                          const oracle_ext* ext, dp_pairs_out* out, dp_counts* counts) {
  if (!sys || !prims || !world || !nodes || !par || !counts) return DP_ERR_ARG;
  if (!(par->alpha > 0.0f && par->alpha < 1.0f)) return DP_ERR_ARG; /* dbastar.cpp:588-590 */
  if (sys->dist_w[0] <= 0.0) return DP_ERR_ARG;
  const int backward = ext && ext->backward;
  if (backward && sys->system == DP_SYS_QUADROTOR) return DP_ERR_ARG;
  if (ext && ext->order && ext->order_rows <= 0) return DP_ERR_ARG;
  series_init();
  const int nx = sys->nx;
  int32_t P = prims->P;
  if (max_motions > 0 && max_motions < P) P = max_motions;

  const double r_app = (double)(par->alpha * par->delta);               /* dbastar.cpp:585-586 */
  const double r_dup = (1. - (double)par->alpha) * (double)par->delta;  /* dbastar.cpp:852 */
  const double goal_bound = par->delta_factor_goal * (double)par->delta; /* dbastar.cpp:827 */

  int32_t maxlen = 0;
  for (int32_t j = 0; j < P; ++j) {
    int32_t l = prims->offsets[j + 1] - prims->offsets[j];
    if (l > maxlen) maxlen = l;
  }
  double* traj = (double*)malloc(sizeof(double) * (size_t)maxlen * nx);
  int32_t* cb = (int32_t*)malloc(sizeof(int32_t) * (world->n_boxes + 1));
  int32_t* cs = (int32_t*)malloc(sizeof(int32_t) * (world->n_spheres + 1));
  double canon[DP_MAX_NX];

  const double cell = (r_dup / sys->dist_w[0]) * (1.0 + 1e-9);
  nnset vis, acc;
  nn_init(&vis, nx, sys->pos_dim, nn_mode == ORACLE_NN_GRID, cell, nv + 16);
  for (int64_t i = 0; i < nv; ++i) nn_add(&vis, visited + i * nx);
  nn_init(&acc, nx, sys->pos_dim, nn_mode == ORACLE_NN_GRID, cell, 1024);

  memset(counts, 0, sizeof(*counts));
  counts->nodes = N;
  int64_t np = 0;
  int overflow = 0;

  for (int32_t i = 0; i < N; ++i) {
    const double* node = nodes + (int64_t)i * nx;
    canonical_state(sys, node, canon);
    int32_t novel_in_node = 0;
    int broke = 0; /* the reference's loop over this node's pairs has ended */
    const int32_t* row = ext && ext->order ? ext->order + (int64_t)(i % ext->order_rows) * P : NULL;
    for (int32_t r = 0; r < P; ++r) {
      const int32_t j = row ? row[r] : r; /* the shuffled applicable list, dbastar.hpp:302-303 */
      if (!(oracle_distance(sys, canon, app_key(sys, prims, j, backward)) <= r_app)) continue; /* nearestR, <= */
      int64_t q = np++;
      int wr = out && q < out->capacity;
      if (out && q >= out->capacity) overflow = 1;
      int32_t L1 = place_primitive(sys, prims, j, node, backward, traj);

      int8_t verdict = DP_V_FREE;
      int32_t first_k = -1;
      for (int32_t k = 0; k < L1; ++k)
        if (!oracle_state_valid(sys, world, traj + k * nx)) {
          first_k = k;
          verdict = DP_V_BOUNDS;
          break;
        }
      if (verdict == DP_V_FREE) {
        counts->checked++;
        int32_t ncb = world->n_boxes, ncs = world->n_spheres;
        const int32_t* pcb = NULL;
        const int32_t* pcs = NULL;
        if (col_mode == ORACLE_COL_SWEPT) {
          swept_candidates(sys, world, traj, L1, cb, &ncb, cs, &ncs);
          pcb = cb;
          pcs = cs;
        }
        for (int32_t k = 0; k < L1; ++k)
          if (collides_subset(sys, world, traj + k * nx, pcb, ncb, pcs, ncs)) {
            first_k = k;
            verdict = DP_V_COLLIDE;
            break;
          }
      }
      if (verdict == DP_V_FREE && ext && ext->n_constraints > 0) {
        const int32_t k = constraint_violation(sys, ext, traj, L1, g_parent[i], par);
        if (k >= 0) {
          first_k = k;
          verdict = DP_V_CONSTRAINT;
        }
      }
      int32_t chosen = -1;
      int8_t res = DP_R_NONE;
      double gq = NAN, hq = NAN;
      /* the new node's state: the last placed state, or the first for the
       * reverse search (check_goal, tdbastar.cpp:453-458) */
      const double* x = traj + (int64_t)(backward ? 0 : L1 - 1) * nx;
      if (verdict == DP_V_BOUNDS) counts->bounds++;
      if (verdict == DP_V_COLLIDE) counts->collide++;
      if (verdict == DP_V_CONSTRAINT) counts->constraint++;
      if (verdict == DP_V_FREE) {
        counts->free_++;
        if (par->check_intermediate_goal) { /* check_goal, dbastar.cpp:492-512 */
          for (int32_t nn = 0; nn < par->num_check_goal; ++nn) {
            int32_t idx = oracle_goal_index(nn, par->num_check_goal, L1);
            if (oracle_distance(sys, traj + (int64_t)idx * nx, par->goal) <= goal_bound) {
              x = traj + (int64_t)idx * nx;
              chosen = idx;
              break;
            }
          }
        }
        hq = oracle_lower_bound_time(sys, x, par->goal);
        double cost_motion = chosen != -1 ? chosen * sys->dt : (L1 - 1) * sys->dt;
        gq = g_parent[i] + cost_motion + par->cost_delta_factor * oracle_lower_bound_time(sys, node, traj);
        if (broke) {
          /* the reference left the loop (dbastar.cpp:916-918) before this
           * pair: it has no goal index, final state, g or h */
          res = DP_R_CAPPED;
          counts->capped++;
          chosen = -1;
          gq = hq = NAN;
        } else {
          if (chosen != -1) counts->goal_hits++;
          int dv = nn_any_within(sys, &vis, x, r_dup);
          int db = dv ? 0 : nn_any_within(sys, &acc, x, r_dup);
          if ((!dv && !db) || chosen != -1) { /* dbastar.cpp:856 */
            res = DP_R_NOVEL;
            counts->novel++;
            novel_in_node++;
            nn_add(&acc, x);
          } else if (dv) {
            res = DP_R_DUP_VISITED;
            counts->dup_visited++;
          } else {
            res = DP_R_DUP_BATCH;
            counts->dup_batch++;
          }
          broke = cap_reached(par, novel_in_node);
        }
      }
      if (wr) {
        if (out->node) out->node[q] = i;
        if (out->prim) out->prim[q] = j;
        if (out->verdict) out->verdict[q] = verdict;
        if (out->first_k) out->first_k[q] = first_k;
        if (out->goal_index) out->goal_index[q] = chosen;
        if (out->resolution) out->resolution[q] = res;
        if (out->final_state)
          for (int c = 0; c < nx; ++c)
            out->final_state[q * nx + c] = verdict == DP_V_FREE && res != DP_R_CAPPED ? x[c] : NAN;
        if (out->g) out->g[q] = gq;
        if (out->h) out->h[q] = hq;
      }
    }
  }
  counts->pairs = np;
  free(traj);
  free(cb);
  free(cs);
  nn_free(&vis);
  nn_free(&acc);
  return overflow ? DP_ERR_CAPACITY : np;
}

/* ------------------------------------------------------------------ */
/* dbrrt's expansion step                                                */
/* ------------------------------------------------------------------ */

/* For each node, dbrrt.cpp:580-660: the applicable primitives
 * (nearestR(canonical node, radius) over first states; Expander built with
 * options_dbrrt.delta, dbrrt.cpp:413-414) in the order of row i % order_rows
 * (ascending when order is NULL), each checked by check_lazy_trajectory
 * (bounds of every state, then collision; dbastar.cpp:391-490).  Among the
 * valid ones the first with the smallest distance(last state, target_i)
 * (strict <, dbrrt.cpp:627-642), or the first valid one with choose_first.
 * Per node: the primitive (-1 when none is valid), its pair index in the
 * batch's pair order, the distance, the last state (__expand_end), the first
 * state (__expand_start), g = g_parent + cost + cost_jump *
 * lower_bound_time(node, first) (:645-648; cost = prims->cost[j], or ref_dt
 * per action without one) and h = lower_bound_time(last, goal) (:641-642).
 * Returns the number of pairs. */
int64_t oracle_select(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions, const dp_world* world,
                      const double* nodes, const double* g_parent, int32_t N, double radius, const double* targets,
                      int32_t choose_first, double cost_jump, const double* goal, const int32_t* order,
                      int32_t order_rows, int32_t* out_prim, int64_t* out_pair, double* out_dist, double* out_state,
                      double* out_start, double* out_g, double* out_h) {
  if (!sys || !prims || !world || !nodes || !g_parent || !targets || !goal || N < 0) return DP_ERR_ARG;
  if (order && order_rows <= 0) return DP_ERR_ARG;
  series_init();
  const int nx = sys->nx;
  int32_t P = prims->P;
  if (max_motions > 0 && max_motions < P) P = max_motions;
  int32_t maxlen = 1;
  for (int32_t j = 0; j < P; ++j) {
    int32_t l = prims->offsets[j + 1] - prims->offsets[j];
    if (l > maxlen) maxlen = l;
  }
  double* traj = (double*)malloc(sizeof(double) * (size_t)maxlen * nx);
  int64_t np = 0;
  for (int32_t i = 0; i < N; ++i) {
    const double* node = nodes + (int64_t)i * nx;
    const double* target = targets + (int64_t)i * nx;
    double canon[DP_MAX_NX];
    canonical_state(sys, node, canon);
    const int32_t* row = order ? order + (int64_t)(i % order_rows) * P : NULL;
    double min_d = DBL_MAX;
    int32_t best = -1;
    int64_t best_q = -1;
    int done = 0;
    for (int32_t r = 0; r < P; ++r) {
      const int32_t j = row ? row[r] : r;
      if (!(oracle_distance(sys, canon, app_key(sys, prims, j, 0)) <= radius)) continue;
      const int64_t q = np++;
      if (done) continue; /* choose_first: the loop has ended (dbrrt.cpp:640-641) */
      const int32_t L1 = oracle_transform_primitive(sys, prims, j, node, traj);
      int valid = 1;
      for (int32_t k = 0; k < L1 && valid; ++k) valid = oracle_state_valid(sys, world, traj + (int64_t)k * nx);
      for (int32_t k = 0; k < L1 && valid; ++k) valid = !oracle_state_collides(sys, world, traj + (int64_t)k * nx);
      if (!valid) continue;
      const double* last = traj + (int64_t)(L1 - 1) * nx;
      const double d = oracle_distance(sys, last, target);
      if (d < min_d) {
        min_d = d;
        best = j;
        best_q = q;
        const double cost = prims->cost ? prims->cost[j] : (L1 - 1) * sys->dt;
        if (out_state) memcpy(out_state + (int64_t)i * nx, last, sizeof(double) * nx);
        if (out_start) memcpy(out_start + (int64_t)i * nx, traj, sizeof(double) * nx);
        if (out_g) out_g[i] = g_parent[i] + cost + cost_jump * oracle_lower_bound_time(sys, node, traj);
        if (out_h) out_h[i] = oracle_lower_bound_time(sys, last, goal);
        if (choose_first) done = 1;
      }
    }
    if (out_prim) out_prim[i] = best;
    if (out_pair) out_pair[i] = best_q;
    if (out_dist) out_dist[i] = best >= 0 ? min_d : NAN;
    if (best < 0) {
      for (int c = 0; c < nx; ++c) {
        if (out_state) out_state[(int64_t)i * nx + c] = NAN;
        if (out_start) out_start[(int64_t)i * nx + c] = NAN;
      }
      if (out_g) out_g[i] = NAN;
      if (out_h) out_h[i] = NAN;
    }
  }
  free(traj);
  return np;
}

/* ------------------------------------------------------------------ */
/* the same batch on several host threads (CPU baseline timing)          */
/* ------------------------------------------------------------------ */

/* Identical results to oracle_expand.  Phase 1 runs the per-pair work of
 * every node in parallel (applicability, translation, bounds, collision,
 * check_goal, g/h and the test against the visited set, which is read-only);
 * phase 2 replays the sequential part of dbastar.cpp:850-918 (branching cap,
 * in-batch duplicates, insertion) in canonical order on one thread. */
int64_t oracle_expand_mt(const dp_system_desc* sys, const dp_prims* prims, int32_t max_motions,
                         const dp_world* world, const double* visited, int64_t nv,
                         const double* nodes, const double* g_parent, int32_t N,
                         const dp_expand_params* par, int32_t nn_mode, int32_t col_mode,
                         dp_pairs_out* out, dp_counts* counts, int32_t nthreads) {
  if (!sys || !prims || !world || !nodes || !par || !counts || N < 0) return DP_ERR_ARG;
  if (!(par->alpha > 0.0f && par->alpha < 1.0f)) return DP_ERR_ARG;
  if (sys->dist_w[0] <= 0.0) return DP_ERR_ARG;
  series_init();
  const int nx = sys->nx;
  int32_t P = prims->P;
  if (max_motions > 0 && max_motions < P) P = max_motions;
  if (nthreads < 1) nthreads = 1;
  const double r_app = (double)(par->alpha * par->delta);
  const double r_dup = (1. - (double)par->alpha) * (double)par->delta;
  const double goal_bound = par->delta_factor_goal * (double)par->delta;
  int32_t maxlen = 0;
  for (int32_t j = 0; j < P; ++j) {
    int32_t l = prims->offsets[j + 1] - prims->offsets[j];
    if (l > maxlen) maxlen = l;
  }
  const double cell = (r_dup / sys->dist_w[0]) * (1.0 + 1e-9);
  nnset vis, acc;
  nn_init(&vis, nx, sys->pos_dim, nn_mode == ORACLE_NN_GRID, cell, nv + 16);
  for (int64_t i = 0; i < nv; ++i) nn_add(&vis, visited + i * nx);
  nn_init(&acc, nx, sys->pos_dim, nn_mode == ORACLE_NN_GRID, cell, 1024);

  /* applicable primitives per node */
  int64_t* first = (int64_t*)calloc((size_t)N + 1, sizeof(int64_t));
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
  for (int32_t i = 0; i < N; ++i) {
    double canon[DP_MAX_NX];
    canonical_state(sys, nodes + (int64_t)i * nx, canon);
    int64_t n = 0;
    for (int32_t j = 0; j < P; ++j)
      if (oracle_distance(sys, canon, prims->states + (int64_t)prims->offsets[j] * nx) <= r_app) ++n;
    first[i + 1] = n;
  }
  for (int32_t i = 0; i < N; ++i) first[i + 1] += first[i];
  const int64_t np = first[N];
  int32_t* pj = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np + 1));
  int8_t* vd = (int8_t*)malloc((size_t)(np + 1));
  int8_t* dv = (int8_t*)malloc((size_t)(np + 1));
  int32_t* fk = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np + 1));
  int32_t* ch = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np + 1));
  double* xf = (double*)malloc(sizeof(double) * (size_t)(np + 1) * nx);
  double* gg = (double*)malloc(sizeof(double) * (size_t)(np + 1));
  double* hh = (double*)malloc(sizeof(double) * (size_t)(np + 1));

#pragma omp parallel num_threads(nthreads)
  {
    double* traj = (double*)malloc(sizeof(double) * (size_t)(maxlen > 0 ? maxlen : 1) * nx);
    int32_t* cb = (int32_t*)malloc(sizeof(int32_t) * (world->n_boxes + 1));
    int32_t* cs = (int32_t*)malloc(sizeof(int32_t) * (world->n_spheres + 1));
#pragma omp for schedule(dynamic, 4)
    for (int32_t i = 0; i < N; ++i) {
      const double* node = nodes + (int64_t)i * nx;
      double canon[DP_MAX_NX];
      canonical_state(sys, node, canon);
      int64_t q = first[i];
      for (int32_t j = 0; j < P; ++j) {
        if (!(oracle_distance(sys, canon, prims->states + (int64_t)prims->offsets[j] * nx) <= r_app)) continue;
        int32_t L1 = oracle_transform_primitive(sys, prims, j, node, traj);
        int8_t verdict = DP_V_FREE;
        int32_t first_k = -1;
        for (int32_t k = 0; k < L1; ++k)
          if (!oracle_state_valid(sys, world, traj + k * nx)) {
            first_k = k;
            verdict = DP_V_BOUNDS;
            break;
          }
        if (verdict == DP_V_FREE) {
          int32_t ncb = world->n_boxes, ncs = world->n_spheres;
          const int32_t* pcb = NULL;
          const int32_t* pcs = NULL;
          if (col_mode == ORACLE_COL_SWEPT) {
            swept_candidates(sys, world, traj, L1, cb, &ncb, cs, &ncs);
            pcb = cb;
            pcs = cs;
          }
          for (int32_t k = 0; k < L1; ++k)
            if (collides_subset(sys, world, traj + k * nx, pcb, ncb, pcs, ncs)) {
              first_k = k;
              verdict = DP_V_COLLIDE;
              break;
            }
        }
        int32_t chosen = -1;
        const double* x = traj + (int64_t)(L1 - 1) * nx;
        double gq = NAN, hq = NAN;
        int8_t dupv = 0;
        if (verdict == DP_V_FREE) {
          if (par->check_intermediate_goal)
            for (int32_t nn = 0; nn < par->num_check_goal; ++nn) {
              int32_t idx = oracle_goal_index(nn, par->num_check_goal, L1);
              if (oracle_distance(sys, traj + (int64_t)idx * nx, par->goal) <= goal_bound) {
                x = traj + (int64_t)idx * nx;
                chosen = idx;
                break;
              }
            }
          hq = oracle_lower_bound_time(sys, x, par->goal);
          double cost_motion = chosen != -1 ? chosen * sys->dt : (L1 - 1) * sys->dt;
          gq = g_parent[i] + cost_motion + par->cost_delta_factor * oracle_lower_bound_time(sys, node, traj);
          dupv = (int8_t)nn_any_within(sys, &vis, x, r_dup);
        }
        pj[q] = j;
        vd[q] = verdict;
        fk[q] = first_k;
        ch[q] = chosen;
        dv[q] = dupv;
        gg[q] = gq;
        hh[q] = hq;
        for (int c = 0; c < nx; ++c) xf[q * nx + c] = verdict == DP_V_FREE ? x[c] : NAN;
        ++q;
      }
    }
    free(traj);
    free(cb);
    free(cs);
  }

  memset(counts, 0, sizeof(*counts));
  counts->nodes = N;
  counts->pairs = np;
  int overflow = 0;
  for (int32_t i = 0; i < N; ++i) {
    int32_t novel_in_node = 0;
    int broke = 0; /* the reference's loop over this node's pairs has ended */
    for (int64_t q = first[i]; q < first[i + 1]; ++q) {
      int8_t res = DP_R_NONE;
      if (vd[q] == DP_V_BOUNDS) counts->bounds++;
      if (vd[q] != DP_V_BOUNDS) counts->checked++;
      if (vd[q] == DP_V_COLLIDE) counts->collide++;
      if (vd[q] == DP_V_FREE) {
        const double* x = xf + q * nx;
        counts->free_++;
        if (broke) {
          res = DP_R_CAPPED;
          counts->capped++;
        } else {
          if (ch[q] != -1) counts->goal_hits++;
          int db = dv[q] ? 0 : nn_any_within(sys, &acc, x, r_dup);
          if ((!dv[q] && !db) || ch[q] != -1) {
            res = DP_R_NOVEL;
            counts->novel++;
            novel_in_node++;
            nn_add(&acc, x);
          } else if (dv[q]) {
            res = DP_R_DUP_VISITED;
            counts->dup_visited++;
          } else {
            res = DP_R_DUP_BATCH;
            counts->dup_batch++;
          }
          broke = cap_reached(par, novel_in_node);
        }
      }
      if (out && q >= out->capacity) overflow = 1;
      if (out && q < out->capacity) {
        if (out->node) out->node[q] = i;
        if (out->prim) out->prim[q] = pj[q];
        if (out->verdict) out->verdict[q] = vd[q];
        if (out->first_k) out->first_k[q] = fk[q];
        const int capped = res == DP_R_CAPPED;
        if (out->goal_index) out->goal_index[q] = capped ? -1 : ch[q];
        if (out->resolution) out->resolution[q] = res;
        if (out->final_state)
          for (int c = 0; c < nx; ++c) out->final_state[q * nx + c] = capped ? NAN : xf[q * nx + c];
        if (out->g) out->g[q] = capped ? NAN : gg[q];
        if (out->h) out->h[q] = capped ? NAN : hh[q];
      }
    }
  }
  free(first);
  free(pj);
  free(vd);
  free(dv);
  free(fk);
  free(ch);
  free(xf);
  free(gg);
  free(hh);
  nn_free(&vis);
  nn_free(&acc);
  return overflow ? DP_ERR_CAPACITY : np;
}
