/*
 * Deterministic synthetic sensor workloads: scenes of axis-aligned boxes,
 * vertical cylinders, a ground plane or a closed room, ray-cast exactly in
 * double precision, plus Gaussian range noise. Used to build the scans that
 * the mapping path is tested and measured on (lidar rings x columns and
 * pinhole depth frames). Output points are float32 x,y,z plus a pad of 1.0
 * (16-byte stride, the pcl::PointXYZ layout the drop-in API hands over).
 *
 * Randomness: splitmix64-seeded xoshiro256**, Box-Muller normals. The same
 * seed gives the same bytes on any x86-64 host with the same libm.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define WL_MAX_BOX 512
#define WL_MAX_CYL 512

typedef struct {
  double lo[3], hi[3];
} wl_box;

typedef struct {
  double cx, cy, r, z0, z1;
} wl_cyl;

typedef struct {
  int32_t nbox, ncyl;
  int32_t has_room, has_ground;
  wl_box room;
  double ground_z;
  double sky_range; /* range reported when a ray hits nothing */
  wl_box box[WL_MAX_BOX];
  wl_cyl cyl[WL_MAX_CYL];
} wl_scene;

typedef struct {
  uint64_t s[4];
} rng_t;

static uint64_t splitmix(uint64_t* x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void rng_seed(rng_t* r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(rng_t* r) {
  uint64_t* s = r->s;
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

static double rng_unit(rng_t* r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static double rng_unif(rng_t* r, double a, double b) { return a + (b - a) * rng_unit(r); }
static double rng_normal(rng_t* r) {
  double u1 = rng_unit(r);
  const double u2 = rng_unit(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

int32_t wl_scene_bytes(void) { return (int32_t)sizeof(wl_scene); }
uint64_t wl_uniform_u64(uint64_t seed) {
  rng_t r;
  rng_seed(&r, seed);
  return rng_next(&r);
}

/* Slab intersection from outside: entry time > 0, or +inf. */
static double hit_box(const wl_box* b, const double o[3], const double d[3]) {
  double t0 = 0.0, t1 = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0.0) {
      if (o[a] < b->lo[a] || o[a] > b->hi[a]) return INFINITY;
      continue;
    }
    double ta = (b->lo[a] - o[a]) / d[a], tb = (b->hi[a] - o[a]) / d[a];
    if (ta > tb) {
      const double t = ta;
      ta = tb;
      tb = t;
    }
    if (ta > t0) t0 = ta;
    if (tb < t1) t1 = tb;
    if (t0 > t1) return INFINITY;
  }
  return t0 > 0.0 ? t0 : INFINITY;
}

/* Exit time of a ray starting inside a box. */
static double exit_box(const wl_box* b, const double o[3], const double d[3]) {
  double t = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (d[a] > 0) {
      const double ta = (b->hi[a] - o[a]) / d[a];
      if (ta < t) t = ta;
    } else if (d[a] < 0) {
      const double ta = (b->lo[a] - o[a]) / d[a];
      if (ta < t) t = ta;
    }
  }
  return t;
}

static double hit_cyl(const wl_cyl* c, const double o[3], const double d[3]) {
  double best = INFINITY;
  const double px = o[0] - c->cx, py = o[1] - c->cy;
  const double a = d[0] * d[0] + d[1] * d[1];
  if (a > 0) {
    const double b = 2.0 * (px * d[0] + py * d[1]);
    const double cc = px * px + py * py - c->r * c->r;
    const double disc = b * b - 4 * a * cc;
    if (disc >= 0) {
      const double t = (-b - sqrt(disc)) / (2 * a);
      if (t > 0) {
        const double z = o[2] + t * d[2];
        if (z >= c->z0 && z <= c->z1) best = t;
      }
    }
  }
  if (d[2] != 0) {
    const double cap = d[2] < 0 ? c->z1 : c->z0;
    const double t = (cap - o[2]) / d[2];
    if (t > 0 && t < best) {
      const double x = px + t * d[0], y = py + t * d[1];
      if (x * x + y * y <= c->r * c->r) best = t;
    }
  }
  return best;
}

/* Range along unit direction d from o: nearest surface, or sky_range. */
static double cast(const wl_scene* s, const double o[3], const double d[3]) {
  double best = INFINITY;
  if (s->has_room) best = exit_box(&s->room, o, d);
  if (s->has_ground && d[2] < 0) {
    const double t = (s->ground_z - o[2]) / d[2];
    if (t > 0 && t < best) best = t;
  }
  const double d2 = sqrt(d[0] * d[0] + d[1] * d[1]);
  for (int i = 0; i < s->nbox; ++i) {
    const wl_box* b = &s->box[i];
    if (d2 > 1e-6) {
      /* 2-D rejection against the box's bounding circle */
      const double cx = 0.5 * (b->lo[0] + b->hi[0]) - o[0], cy = 0.5 * (b->lo[1] + b->hi[1]) - o[1];
      const double hx = 0.5 * (b->hi[0] - b->lo[0]), hy = 0.5 * (b->hi[1] - b->lo[1]);
      const double rad = sqrt(hx * hx + hy * hy);
      const double along = (cx * d[0] + cy * d[1]) / d2;
      const double perp = fabs(cx * d[1] - cy * d[0]) / d2;
      if (perp > rad || along < -rad || along - rad > best * d2) continue;
    }
    const double t = hit_box(b, o, d);
    if (t < best) best = t;
  }
  for (int i = 0; i < s->ncyl; ++i) {
    const wl_cyl* c = &s->cyl[i];
    if (d2 > 1e-6) {
      const double cx = c->cx - o[0], cy = c->cy - o[1];
      const double along = (cx * d[0] + cy * d[1]) / d2;
      const double perp = fabs(cx * d[1] - cy * d[0]) / d2;
      if (perp > c->r || along < -c->r || along - c->r > best * d2) continue;
    }
    const double t = hit_cyl(c, o, d);
    if (t < best) best = t;
  }
  return isfinite(best) ? best : s->sky_range;
}

/* 20 x 20 x 5 m room around the origin (floor at z = -1.5) with nbox boxes
 * of sides in [side_min, side_max] standing on the floor, none within
 * 0.5 m (xy) of the origin. */
void wl_scene_room(uint64_t seed, int32_t nbox, double side_min, double side_max, wl_scene* s) {
  memset(s, 0, sizeof *s);
  rng_t r;
  rng_seed(&r, seed);
  s->has_room = 1;
  s->room.lo[0] = -10;
  s->room.lo[1] = -10;
  s->room.lo[2] = -1.5;
  s->room.hi[0] = 10;
  s->room.hi[1] = 10;
  s->room.hi[2] = 3.5;
  s->sky_range = 1e3;
  if (nbox > WL_MAX_BOX) nbox = WL_MAX_BOX;
  while (s->nbox < nbox) {
    const double sx = rng_unif(&r, side_min, side_max), sy = rng_unif(&r, side_min, side_max),
                 sz = rng_unif(&r, side_min, side_max);
    const double cx = rng_unif(&r, -9.0, 9.0), cy = rng_unif(&r, -9.0, 9.0);
    wl_box b;
    b.lo[0] = cx - sx / 2;
    b.hi[0] = cx + sx / 2;
    b.lo[1] = cy - sy / 2;
    b.hi[1] = cy + sy / 2;
    b.lo[2] = -1.5;
    b.hi[2] = -1.5 + sz;
    if (b.lo[0] < -10) b.lo[0] = -10;
    if (b.hi[0] > 10) b.hi[0] = 10;
    if (b.lo[1] < -10) b.lo[1] = -10;
    if (b.hi[1] > 10) b.hi[1] = 10;
    if (b.lo[0] < 0.5 && b.hi[0] > -0.5 && b.lo[1] < 0.5 && b.hi[1] > -0.5) continue;
    s->box[s->nbox++] = b;
  }
}

/* Trajectory of n poses (x, y, z, roll, pitch, yaw): steps of `step` metres,
 * heading drifting by at most max_yaw_deg per step. */
void wl_trajectory(uint64_t seed, int32_t n, double step, double max_yaw_deg, double z,
                   double* poses) {
  rng_t r;
  rng_seed(&r, seed);
  double x = 0, y = 0, yaw = rng_unif(&r, -3.141592653589793, 3.141592653589793);
  const double dy = max_yaw_deg * 3.141592653589793 / 180.0;
  for (int i = 0; i < n; ++i) {
    double* p = poses + 6 * i;
    p[0] = x;
    p[1] = y;
    p[2] = z;
    p[3] = 0;
    p[4] = 0;
    p[5] = yaw;
    yaw += rng_unif(&r, -dy, dy);
    x += step * cos(yaw);
    y += step * sin(yaw);
  }
}

/* Outdoor scene: ground plane z = 0, nbox boxes and ncyl vertical cylinders
 * in [-half, half]^2, none within `clear` metres (xy) of any trajectory pose. */
void wl_scene_outdoor(uint64_t seed, int32_t nbox, int32_t ncyl, double half, const double* poses,
                      int32_t npose, double clear, double sky_range, wl_scene* s) {
  memset(s, 0, sizeof *s);
  rng_t r;
  rng_seed(&r, seed);
  s->has_ground = 1;
  s->ground_z = 0;
  s->sky_range = sky_range;
  if (nbox > WL_MAX_BOX) nbox = WL_MAX_BOX;
  if (ncyl > WL_MAX_CYL) ncyl = WL_MAX_CYL;
  while (s->nbox < nbox || s->ncyl < ncyl) {
    const int is_box = s->nbox < nbox && (s->ncyl >= ncyl || (rng_next(&r) & 1));
    const double cx = rng_unif(&r, -half, half), cy = rng_unif(&r, -half, half);
    double rad;
    wl_box b;
    wl_cyl c;
    if (is_box) {
      const double sx = rng_unif(&r, 0.5, 6.0), sy = rng_unif(&r, 0.5, 6.0), h = rng_unif(&r, 0.5, 12.0);
      b.lo[0] = cx - sx / 2;
      b.hi[0] = cx + sx / 2;
      b.lo[1] = cy - sy / 2;
      b.hi[1] = cy + sy / 2;
      b.lo[2] = 0;
      b.hi[2] = h;
      rad = 0.5 * sqrt(sx * sx + sy * sy);
    } else {
      c.cx = cx;
      c.cy = cy;
      c.r = rng_unif(&r, 0.15, 1.5);
      c.z0 = 0;
      c.z1 = rng_unif(&r, 1.0, 15.0);
      rad = c.r;
    }
    int ok = 1;
    for (int i = 0; i < npose && ok; ++i) {
      const double ex = poses[6 * i] - cx, ey = poses[6 * i + 1] - cy;
      if (sqrt(ex * ex + ey * ey) < rad + clear) ok = 0;
    }
    if (!ok) continue;
    if (is_box)
      s->box[s->nbox++] = b;
    else
      s->cyl[s->ncyl++] = c;
  }
}

static void rotation(const double pose[6], double R[9]) {
  const double cr = cos(pose[3]), sr = sin(pose[3]), cp = cos(pose[4]), sp = sin(pose[4]),
               cy = cos(pose[5]), sy = sin(pose[5]);
  /* R = Rz(yaw) * Ry(pitch) * Rx(roll) */
  R[0] = cy * cp;
  R[1] = cy * sp * sr - sy * cr;
  R[2] = cy * sp * cr + sy * sr;
  R[3] = sy * cp;
  R[4] = sy * sp * sr + cy * cr;
  R[5] = sy * sp * cr - cy * sr;
  R[6] = -sp;
  R[7] = cp * sr;
  R[8] = cp * cr;
}

/* Spinning-lidar scan: npts points, point i on ring (i % rings) and column
 * (i / rings); ring elevations evenly spaced over [elev_min, elev_max] deg,
 * column azimuths evenly spaced over 360 deg with `cols` columns. Points in
 * the world (map) frame. Threads: up to `threads` OpenMP threads. */
void wl_lidar_scan(const wl_scene* s, const double pose[6], int32_t rings, int32_t cols,
                   int64_t npts, double elev_min_deg, double elev_max_deg, double sigma,
                   uint64_t seed, int32_t threads, float* out4) {
  double R[9];
  rotation(pose, R);
  const double deg = 3.141592653589793 / 180.0;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
  for (int64_t blk = 0; blk < (npts + 1023) / 1024; ++blk) {
    rng_t r;
    rng_seed(&r, seed * 0x100000001B3ull + (uint64_t)blk);
    const int64_t end = (blk + 1) * 1024 < npts ? (blk + 1) * 1024 : npts;
    for (int64_t i = blk * 1024; i < end; ++i) {
      const int32_t ring = (int32_t)(i % rings);
      const int64_t col = i / rings;
      const double el = (rings > 1 ? elev_min_deg + (elev_max_deg - elev_min_deg) * ring / (rings - 1)
                                   : 0.5 * (elev_min_deg + elev_max_deg)) * deg;
      const double az = 2.0 * 3.141592653589793 * (double)col / (double)cols;
      const double ds[3] = {cos(el) * cos(az), cos(el) * sin(az), sin(el)};
      double d[3];
      for (int a = 0; a < 3; ++a) d[a] = R[3 * a] * ds[0] + R[3 * a + 1] * ds[1] + R[3 * a + 2] * ds[2];
      double rg = cast(s, pose, d) + sigma * rng_normal(&r);
      if (rg < 0) rg = 0;
      float* p = out4 + 4 * i;
      p[0] = (float)(pose[0] + d[0] * rg);
      p[1] = (float)(pose[1] + d[1] * rg);
      p[2] = (float)(pose[2] + d[2] * rg);
      p[3] = 1.0f;
    }
  }
}

/* Pinhole depth frame (camera looks along its +x axis, image u to -y, v to
 * -z): depth along the optical axis, noise sigma = k * depth^2, pixels whose
 * noisy depth falls outside [dmin, dmax] are NaN (no return). */
void wl_depth_frame(const wl_scene* s, const double pose[6], int32_t width, int32_t height,
                    double fx, double fy, double cx, double cy, double dmin, double dmax,
                    double k_noise, uint64_t seed, int32_t threads, float* out4) {
  double R[9];
  rotation(pose, R);
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
  for (int32_t v = 0; v < height; ++v) {
    rng_t r;
    rng_seed(&r, seed * 0x100000001B3ull + (uint64_t)v);
    for (int32_t u = 0; u < width; ++u) {
      const double xs = 1.0, ys = -((double)u - cx) / fx, zs = -((double)v - cy) / fy;
      const double n = sqrt(xs * xs + ys * ys + zs * zs);
      const double ds[3] = {xs / n, ys / n, zs / n};
      double d[3];
      for (int a = 0; a < 3; ++a) d[a] = R[3 * a] * ds[0] + R[3 * a + 1] * ds[1] + R[3 * a + 2] * ds[2];
      const double rg = cast(s, pose, d);
      double depth = rg / n; /* range -> depth along the optical axis */
      depth += k_noise * depth * depth * rng_normal(&r);
      float* p = out4 + 4 * ((int64_t)v * width + u);
      if (!(depth >= dmin && depth <= dmax)) {
        p[0] = p[1] = p[2] = NAN;
      } else {
        const double rr = depth * n;
        p[0] = (float)(pose[0] + d[0] * rr);
        p[1] = (float)(pose[1] + d[1] * rr);
        p[2] = (float)(pose[2] + d[2] * rr);
      }
      p[3] = 1.0f;
    }
  }
}

/* Indoor scene for depth frames: 8 x 8 x 3 m room (floor z = 0), shelves
 * (tall thin boxes) along the walls and free-standing boxes, none within
 * `clear` metres of the camera path. */
void wl_scene_indoor(uint64_t seed, int32_t nshelf, int32_t nbox, const double* poses,
                     int32_t npose, double clear, wl_scene* s) {
  memset(s, 0, sizeof *s);
  rng_t r;
  rng_seed(&r, seed);
  s->has_room = 1;
  s->room.lo[0] = -4;
  s->room.lo[1] = -4;
  s->room.lo[2] = 0;
  s->room.hi[0] = 4;
  s->room.hi[1] = 4;
  s->room.hi[2] = 3;
  s->sky_range = 1e3;
  int made = 0, tries = 0;
  while (made < nshelf + nbox && tries < 100000) {
    ++tries;
    wl_box b;
    if (made < nshelf) {
      /* shelf: 0.3-0.5 deep, 0.8-2 wide, 1.5-2.2 tall, against a wall */
      const int wall = (int)(rng_next(&r) & 3);
      const double depth = rng_unif(&r, 0.3, 0.5), width = rng_unif(&r, 0.8, 2.0),
                   h = rng_unif(&r, 1.5, 2.2), along = rng_unif(&r, -3.0, 3.0);
      if (wall < 2) {
        const double x0 = wall == 0 ? -4.0 : 4.0 - depth;
        b.lo[0] = x0;
        b.hi[0] = x0 + depth;
        b.lo[1] = along - width / 2;
        b.hi[1] = along + width / 2;
      } else {
        const double y0 = wall == 2 ? -4.0 : 4.0 - depth;
        b.lo[1] = y0;
        b.hi[1] = y0 + depth;
        b.lo[0] = along - width / 2;
        b.hi[0] = along + width / 2;
      }
      b.lo[2] = 0;
      b.hi[2] = h;
    } else {
      const double sx = rng_unif(&r, 0.2, 0.8), sy = rng_unif(&r, 0.2, 0.8), sz = rng_unif(&r, 0.2, 1.2);
      const double cx = rng_unif(&r, -3.5, 3.5), cy = rng_unif(&r, -3.5, 3.5);
      b.lo[0] = cx - sx / 2;
      b.hi[0] = cx + sx / 2;
      b.lo[1] = cy - sy / 2;
      b.hi[1] = cy + sy / 2;
      b.lo[2] = 0;
      b.hi[2] = sz;
    }
    int ok = 1;
    for (int i = 0; i < npose && ok; ++i) {
      const double px = poses[6 * i], py = poses[6 * i + 1];
      const double ex = px < b.lo[0] ? b.lo[0] - px : (px > b.hi[0] ? px - b.hi[0] : 0);
      const double ey = py < b.lo[1] ? b.lo[1] - py : (py > b.hi[1] ? py - b.hi[1] : 0);
      if (sqrt(ex * ex + ey * ey) < clear) ok = 0;
    }
    if (!ok) continue;
    s->box[s->nbox++] = b;
    ++made;
  }
}
