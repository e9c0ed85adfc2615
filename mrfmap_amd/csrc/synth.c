/*
 * Deterministic synthetic scenes and depth keyframes (host-only C).
 * See include/mrfmap_synth.h for conventions.  Random numbers come from a
 * counter-based SplitMix64 hash so every pixel's noise depends only on
 * (noise_seed, pixel index) and the output is identical on every machine.
 */
#include "mrfmap_synth.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAX_OBJ 4096

struct mrf_synth_scene {
  mrf_synth_cfg cfg;
  int nb, ns;
  float blo[MAX_OBJ][3], bhi[MAX_OBJ][3];
  float sc[MAX_OBJ][3], sr[MAX_OBJ];
};

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* uniform double in [0,1) from a 64-bit hash */
static double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

typedef struct {
  uint64_t state;
} rng_t;
static double rng_u(rng_t* r) {
  r->state = splitmix64(r->state);
  return u01(r->state);
}
static double rng_range(rng_t* r, double a, double b) { return a + (b - a) * rng_u(r); }

void mrf_synth_default_room(mrf_synth_cfg* c, uint64_t seed) {
  memset(c, 0, sizeof(*c));
  c->room[0] = 4.f;
  c->room[1] = 4.f;
  c->room[2] = 3.f;
  c->num_boxes = 12;
  c->box_min = 0.2f;
  c->box_max = 1.0f;
  c->num_spheres = 4;
  c->sph_min = 0.1f;
  c->sph_max = 0.4f;
  c->ring_radius = 2.f;
  c->ring_height = 1.5f;
  c->yaw_jitter_deg = 5.f;
  c->max_range = 0.f;
  c->seed = seed;
}

void mrf_synth_large_room(mrf_synth_cfg* c, uint64_t seed) {
  mrf_synth_default_room(c, seed);
  c->room[0] = 10.f;
  c->room[1] = 10.f;
  c->room[2] = 3.f;
  c->num_boxes = 60;
  c->num_spheres = 4;
  c->ring_radius = 4.f;
  c->max_range = 10.f;
}

/* An object may not come within `margin` of the camera ring at ring height. */
static int blocks_ring(const mrf_synth_cfg* c, const float lo[3], const float hi[3]) {
  const float margin = 0.3f;
  if (hi[2] < c->ring_height - margin || lo[2] > c->ring_height + margin) return 0;
  /* nearest and farthest horizontal distance of the box from the room centre */
  float nx = fmaxf(0.f, fmaxf(lo[0], -hi[0]));
  float ny = fmaxf(0.f, fmaxf(lo[1], -hi[1]));
  float fx = fmaxf(fabsf(lo[0]), fabsf(hi[0]));
  float fy = fmaxf(fabsf(lo[1]), fabsf(hi[1]));
  float dmin = sqrtf(nx * nx + ny * ny), dmax = sqrtf(fx * fx + fy * fy);
  return dmax > c->ring_radius - margin && dmin < c->ring_radius + margin;
}

int mrf_synth_create(const mrf_synth_cfg* cfg, mrf_synth_scene** out) {
  if (!cfg || !out) return -1;
  if (cfg->num_boxes < 0 || cfg->num_boxes > MAX_OBJ || cfg->num_spheres < 0 ||
      cfg->num_spheres > MAX_OBJ)
    return -2;
  mrf_synth_scene* s = (mrf_synth_scene*)calloc(1, sizeof(mrf_synth_scene));
  if (!s) return -3;
  s->cfg = *cfg;
  rng_t r = {splitmix64(cfg->seed ^ 0x5CE7E5CE7Eull)};
  const float hx = cfg->room[0] * 0.5f, hy = cfg->room[1] * 0.5f, hz = cfg->room[2];
  for (int i = 0; i < cfg->num_boxes; ++i) {
    for (int tries = 0; tries < 1000; ++tries) {
      float e[3];
      for (int a = 0; a < 3; ++a) e[a] = (float)rng_range(&r, cfg->box_min, cfg->box_max);
      float cx = (float)rng_range(&r, -hx + e[0] * 0.5, hx - e[0] * 0.5);
      float cy = (float)rng_range(&r, -hy + e[1] * 0.5, hy - e[1] * 0.5);
      int on_floor = rng_u(&r) < 0.5;
      float z0 = on_floor ? 0.f : (float)rng_range(&r, 0.0, fmax(0.0, hz - e[2]));
      float lo[3] = {cx - e[0] * 0.5f, cy - e[1] * 0.5f, z0};
      float hi[3] = {cx + e[0] * 0.5f, cy + e[1] * 0.5f, fminf(z0 + e[2], hz)};
      if (blocks_ring(cfg, lo, hi)) continue;
      memcpy(s->blo[s->nb], lo, sizeof(lo));
      memcpy(s->bhi[s->nb], hi, sizeof(hi));
      s->nb++;
      break;
    }
  }
  for (int i = 0; i < cfg->num_spheres; ++i) {
    for (int tries = 0; tries < 1000; ++tries) {
      float rad = (float)rng_range(&r, cfg->sph_min, cfg->sph_max);
      float c[3] = {(float)rng_range(&r, -hx + rad, hx - rad), (float)rng_range(&r, -hy + rad, hy - rad),
                    (float)rng_range(&r, rad, hz - rad)};
      float lo[3] = {c[0] - rad, c[1] - rad, c[2] - rad}, hi[3] = {c[0] + rad, c[1] + rad, c[2] + rad};
      if (blocks_ring(cfg, lo, hi)) continue;
      memcpy(s->sc[s->ns], c, sizeof(c));
      s->sr[s->ns] = rad;
      s->ns++;
      break;
    }
  }
  *out = s;
  return 0;
}

void mrf_synth_destroy(mrf_synth_scene* s) { free(s); }

int mrf_synth_num_objects(const mrf_synth_scene* s, int* boxes, int* spheres) {
  if (!s) return -1;
  if (boxes) *boxes = s->nb;
  if (spheres) *spheres = s->ns;
  return 0;
}
int mrf_synth_box(const mrf_synth_scene* s, int i, float lo[3], float hi[3]) {
  if (!s || i < 0 || i >= s->nb) return -1;
  memcpy(lo, s->blo[i], 12);
  memcpy(hi, s->bhi[i], 12);
  return 0;
}
int mrf_synth_sphere(const mrf_synth_scene* s, int i, float c[3], float* r) {
  if (!s || i < 0 || i >= s->ns) return -1;
  memcpy(c, s->sc[i], 12);
  *r = s->sr[i];
  return 0;
}

int mrf_synth_ring_pose(const mrf_synth_scene* s, int k, int n, float pose[16]) {
  if (!s || n <= 0 || k < 0 || k >= n) return -1;
  const mrf_synth_cfg* c = &s->cfg;
  double th = 2.0 * M_PI * (double)k / (double)n;
  double px = c->ring_radius * cos(th), py = c->ring_radius * sin(th), pz = c->ring_height;
  double jit = (2.0 * u01(splitmix64(c->seed * 0x100000001B3ull + 7919ull * (uint64_t)(k + 1))) - 1.0) *
               c->yaw_jitter_deg * M_PI / 180.0;
  /* forward: towards (0,0,height), rotated by the jitter about world z */
  double fx0 = -px, fy0 = -py, nrm = sqrt(fx0 * fx0 + fy0 * fy0);
  fx0 /= nrm;
  fy0 /= nrm;
  double f[3] = {cos(jit) * fx0 - sin(jit) * fy0, sin(jit) * fx0 + cos(jit) * fy0, 0.0};
  double up[3] = {0.0, 0.0, 1.0};
  /* right = f x up ; down = f x right */
  double rx[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double rn = sqrt(rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2]);
  for (int a = 0; a < 3; ++a) rx[a] /= rn;
  double dn[3] = {f[1] * rx[2] - f[2] * rx[1], f[2] * rx[0] - f[0] * rx[2], f[0] * rx[1] - f[1] * rx[0]};
  double pos[3] = {px, py, pz};
  for (int i = 0; i < 3; ++i) {
    pose[i * 4 + 0] = (float)rx[i];
    pose[i * 4 + 1] = (float)dn[i];
    pose[i * 4 + 2] = (float)f[i];
    pose[i * 4 + 3] = (float)pos[i];
  }
  pose[12] = pose[13] = pose[14] = 0.f;
  pose[15] = 1.f;
  return 0;
}

/* First hit (t > eps) of o + t d with the scene; returns +inf if none.  t is the
 * parameter along d, which equals planar depth because d has unit camera z. */
static double cast_ray(const mrf_synth_scene* s, const double o[3], const double d[3]) {
  const double eps = 1e-9;
  const mrf_synth_cfg* c = &s->cfg;
  double lo_room[3] = {-0.5 * c->room[0], -0.5 * c->room[1], 0.0};
  double hi_room[3] = {0.5 * c->room[0], 0.5 * c->room[1], c->room[2]};
  double best = INFINITY;
  /* room walls, seen from inside: exit distance of the room box */
  for (int a = 0; a < 3; ++a) {
    if (d[a] > 0) {
      double t = (hi_room[a] - o[a]) / d[a];
      if (t > eps && t < best) best = t;
    } else if (d[a] < 0) {
      double t = (lo_room[a] - o[a]) / d[a];
      if (t > eps && t < best) best = t;
    }
  }
  for (int i = 0; i < s->nb; ++i) {
    double tn = -INFINITY, tf = INFINITY;
    int miss = 0;
    for (int a = 0; a < 3 && !miss; ++a) {
      if (d[a] == 0.0) {
        if (o[a] < s->blo[i][a] || o[a] > s->bhi[i][a]) miss = 1;
        continue;
      }
      double t0 = (s->blo[i][a] - o[a]) / d[a], t1 = (s->bhi[i][a] - o[a]) / d[a];
      if (t0 > t1) {
        double tt = t0;
        t0 = t1;
        t1 = tt;
      }
      if (t0 > tn) tn = t0;
      if (t1 < tf) tf = t1;
    }
    if (miss || tn > tf) continue;
    double t = tn > eps ? tn : tf;
    if (t > eps && t < best) best = t;
  }
  for (int i = 0; i < s->ns; ++i) {
    double oc[3] = {o[0] - s->sc[i][0], o[1] - s->sc[i][1], o[2] - s->sc[i][2]};
    double A = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double B = 2.0 * (oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2]);
    double C = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - (double)s->sr[i] * s->sr[i];
    double disc = B * B - 4 * A * C;
    if (disc < 0) continue;
    double sq = sqrt(disc);
    double t0 = (-B - sq) / (2 * A), t1 = (-B + sq) / (2 * A);
    double t = t0 > eps ? t0 : t1;
    if (t > eps && t < best) best = t;
  }
  return best;
}

/* standard normal from two hashed uniforms (Box-Muller) */
static double gauss(uint64_t h1, uint64_t h2) {
  double u1 = u01(h1), u2 = u01(h2);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

int mrf_synth_render(const mrf_synth_scene* s, const float pose[16], float fx, float fy, float cx,
                     float cy, int rows, int cols, int noise, float dropout, uint64_t noise_seed,
                     float* depth) {
  if (!s || !pose || !depth || rows <= 0 || cols <= 0 || fx <= 0 || fy <= 0) return -1;
  double o[3] = {pose[3], pose[7], pose[11]};
  for (int y = 0; y < rows; ++y) {
    for (int x = 0; x < cols; ++x) {
      double dc[3] = {((double)x - cx) / fx, ((double)y - cy) / fy, 1.0};
      double d[3];
      for (int i = 0; i < 3; ++i) d[i] = pose[i * 4 + 0] * dc[0] + pose[i * 4 + 1] * dc[1] + pose[i * 4 + 2] * dc[2];
      double z = cast_ray(s, o, d);
      uint64_t idx = (uint64_t)y * (uint64_t)cols + (uint64_t)x;
      if (!isfinite(z)) z = 0.0;
      if (noise && z > 0) {
        uint64_t h = splitmix64(noise_seed * 0x9E3779B97F4A7C15ull + idx);
        if (u01(splitmix64(h ^ 0xD80Dull)) < dropout) {
          z = 0.0;
        } else {
          double sig = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
          z += sig * gauss(splitmix64(h ^ 0x1111ull), splitmix64(h ^ 0x2222ull));
          if (z < 0) z = 0;
        }
      }
      if (s->cfg.max_range > 0 && z > s->cfg.max_range) z = 0.0;
      depth[idx] = (float)z;
    }
  }
  return 0;
}

int mrf_synth_occupancy(const mrf_synth_scene* s, float res, const float* p, int64_t n, uint8_t* occ) {
  (void)res;
  if (!s || !p || !occ || n < 0) return -1;
  const mrf_synth_cfg* c = &s->cfg;
  for (int64_t i = 0; i < n; ++i) {
    float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
    int o = (x < -0.5f * c->room[0] || x >= 0.5f * c->room[0] || y < -0.5f * c->room[1] ||
             y >= 0.5f * c->room[1] || z < 0 || z >= c->room[2]);
    for (int b = 0; b < s->nb && !o; ++b)
      o = x >= s->blo[b][0] && x <= s->bhi[b][0] && y >= s->blo[b][1] && y <= s->bhi[b][1] &&
          z >= s->blo[b][2] && z <= s->bhi[b][2];
    for (int q = 0; q < s->ns && !o; ++q) {
      float dx = x - s->sc[q][0], dy = y - s->sc[q][1], dz = z - s->sc[q][2];
      o = dx * dx + dy * dy + dz * dz <= s->sr[q] * s->sr[q];
    }
    occ[i] = (uint8_t)o;
  }
  return 0;
}
