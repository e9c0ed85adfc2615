/*
 * OctoMap-equivalent point-cloud insertion — CPU COMPARATOR, TEST/BENCH ONLY.
 *
 * Restates the published algorithm of octomap::OccupancyOcTreeBase::
 * insertPointCloud(scan, origin) (OctoMap 1.9, icoderaven/octomap@master is
 * the reference's un-vendored dependency, CMakeLists.txt:63-66) as
 * configured by GVDBOctomapWrapper::init (src/GVDBOctomapWrapper.cpp:3-17):
 * resolution = voxel size, clamping [0.0001, 0.9999], BBX-limited updates.
 *   1. computeUpdate: per point, the endpoint key joins the occupied set if
 *      it lies in the BBX; computeRayKeys(origin, point) walks the key grid
 *      with OctoMap's 3-D DDA (tMax/tDelta from cell borders) and every key
 *      before the endpoint joins the free set until one leaves the BBX.
 *   2. free keys not in the occupied set get a miss (log-odds of 0.4),
 *      occupied keys a hit (log-odds of 0.7), clamped; inner nodes of the
 *      16-level octree take the max of their children.
 * This is not OctoMap itself (not available offline) and is labelled
 * "OctoMap-equivalent" wherever it is timed.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define TREE_DEPTH 16
#define TREE_MAX_VAL 32768

typedef struct onode {
  float lo;
  struct onode** ch;
} onode;

struct oct_map {
  double res;
  uint16_t bmin[3], bmax[3];
  double bbx_min[3], bbx_max[3];
  float hit, miss, cmin, cmax;
  onode* root;
  int64_t nodes;
  /* per-scan key sets (open addressing, key+1 stored, 0 = empty) */
  uint64_t *fset, *oset;
  int64_t fcap, ocap, fn, on;
};

static float logodds(double p) { return (float)log(p / (1.0 - p)); }

static int coord_to_key(double res, double c, uint16_t* k) {
  int64_t v = (int64_t)floor(c / res) + TREE_MAX_VAL;
  if (v < 0 || v >= 2 * TREE_MAX_VAL) return 0;
  *k = (uint16_t)v;
  return 1;
}
static double key_to_coord(double res, uint16_t k) { return ((double)((int)k - TREE_MAX_VAL) + 0.5) * res; }

static uint64_t pack(const uint16_t k[3]) { return ((uint64_t)k[0] << 32) | ((uint64_t)k[1] << 16) | k[2]; }
static uint64_t mix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  return x;
}

static int set_insert(uint64_t** tab, int64_t* cap, int64_t* n, uint64_t key) {
  if ((*n + 1) * 2 > *cap) {
    int64_t nc = *cap ? *cap * 2 : 1 << 16;
    uint64_t* nt = (uint64_t*)calloc((size_t)nc, sizeof(uint64_t));
    for (int64_t i = 0; i < *cap; ++i)
      if ((*tab)[i]) {
        int64_t h = (int64_t)(mix((*tab)[i] - 1) & (uint64_t)(nc - 1));
        while (nt[h]) h = (h + 1) & (nc - 1);
        nt[h] = (*tab)[i];
      }
    free(*tab);
    *tab = nt;
    *cap = nc;
  }
  int64_t h = (int64_t)(mix(key) & (uint64_t)(*cap - 1));
  while ((*tab)[h]) {
    if ((*tab)[h] == key + 1) return 0;
    h = (h + 1) & (*cap - 1);
  }
  (*tab)[h] = key + 1;
  ++*n;
  return 1;
}

static int set_has(const uint64_t* tab, int64_t cap, uint64_t key) {
  if (!cap) return 0;
  int64_t h = (int64_t)(mix(key) & (uint64_t)(cap - 1));
  while (tab[h]) {
    if (tab[h] == key + 1) return 1;
    h = (h + 1) & (cap - 1);
  }
  return 0;
}

int oct_create(double res, const double bbx_min[3], const double bbx_max[3], double occ_thresh, oct_map** out) {
  (void)occ_thresh;
  oct_map* m = (oct_map*)calloc(1, sizeof(oct_map));
  m->res = res;
  for (int a = 0; a < 3; ++a) {
    m->bbx_min[a] = bbx_min[a];
    m->bbx_max[a] = bbx_max[a];
    if (!coord_to_key(res, bbx_min[a], &m->bmin[a]) || !coord_to_key(res, bbx_max[a], &m->bmax[a])) {
      free(m);
      return -1;
    }
  }
  m->hit = logodds(0.7);
  m->miss = logodds(0.4);
  m->cmin = logodds(0.0001);
  m->cmax = logodds(0.9999);
  m->root = (onode*)calloc(1, sizeof(onode));
  m->nodes = 1;
  *out = m;
  return 0;
}

static void free_node(onode* n) {
  if (!n) return;
  if (n->ch) {
    for (int i = 0; i < 8; ++i) free_node(n->ch[i]);
    free(n->ch);
  }
  free(n);
}

void oct_destroy(oct_map* m) {
  if (!m) return;
  free_node(m->root);
  free(m->fset);
  free(m->oset);
  free(m);
}

static int in_bbx_key(const oct_map* m, const uint16_t k[3]) {
  for (int a = 0; a < 3; ++a)
    if (k[a] < m->bmin[a] || k[a] > m->bmax[a]) return 0;
  return 1;
}
static int in_bbx_point(const oct_map* m, const double p[3]) {
  for (int a = 0; a < 3; ++a)
    if (p[a] < m->bbx_min[a] || p[a] > m->bbx_max[a]) return 0;
  return 1;
}

/* descend, update the leaf, refresh inner nodes with the max of children */
static void update_recurs(oct_map* m, onode* n, const uint16_t k[3], int depth, float delta) {
  if (depth == TREE_DEPTH) {
    float v = n->lo + delta;
    n->lo = v < m->cmin ? m->cmin : (v > m->cmax ? m->cmax : v);
    return;
  }
  const int bit = TREE_DEPTH - 1 - depth;
  const int idx = ((k[0] >> bit) & 1) | (((k[1] >> bit) & 1) << 1) | (((k[2] >> bit) & 1) << 2);
  if (!n->ch) n->ch = (onode**)calloc(8, sizeof(onode*));
  if (!n->ch[idx]) {
    n->ch[idx] = (onode*)calloc(1, sizeof(onode));
    m->nodes++;
  }
  update_recurs(m, n->ch[idx], k, depth + 1, delta);
  float mx = -INFINITY;
  for (int i = 0; i < 8; ++i)
    if (n->ch[i] && n->ch[i]->lo > mx) mx = n->ch[i]->lo;
  n->lo = mx;
}

/* OcTreeBaseImpl::computeRayKeys: keys from origin up to (excluding) end */
static int64_t ray_keys(oct_map* m, const double o[3], const double e[3], int64_t* nfree) {
  uint16_t ko[3], ke[3], cur[3];
  int64_t upd = 0;
  if (!coord_to_key(m->res, o[0], &ko[0]) || !coord_to_key(m->res, o[1], &ko[1]) || !coord_to_key(m->res, o[2], &ko[2]) ||
      !coord_to_key(m->res, e[0], &ke[0]) || !coord_to_key(m->res, e[1], &ke[1]) || !coord_to_key(m->res, e[2], &ke[2]))
    return 0;
  if (ko[0] == ke[0] && ko[1] == ke[1] && ko[2] == ke[2]) return 0;
  double dir[3] = {e[0] - o[0], e[1] - o[1], e[2] - o[2]};
  double length = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  for (int a = 0; a < 3; ++a) dir[a] /= length;
  int step[3];
  double tmax[3], tdel[3];
  memcpy(cur, ko, sizeof(cur));
  for (int a = 0; a < 3; ++a) {
    step[a] = dir[a] > 0 ? 1 : (dir[a] < 0 ? -1 : 0);
    if (step[a]) {
      double border = key_to_coord(m->res, cur[a]) + (double)step[a] * m->res * 0.5;
      tmax[a] = (border - o[a]) / dir[a];
      tdel[a] = m->res / fabs(dir[a]);
    } else {
      tmax[a] = tdel[a] = 1e300;
    }
  }
  /* the origin key is part of the ray; stop adding once a key leaves the BBX */
  int inside = in_bbx_key(m, cur);
  if (inside) {
    set_insert(&m->fset, &m->fcap, &m->fn, pack(cur));
    ++upd;
  }
  for (;;) {
    int dim;
    if (tmax[0] < tmax[1])
      dim = tmax[0] < tmax[2] ? 0 : 2;
    else
      dim = tmax[1] < tmax[2] ? 1 : 2;
    cur[dim] = (uint16_t)(cur[dim] + step[dim]);
    tmax[dim] += tdel[dim];
    if (cur[0] == ke[0] && cur[1] == ke[1] && cur[2] == ke[2]) break;
    double dmin = fmin(tmax[0], fmin(tmax[1], tmax[2]));
    if (dmin > length) break;
    if (!inside) break;
    if (!in_bbx_key(m, cur)) {
      inside = 0;
      break;
    }
    set_insert(&m->fset, &m->fcap, &m->fn, pack(cur));
    ++upd;
  }
  *nfree += upd;
  return upd;
}

int64_t oct_insert_point_cloud(oct_map* m, const float* pts, int64_t n, const double origin[3]) {
  if (!m || (!pts && n)) return -1;
  if (m->fset) memset(m->fset, 0, sizeof(uint64_t) * (size_t)m->fcap);
  if (m->oset) memset(m->oset, 0, sizeof(uint64_t) * (size_t)m->ocap);
  m->fn = m->on = 0;
  int64_t traversed = 0;
  for (int64_t i = 0; i < n; ++i) {
    double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    uint16_t k[3];
    if (in_bbx_point(m, p) && coord_to_key(m->res, p[0], &k[0]) && coord_to_key(m->res, p[1], &k[1]) &&
        coord_to_key(m->res, p[2], &k[2]))
      set_insert(&m->oset, &m->ocap, &m->on, pack(k));
    ray_keys(m, origin, p, &traversed);
  }
  int64_t updates = 0;
  for (int64_t i = 0; i < m->fcap; ++i) {
    uint64_t key = m->fset[i];
    if (!key) continue;
    key -= 1;
    if (set_has(m->oset, m->ocap, key)) continue;
    uint16_t k[3] = {(uint16_t)(key >> 32), (uint16_t)(key >> 16), (uint16_t)key};
    update_recurs(m, m->root, k, 0, m->miss);
    ++updates;
  }
  for (int64_t i = 0; i < m->ocap; ++i) {
    uint64_t key = m->oset[i];
    if (!key) continue;
    key -= 1;
    uint16_t k[3] = {(uint16_t)(key >> 32), (uint16_t)(key >> 16), (uint16_t)key};
    update_recurs(m, m->root, k, 0, m->hit);
    ++updates;
  }
  return traversed + m->on;
}

int64_t oct_num_nodes(oct_map* m) { return m ? m->nodes : -1; }

double oct_logodds_at(oct_map* m, const double p[3]) {
  uint16_t k[3];
  if (!m || !coord_to_key(m->res, p[0], &k[0]) || !coord_to_key(m->res, p[1], &k[1]) ||
      !coord_to_key(m->res, p[2], &k[2]))
    return NAN;
  onode* n = m->root;
  for (int depth = 0; depth < TREE_DEPTH; ++depth) {
    const int bit = TREE_DEPTH - 1 - depth;
    const int idx = ((k[0] >> bit) & 1) | (((k[1] >> bit) & 1) << 1) | (((k[2] >> bit) & 1) << 2);
    if (!n->ch || !n->ch[idx]) return NAN;
    n = n->ch[idx];
  }
  return n->lo;
}
