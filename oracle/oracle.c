/*
 * CPU oracle of the MRFMap BP inference path — TEST INFRASTRUCTURE ONLY
 * (see oracle.h).  Scalar, FP64, single-threaded.
 *
 * The traversal is deliberately written differently from the GPU's
 * incremental DDA: all boundary crossings of the ray inside the grid are
 * collected per axis as exact rationals, merged, and each interval between
 * consecutive distinct crossings is assigned the voxel that contains its
 * midpoint (exact 128-bit arithmetic).  The float outputs (chord lengths,
 * entry depth) are then derived by the same IEEE formulas the traversal
 * contract fixes, so GPU and oracle lists must agree bit for bit.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FIX 16
#define FIXS 65536.0
#define F_EPS 1e-4
#define GAP 0xFFFFFFFFu
#define LOOKUP_PX 20
#define RAN_OUT 0x80000000u
/* the epilogue's gate on a camera's summed chord length, C > 0.1f in float
 * (gvdbkernels.cu:408), applied to the FP64 sum */
#define CUM_LEN_GATE ((double)0.1f)

typedef __int128 i128;
typedef struct {
  int64_t n, d; /* n / d, d > 0 */
} rat;

static int rlt(rat a, rat b) { return (i128)a.n * b.d < (i128)b.n * a.d; }
static int req(rat a, rat b) { return (i128)a.n * b.d == (i128)b.n * a.d; }
static double rdbl(rat a) { return (double)a.n / (double)a.d; }

/* floor(a / b) for b > 0 */
static i128 floordiv128(i128 a, i128 b) {
  i128 q = a / b;
  if ((a % b != 0) && (a < 0)) q -= 1;
  return q;
}

static double poly(const double c[3], int n, double s) {
  /* same operation order as the traversal contract */
  if (n == 3) {
    double p = (s * s) * c[2];
    p = p + s * c[1];
    return p + c[0];
  }
  double p = s * c[1];
  return p + c[0];
}

double orc_depth_psi(double s, double d, const double sig[3], const double bia[3], int lookup_n) {
  /* depth_potential, gvdbkernels.cu:57-72: N(d; s + bias(s), (3 sigma(s))^2) */
  double diff = fabs(d - (s + poly(bia, lookup_n, s)));
  double sg3 = 3.0 * poly(sig, lookup_n, s);
  double u = diff / sg3;
  return 0.398942 * exp(-0.5 * u * u) / sg3;
}

/* ------------------------------------------------------------------ walk */

typedef struct {
  int32_t v[3];
  rat z0, z1;
} seg_t;

typedef struct {
  seg_t* s;
  int n, cap;
} segv;

static void segv_push(segv* sv, const int32_t v[3], rat a, rat b) {
  if (sv->n == sv->cap) {
    sv->cap = sv->cap ? 2 * sv->cap : 256;
    sv->s = (seg_t*)realloc(sv->s, sizeof(seg_t) * sv->cap);
  }
  seg_t* e = &sv->s[sv->n++];
  memcpy(e->v, v, sizeof(e->v));
  e->z0 = a;
  e->z1 = b;
}

/* All voxels pierced by the ray inside the grid, in order (merge + midpoint). */
static int ray_segments(const int64_t Oi[3], const int64_t Di[3], const int N[3], segv* out) {
  out->n = 0;
  rat zin = {0, 1}, zout = {0, 1};
  int has_out = 0;
  for (int a = 0; a < 3; ++a) {
    int64_t top = (int64_t)N[a] << FIX;
    if (Di[a] == 0) {
      if (Oi[a] < 0 || Oi[a] >= top) return 0;
      continue;
    }
    rat lo, hi;
    if (Di[a] > 0) {
      lo.n = -Oi[a]; lo.d = Di[a];
      hi.n = top - Oi[a]; hi.d = Di[a];
    } else {
      lo.n = Oi[a] - top; lo.d = -Di[a];
      hi.n = Oi[a]; hi.d = -Di[a];
    }
    if (rlt(zin, lo)) zin = lo;
    if (!has_out || rlt(hi, zout)) {
      zout = hi;
      has_out = 1;
    }
  }
  if (!has_out || !rlt(zin, zout)) return 0;
  /* crossings strictly inside (zin, zout), per axis in increasing z */
  rat* cr[3] = {NULL, NULL, NULL};
  int nc[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    if (Di[a] == 0) continue;
    cr[a] = (rat*)malloc(sizeof(rat) * (size_t)(N[a] + 1));
    for (int kk = 1; kk < N[a]; ++kk) {
      int k = Di[a] > 0 ? kk : N[a] - kk;
      rat z;
      if (Di[a] > 0) {
        z.n = ((int64_t)k << FIX) - Oi[a];
        z.d = Di[a];
      } else {
        z.n = Oi[a] - ((int64_t)k << FIX);
        z.d = -Di[a];
      }
      if (rlt(zin, z) && rlt(z, zout)) cr[a][nc[a]++] = z;
    }
  }
  /* merge distinct breakpoints */
  int i0 = 0, i1 = 0, i2 = 0;
  rat prev = zin;
  for (;;) {
    rat best;
    int found = 0;
    if (i0 < nc[0]) { best = cr[0][i0]; found = 1; }
    if (i1 < nc[1] && (!found || rlt(cr[1][i1], best))) { best = cr[1][i1]; found = 1; }
    if (i2 < nc[2] && (!found || rlt(cr[2][i2], best))) { best = cr[2][i2]; found = 1; }
    rat next = found ? best : zout;
    /* voxel containing the midpoint of [prev, next] */
    int32_t v[3];
    i128 den = (i128)2 * prev.d * next.d;
    i128 num = (i128)prev.n * next.d + (i128)next.n * prev.d; /* mid = num / den */
    for (int a = 0; a < 3; ++a) {
      i128 coord = (i128)Oi[a] * den + num * (i128)Di[a]; /* in 2^-FIX units times den */
      v[a] = (int32_t)floordiv128(coord, den << FIX);
    }
    segv_push(out, v, prev, next);
    if (!found) break;
    if (i0 < nc[0] && req(cr[0][i0], next)) ++i0;
    if (i1 < nc[1] && req(cr[1][i1], next)) ++i1;
    if (i2 < nc[2] && req(cr[2][i2], next)) ++i2;
    prev = next;
  }
  for (int a = 0; a < 3; ++a) free(cr[a]);
  return out->n;
}

/* Truncated, activation-filtered walk.  slot == NULL: every brick active and
 * visited bricks are flagged in mark[].  Emits records via the callbacks. */
typedef struct {
  uint32_t* rec;   /* 2 words per record (may be NULL for counting) */
  double* len;     /* exact chord length per record */
  double* zex;     /* exact exit depth per record */
  int32_t* vxyz;   /* voxel coordinates (3 per record, gap = -1) */
  int64_t n, cap;
} recbuf;

static void rb_push(recbuf* b, uint32_t vox, float lf, double len, double zex, const int32_t* v) {
  if (b->n == b->cap) {
    b->cap = b->cap ? 2 * b->cap : 1024;
    b->rec = (uint32_t*)realloc(b->rec, sizeof(uint32_t) * 2 * (size_t)b->cap);
    b->len = (double*)realloc(b->len, sizeof(double) * (size_t)b->cap);
    b->zex = (double*)realloc(b->zex, sizeof(double) * (size_t)b->cap);
    b->vxyz = (int32_t*)realloc(b->vxyz, sizeof(int32_t) * 3 * (size_t)b->cap);
  }
  uint32_t lb;
  memcpy(&lb, &lf, 4);
  b->rec[2 * b->n] = vox;
  b->rec[2 * b->n + 1] = lb;
  b->len[b->n] = len;
  b->zex[b->n] = zex;
  if (v) {
    memcpy(&b->vxyz[3 * b->n], v, 12);
  } else {
    b->vxyz[3 * b->n] = b->vxyz[3 * b->n + 1] = b->vxyz[3 * b->n + 2] = -1;
  }
  b->n++;
}

typedef struct {
  int bx, by, bz;
  const int32_t* slot; /* NULL: all active, mark visited */
  uint8_t* mark;
} brickmap;

static int brick_lin(const brickmap* bm, const int32_t v[3]) {
  return ((v[0] >> 3) * bm->by + (v[1] >> 3)) * bm->bz + (v[2] >> 3);
}

/* returns number of records emitted; *z0f, *ran_out set */
static int64_t walk_filtered(const segv* sv, const brickmap* bm, double norm, double depth, const double sig[3],
                             const double bia[3], int lookup_n, recbuf* out, float* z0f, int* ran_out) {
  int64_t start = out ? out->n : 0, cnt = 0;
  int any = 0, skipped = 0;
  double last_exit = 0.0;
  *ran_out = 1;
  *z0f = 0.f;
  for (int j = 0; j < sv->n; ++j) {
    const seg_t* e = &sv->s[j];
    int lin = brick_lin(bm, e->v);
    int active;
    if (!bm->slot) {
      bm->mark[lin] = 1;
      active = 1;
    } else {
      active = bm->slot[lin] >= 0;
    }
    double zc = rdbl(e->z0), zn = rdbl(e->z1);
    if (active && bm->slot) {
      if (any && skipped) {
        double gl = (zc - last_exit) * norm;
        if (out) rb_push(out, GAP, (float)gl, gl, zc, NULL);
        ++cnt;
        skipped = 0;
      }
      if (!any) {
        *z0f = (float)zc;
        any = 1;
      }
      double l = (zn - zc) * norm;
      uint32_t local = (uint32_t)(((e->v[0] & 7) * 8 + (e->v[1] & 7)) * 8 + (e->v[2] & 7));
      uint32_t vox = (uint32_t)bm->slot[lin] * 512u + local;
      if (out) rb_push(out, vox, (float)l, l, zn, e->v);
      ++cnt;
      last_exit = zn;
    } else if (!active && any) {
      skipped = 1;
    }
    int brick_exit = (j + 1 == sv->n) || brick_lin(bm, sv->s[j + 1].v) != lin;
    if (brick_exit && active) {
      double lhs = zn + poly(bia, lookup_n, zn);
      double rhs = depth + 3.0 * poly(sig, lookup_n, zn);
      if (lhs >= rhs) {
        *ran_out = 0;
        break;
      }
    }
  }
  (void)start;
  return cnt;
}

/* ------------------------------------------------------------ context */

typedef struct {
  float pose[16], K[9];
  double R[9], t[3], fx, fy, cx, cy;
  int rows, cols, lookup_u;
  float* depth;
  /* traversal */
  int64_t nrays;
  uint32_t* rays; /* 6 words per ray */
  recbuf rb;
  /* pairs */
  int64_t npairs;
  uint32_t* pair_vox;
  double* pair_c;
  uint32_t* rpair; /* per record: its pair (GAP for gap records) */
  double* M;
} ocam;

struct orc_ctx {
  orc_params p;
  int N[3], Mx[3], B3[3], nbricks;
  double res, off[3], prior_lo;
  double sig[3], bia[3];
  float *sig_tab, *bia_tab; /* per-cell tables (NULL: sig/bia in every cell) */
  int32_t* slot;
  int32_t* slot_lin;
  int64_t nslots;
  double *B, *A;
  uint8_t* upd;
  int64_t pool_cap;
  ocam* cams;
  int ncams, capcams;
  int dirty;
};

int orc_create(const orc_params* p, orc_ctx** out) {
  if (!p || !out || !(p->res > 0) || (p->lookup_n != 2 && p->lookup_n != 3)) return -1;
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->p = *p;
  c->res = (double)p->res;
  for (int a = 0; a < 3; ++a) {
    c->Mx[a] = (int)lround((double)p->dims[a] / (double)p->res);
    if (c->Mx[a] <= 0 || c->Mx[a] > 2048) {
      free(c);
      return -7;
    }
    c->N[a] = (c->Mx[a] + 7) & ~7;
    c->B3[a] = c->N[a] >> 3;
    c->sig[a] = p->sigma_poly[a];
    c->bia[a] = p->bias_poly[a];
  }
  c->off[0] = 0.5 * c->Mx[0];
  c->off[1] = 0.5 * c->Mx[1];
  c->off[2] = 0.0;
  if (p->sigma_lookup && p->bias_lookup) {
    if (p->rows <= 0 || p->cols <= 0) {
      free(c);
      return -1;
    }
    const size_t n = (size_t)((p->rows + LOOKUP_PX - 1) / LOOKUP_PX) * ((p->cols + LOOKUP_PX - 1) / LOOKUP_PX) *
                     (size_t)p->lookup_n;
    c->sig_tab = (float*)malloc(sizeof(float) * n);
    c->bia_tab = (float*)malloc(sizeof(float) * n);
    memcpy(c->sig_tab, p->sigma_lookup, sizeof(float) * n);
    memcpy(c->bia_tab, p->bias_lookup, sizeof(float) * n);
  }
  c->p.sigma_lookup = c->p.bias_lookup = NULL;
  c->nbricks = c->B3[0] * c->B3[1] * c->B3[2];
  c->slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)c->nbricks);
  for (int i = 0; i < c->nbricks; ++i) c->slot[i] = -1;
  c->slot_lin = (int32_t*)malloc(sizeof(int32_t) * (size_t)c->nbricks);
  /* logit of the prior in float, as the reference computes it (GVDBMRFMap.cpp:51) */
  c->prior_lo = (double)(logf(p->prior) - logf(1.0f - p->prior));
  c->dirty = 1;
  *out = c;
  return 0;
}

static void free_cam_lists(ocam* k) {
  free(k->rays);
  free(k->rb.rec);
  free(k->rb.len);
  free(k->rb.zex);
  free(k->rb.vxyz);
  memset(&k->rb, 0, sizeof(k->rb));
  k->rays = NULL;
  k->nrays = 0;
  free(k->pair_vox);
  free(k->pair_c);
  free(k->rpair);
  k->pair_vox = NULL;
  k->pair_c = NULL;
  k->rpair = NULL;
  k->npairs = 0;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  for (int i = 0; i < c->ncams; ++i) {
    free_cam_lists(&c->cams[i]);
    free(c->cams[i].depth);
    free(c->cams[i].M);
  }
  free(c->cams);
  free(c->sig_tab);
  free(c->bia_tab);
  free(c->slot);
  free(c->slot_lin);
  free(c->B);
  free(c->A);
  free(c->upd);
  free(c);
}

static void set_pose(ocam* k, const float pose[16]) {
  memcpy(k->pose, pose, sizeof(k->pose));
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) k->R[i * 3 + j] = (double)pose[i * 4 + j];
    k->t[i] = (double)pose[i * 4 + 3];
  }
}

int orc_add_camera(orc_ctx* c, const float pose[16], const float K[9], const float* depth, int rows, int cols) {
  if (!c || !pose || !K || !depth || rows <= 0 || cols <= 0) return -1;
  if (c->sig_tab && (rows != c->p.rows || cols != c->p.cols)) return -1; /* tables are per params image */
  if (c->ncams == c->capcams) {
    c->capcams = c->capcams ? 2 * c->capcams : 16;
    c->cams = (ocam*)realloc(c->cams, sizeof(ocam) * (size_t)c->capcams);
  }
  ocam* k = &c->cams[c->ncams++];
  memset(k, 0, sizeof(*k));
  set_pose(k, pose);
  memcpy(k->K, K, sizeof(k->K));
  k->fx = K[0];
  k->fy = K[4];
  k->cx = K[2];
  k->cy = K[5];
  k->rows = rows;
  k->cols = cols;
  k->lookup_u = (cols + LOOKUP_PX - 1) / LOOKUP_PX;
  k->depth = (float*)malloc(sizeof(float) * (size_t)rows * cols);
  memcpy(k->depth, depth, sizeof(float) * (size_t)rows * cols);
  c->dirty = 1;
  return c->ncams - 1;
}

/* set_pose (GVDBInference.cpp:66-82): the camera's rays are removed from the
 * summed chord lengths and every voxel that no remaining ray reaches (summed
 * length of the other cameras' traversals < 0.05) returns to the prior with
 * its marginal reading 0 and the camera's message cleared
 * (reset_empty_voxels, gvdbkernels.cu:498-517); then the pose changes and
 * the camera is traversed again at the next inference.  The reference reads
 * the lengths from its single cum-len channel, which holds only the last
 * camera pass's lengths at that point; the other cameras' summed lengths are
 * what its comment ("voxels that have no ray passing through them") asks for. */
int orc_set_pose(orc_ctx* c, int cam, const float pose[16]) {
  if (!c || cam < 0 || cam >= c->ncams) return -5;
  const int64_t nv = c->nslots * 512;
  if (nv > 0) {
    double* cum = (double*)calloc((size_t)nv, sizeof(double));
    for (int ci = 0; ci < c->ncams; ++ci) {
      if (ci == cam) continue;
      const ocam* k = &c->cams[ci];
      for (int64_t q = 0; q < k->npairs; ++q) cum[k->pair_vox[q]] += k->pair_c[q];
    }
    double* M = c->cams[cam].M;
    for (int64_t v = 0; v < nv; ++v)
      if (cum[v] < 0.05) {
        c->B[v] = c->prior_lo;
        c->upd[v] = 0;
        if (M) M[v] = 0.0;
      }
    free(cum);
  }
  set_pose(&c->cams[cam], pose);
  c->dirty = 1;
  return 0;
}

static int pixel_ok(const orc_ctx* c, const ocam* k, int x, int y, float d) {
  if (!(d > 0.0f)) return 0;
  if (c->p.mask_848_border && k->cols == 848 && k->rows == 480 &&
      (x < 24 || x >= k->cols - 24 || y < 20 || y >= k->rows - 20))
    return 0;
  return 1;
}

/* offset of pixel (x, y)'s sensor-model cell: 20 x 20 pixel cells of the
 * vertically flipped image, lookup_n coefficients each (gvdbkernels.cu:46-55) */
static uint32_t cell_offset(const ocam* k, int ln, int x, int y) {
  const int v = k->rows - 1 - y;
  return (uint32_t)((v / LOOKUP_PX) * (k->lookup_u * ln) + (x / LOOKUP_PX) * ln);
}

/* sigma and bias polynomials of the cell at `cell` (the constant ones without tables) */
static void cell_polys(const orc_ctx* c, uint32_t cell, double sig[3], double bia[3]) {
  const int ln = c->p.lookup_n;
  for (int j = 0; j < 3; ++j) {
    if (!c->sig_tab) {
      sig[j] = c->sig[j];
      bia[j] = c->bia[j];
    } else {
      sig[j] = j < ln ? (double)c->sig_tab[cell + (uint32_t)j] : 0.0;
      bia[j] = j < ln ? (double)c->bia_tab[cell + (uint32_t)j] : 0.0;
    }
  }
}

static int setup_ray(const orc_ctx* c, const ocam* k, int x, int y, int64_t Oi[3], int64_t Di[3], double* norm) {
  double ux = ((double)x - k->cx) / k->fx;
  double uy = ((double)y - k->cy) / k->fy;
  double n2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    double dw = k->R[a * 3 + 0] * ux;
    dw = dw + k->R[a * 3 + 1] * uy;
    dw = dw + k->R[a * 3 + 2];
    double D = dw / c->res;
    double O = k->t[a] / c->res;
    O = O + c->off[a];
    double oi = rint(O * FIXS), di = rint(D * FIXS);
    if (!(fabs(oi) < 536870912.0) || !(fabs(di) < 67108864.0)) return 0;
    Oi[a] = (int64_t)oi;
    Di[a] = (int64_t)di;
    double dd = (double)Di[a] / FIXS;
    n2 = n2 + dd * dd;
  }
  *norm = sqrt(n2);
  return *norm > 0.0;
}

static void ensure_pool(orc_ctx* c) {
  int64_t need = c->nslots * 512;
  if (need <= c->pool_cap) return;
  c->B = (double*)realloc(c->B, sizeof(double) * (size_t)need);
  c->A = (double*)realloc(c->A, sizeof(double) * (size_t)need);
  c->upd = (uint8_t*)realloc(c->upd, (size_t)need);
  for (int64_t i = c->pool_cap; i < need; ++i) {
    c->B[i] = c->prior_lo;
    c->A[i] = 0;
    c->upd[i] = 0;
  }
  c->pool_cap = need;
}

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

int orc_prepare(orc_ctx* c) {
  if (!c) return -1;
  if (!c->dirty) return 0;
  const int ln = c->p.lookup_n;
  uint8_t* mark = (uint8_t*)calloc((size_t)c->nbricks, 1);
  brickmap bm = {c->B3[0], c->B3[1], c->B3[2], NULL, mark};
  segv sv = {0};
  /* activation */
  for (int ci = 0; ci < c->ncams; ++ci) {
    ocam* k = &c->cams[ci];
    for (int y = 0; y < k->rows; ++y)
      for (int x = 0; x < k->cols; ++x) {
        float d = k->depth[(size_t)y * k->cols + x];
        int64_t Oi[3], Di[3];
        double norm;
        if (!pixel_ok(c, k, x, y, d) || !setup_ray(c, k, x, y, Oi, Di, &norm)) continue;
        if (c->p.activation == 1) {
          double sg[3], bi[3];
          cell_polys(c, cell_offset(k, ln, x, y), sg, bi);
          ray_segments(Oi, Di, c->N, &sv);
          float z0;
          int ro;
          walk_filtered(&sv, &bm, norm, (double)d, sg, bi, ln, NULL, &z0, &ro);
        } else {
          int lo[3], hi[3], ok = 1;
          for (int a = 0; a < 3 && ok; ++a) {
            double P = (double)d * (double)Di[a];
            P = (double)Oi[a] + P;
            P = P / FIXS;
            if (!(P >= 0.0 && P < (double)c->Mx[a])) {
              ok = 0;
              break;
            }
            int l = (int)floor(P - 1.0), h = (int)floor(P + 1.0);
            if (l < 0) l = 0;
            if (h > c->N[a] - 1) h = c->N[a] - 1;
            lo[a] = l >> 3;
            hi[a] = h >> 3;
          }
          if (!ok) continue;
          for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
              for (int kk = lo[2]; kk <= hi[2]; ++kk) mark[(i * c->B3[1] + j) * c->B3[2] + kk] = 1;
        }
      }
  }
  for (int lin = 0; lin < c->nbricks; ++lin)
    if (mark[lin] && c->slot[lin] < 0) {
      c->slot[lin] = (int32_t)c->nslots;
      c->slot_lin[c->nslots++] = lin;
    }
  free(mark);
  ensure_pool(c);
  /* traversal */
  bm.slot = c->slot;
  bm.mark = NULL;
  double* csum = (double*)calloc((size_t)c->nslots * 512, sizeof(double));
  uint32_t* vpair = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(c->nslots * 512 + 1));
  uint32_t* touched = NULL;
  int64_t ntouched = 0, captouched = 0;
  for (int ci = 0; ci < c->ncams; ++ci) {
    ocam* k = &c->cams[ci];
    free_cam_lists(k);
    int64_t cap = 1024;
    k->rays = (uint32_t*)malloc(sizeof(uint32_t) * 6 * (size_t)cap);
    /* rays are listed in 8x8 pixel tiles (tiles row-major, pixels row-major
       within a tile), the order the traversal contract fixes */
    const int ntx = (k->cols + 7) / 8, nty = (k->rows + 7) / 8;
    for (int tile = 0; tile < ntx * nty; ++tile)
      for (int l = 0; l < 64; ++l) {
        const int x = (tile % ntx) * 8 + (l % 8), y = (tile / ntx) * 8 + (l / 8);
        if (x >= k->cols || y >= k->rows) continue;
        float d = k->depth[(size_t)y * k->cols + x];
        int64_t Oi[3], Di[3];
        double norm;
        if (!pixel_ok(c, k, x, y, d) || !setup_ray(c, k, x, y, Oi, Di, &norm)) continue;
        const uint32_t cell = cell_offset(k, ln, x, y);
        double sg[3], bi[3];
        cell_polys(c, cell, sg, bi);
        ray_segments(Oi, Di, c->N, &sv);
        int64_t off = k->rb.n;
        float z0;
        int ro;
        int64_t n = walk_filtered(&sv, &bm, norm, (double)d, sg, bi, ln, &k->rb, &z0, &ro);
        if (n == 0) continue;
        if (k->nrays == cap) {
          cap *= 2;
          k->rays = (uint32_t*)realloc(k->rays, sizeof(uint32_t) * 6 * (size_t)cap);
        }
        uint32_t* r = &k->rays[6 * k->nrays++];
        float inv = (float)(1.0 / norm);
        r[0] = (uint32_t)off;
        r[1] = (uint32_t)n;
        memcpy(&r[2], &z0, 4);
        memcpy(&r[3], &inv, 4);
        memcpy(&r[4], &d, 4);
        r[5] = cell | (ro ? RAN_OUT : 0u);
      }
    /* pairs: summed chord length per touched voxel */
    ntouched = 0;
    for (int64_t i = 0; i < k->rb.n; ++i) {
      uint32_t vox = k->rb.rec[2 * i];
      if (vox == GAP) continue;
      if (csum[vox] == 0.0) {
        if (ntouched == captouched) {
          captouched = captouched ? 2 * captouched : 4096;
          touched = (uint32_t*)realloc(touched, sizeof(uint32_t) * (size_t)captouched);
        }
        touched[ntouched++] = vox;
      }
      /* C sums the chord lengths the records carry, rounded to float as the
       * traversal contract defines them (render_cumulative_length_kernel adds
       * float lengths, gvdbkernels.cu:555-557); the messages keep FP64 lengths */
      csum[vox] += (double)(float)k->rb.len[i];
    }
    qsort(touched, (size_t)ntouched, sizeof(uint32_t), cmp_u32);
    k->npairs = ntouched;
    k->pair_vox = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ntouched ? ntouched : 1));
    k->pair_c = (double*)malloc(sizeof(double) * (size_t)(ntouched ? ntouched : 1));
    for (int64_t i = 0; i < ntouched; ++i) {
      k->pair_vox[i] = touched[i];
      k->pair_c[i] = csum[touched[i]];
      csum[touched[i]] = 0.0;
      vpair[touched[i]] = (uint32_t)i;
    }
    k->rpair = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(k->rb.n ? k->rb.n : 1));
    for (int64_t i = 0; i < k->rb.n; ++i) {
      const uint32_t vox = k->rb.rec[2 * i];
      k->rpair[i] = vox == GAP ? GAP : vpair[vox];
    }
    free(k->M);
    k->M = (double*)calloc((size_t)c->nslots * 512, sizeof(double));
  }
  free(touched);
  free(csum);
  free(vpair);
  free(sv.s);
  c->dirty = 0;
  return 0;
}

int64_t orc_num_bricks(orc_ctx* c) {
  if (!c || orc_prepare(c)) return -1;
  return c->nslots;
}

int orc_brick_coords(orc_ctx* c, int32_t* xyz) {
  if (!c || orc_prepare(c)) return -1;
  for (int64_t s = 0; s < c->nslots; ++s) {
    int lin = c->slot_lin[s];
    xyz[3 * s + 2] = lin % c->B3[2];
    xyz[3 * s + 1] = (lin / c->B3[2]) % c->B3[1];
    xyz[3 * s + 0] = lin / (c->B3[2] * c->B3[1]);
  }
  return 0;
}

int orc_get_traversal(orc_ctx* c, int cam, int64_t* n_rays, int64_t* n_records, uint32_t* rays,
                      uint32_t* records) {
  if (!c || cam < 0 || cam >= c->ncams) return -5;
  if (orc_prepare(c)) return -1;
  ocam* k = &c->cams[cam];
  if (n_rays) *n_rays = k->nrays;
  if (n_records) *n_records = k->rb.n;
  if (rays) memcpy(rays, k->rays, sizeof(uint32_t) * 6 * (size_t)k->nrays);
  if (records) memcpy(records, k->rb.rec, sizeof(uint32_t) * 2 * (size_t)k->rb.n);
  return 0;
}

/* ------------------------------------------------------------ messages */

static double* scratch_buf(int n) {
  static __thread double* buf = NULL;
  static __thread int cap = 0;
  if (n > cap) {
    cap = n * 2;
    buf = (double*)realloc(buf, sizeof(double) * 3 * (size_t)cap);
  }
  return buf;
}

/* The reference's arithmetic runs in FP32 with denormals flushed to zero
 * (--use_fast_math, CMakeLists.txt:192): the pre pass's tau_i = (V_i (1 -
 * n_i)) (0.398942 e^{-u^2/2} / (3 sigma)) (gvdbkernels.cu:57-72, 792-795) is
 * exactly 0 as soon as one of its products falls below the smallest normal
 * float, 2^-126.  This is that rule in FP64: tau_i counts as non-zero when
 * a = 0.398942 e^{-u^2/2}, psi = a / (3 sigma), b = V (1 - n) and tau = b psi
 * are all at least 2^-126. */
#define FTZ_MIN 0x1p-126

static int tau_nonzero_ftz(double sv, double d, const double sig[3], const double bia[3], int ln, double V,
                           double n, double* tau) {
  const double diff = fabs(d - (sv + poly(bia, ln, sv)));
  const double sg3 = 3.0 * poly(sig, ln, sv);
  const double u = diff / sg3;
  const double a = 0.398942 * exp(-0.5 * u * u);
  const double ps = a / sg3;
  const double b = V * (1.0 - n);
  *tau = b * ps;
  return a >= FTZ_MIN && ps >= FTZ_MIN && b >= FTZ_MIN && *tau >= FTZ_MIN;
}

double orc_ray_messages(int n, const double* nin, const double* len, const double* s, const int* live, double d,
                        const double sig[3], const double bia[3], int lookup_n, int ran_out, double psi_inf,
                        int mode, double* p) {
  return orc_ray_messages_ex(n, nin, len, s, live, NULL, d, sig, bia, lookup_n, ran_out, psi_inf, mode, p);
}

double orc_ray_messages_ex(int n, const double* nin, const double* len, const double* s, const int* live,
                           const int* bstart, double d, const double sig[3], const double bia[3], int lookup_n,
                           int ran_out, double psi_inf, int mode, double* p) {
  if (n <= 0) return 0.0;
  double* buf = scratch_buf(n);
  double *tau = buf, *vpsi = buf + n, *R = buf + 2 * n;
  double L = 0.0, T = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!live[i]) {
      tau[i] = vpsi[i] = 0.0;
      continue;
    }
    double V;
    if (mode == 0) { /* visibility includes voxel i itself, weighted by length (gvdbkernels.cu:781-783) */
      L += len[i] * log(nin[i]);
      V = exp(L);
    } else { /* textbook: product over j < i, unit weights */
      V = exp(L);
      L += log(nin[i]);
    }
    double ps = orc_depth_psi(s[i], d, sig, bia, lookup_n);
    vpsi[i] = V * ps;
    tau[i] = vpsi[i] * (1.0 - nin[i]);
    T += tau[i];
  }
  double tinf = 0.0;
  int inf_term = 0;
  if (mode == 0) {
    /* gvdbkernels.cu:1196-1208: only when the walk ran out, and only if d lies beyond */
    if (ran_out) {
      double sn = s[n - 1];
      inf_term = d > sn + poly(bia, lookup_n, sn);
    }
  } else {
    inf_term = 1;
  }
  if (inf_term) tinf = exp(L) * psi_inf;
  T += tinf;
  if (mode == 0 && bstart) {
    /* The reference's pre pass (gvdbkernels.cu:689-819) carries its
     * visibility sum from brick to brick only once the accumulated tau is
     * non-zero (the restore at :708-715 tests clr.x != 0): until then every
     * brick starts again from V = 1, and T, the infinity term (:1196-1208)
     * included, is summed with these brick-local visibilities.  The post
     * pass above carries along the whole ray (:844-856) and takes R_i =
     * T - sum_{j<=i} tau_j from this T. */
    double Lp = 0.0, Tp = 0.0;
    int nz = 0;
    for (int i = 0; i < n; ++i) {
      if (!live[i]) continue;
      if (bstart[i] && !nz) Lp = 0.0;
      Lp += len[i] * log(nin[i]);
      double tp;
      if (tau_nonzero_ftz(s[i], d, sig, bia, lookup_n, exp(Lp), nin[i], &tp)) {
        Tp += tp;
        nz = 1;
      }
    }
    if (inf_term) Tp += exp(Lp) * psi_inf;
    T = Tp;
    double pre = 0.0;
    for (int i = 0; i < n; ++i) {
      pre += tau[i];
      R[i] = T - pre;
    }
  } else {
    /* suffix sums R_i = sum_{j>i} tau_j + tinf */
    double acc = tinf;
    for (int i = n - 1; i >= 0; --i) {
      R[i] = acc;
      acc += tau[i];
    }
  }
  double P = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!live[i]) {
      p[i] = 0.0;
      continue;
    }
    double mu1 = P + vpsi[i];
    double mu0 = P + R[i] / nin[i];
    double q;
    if (mode == 0) {
      q = mu1 / (mu1 + mu0 + 1e-7);
      q = q < F_EPS ? F_EPS : (q > 1.0 - F_EPS ? 1.0 - F_EPS : q);
    } else {
      q = mu1 / (mu1 + mu0);
    }
    p[i] = q;
    P += tau[i];
  }
  return T;
}

int orc_ray_messages_bruteforce(int n, const double* nin, const double* psi, double psi_inf, double* p) {
  if (n <= 0 || n > 20) return -1;
  double* m1 = (double*)calloc((size_t)n, sizeof(double));
  double* m0 = (double*)calloc((size_t)n, sizeof(double));
  for (uint32_t o = 0; o < (1u << n); ++o) {
    int first = -1;
    for (int k = 0; k < n; ++k)
      if (o >> k & 1u) {
        first = k;
        break;
      }
    double f = first < 0 ? psi_inf : psi[first];
    for (int i = 0; i < n; ++i) {
      double w = f;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        w *= (o >> j & 1u) ? (1.0 - nin[j]) : nin[j];
      }
      if (o >> i & 1u)
        m1[i] += w;
      else
        m0[i] += w;
    }
  }
  for (int i = 0; i < n; ++i) p[i] = m1[i] / (m1[i] + m0[i]);
  free(m1);
  free(m0);
  return 0;
}

/* ------------------------------------------------------------ BP schedule */

static double marginal(double b) {
  double neg = 1.0 / (1.0 + exp(b));
  neg = neg < F_EPS ? F_EPS : (neg > 1.0 - F_EPS ? 1.0 - F_EPS : neg);
  return 1.0 - neg;
}

static double clampd(double x) { return x < F_EPS ? F_EPS : (x > 1.0 - F_EPS ? 1.0 - F_EPS : x); }

static int bp_iterate(orc_ctx* c, int max_iters, double tol, double* residual, int64_t* updates_per_iter);

int orc_perform_inference(orc_ctx* c, const int* ids, int n_ids, int max_iters, double tol, double* residual,
                          int64_t* updates_per_iter) {
  if (!c) return -1;
  if (orc_prepare(c)) return -1;
  const int64_t nv = c->nslots * 512;
  if (ids) {
    for (int i = 0; i < n_ids; ++i) {
      if (ids[i] < 0 || ids[i] >= c->ncams) return -5;
      memset(c->cams[ids[i]].M, 0, sizeof(double) * (size_t)nv);
    }
  } else {
    for (int i = 0; i < c->ncams; ++i) memset(c->cams[i].M, 0, sizeof(double) * (size_t)nv);
  }
  return bp_iterate(c, max_iters, tol, residual, updates_per_iter);
}

int orc_continue_inference(orc_ctx* c, int max_iters, double tol, double* residual) {
  if (!c) return -1;
  if (orc_prepare(c)) return -1;
  return bp_iterate(c, max_iters, tol, residual, NULL);
}

/* Threads of the BP loop (0: OpenMP's default, i.e. OMP_NUM_THREADS or
 * every core).  The rays of a camera pass are split into one contiguous block
 * per thread; each thread sums its rays' p * l per (camera, voxel) pair in an
 * array of its own, and the per-pair sums S add the threads' arrays in
 * thread order, so a result depends on the thread count only through FP64
 * rounding and is reproducible for a given count. */
static int g_threads = 0;

void orc_set_threads(int n) { g_threads = n < 0 ? 0 : n; }

int orc_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* Messages of rays [r0, r1) of camera k into St (p * l summed per pair). */
static void camera_rays(const orc_ctx* c, const ocam* k, int64_t r0, int64_t r1, double* St, double* nin,
                        double* pout, int* live, int* bst) {
  const int ln = c->p.lookup_n;
  for (int64_t r = r0; r < r1; ++r) {
    const uint32_t* rr = &k->rays[6 * r];
    const int64_t off = rr[0];
    const int n = (int)rr[1];
    float d;
    memcpy(&d, &rr[4], 4);
    const int ro = (rr[5] & RAN_OUT) != 0;
    for (int i = 0; i < n; ++i) {
      uint32_t vox = k->rb.rec[2 * (off + i)];
      live[i] = vox != GAP;
      nin[i] = live[i] ? clampd(1.0 / (1.0 + exp(c->B[vox] - k->M[vox]))) : 1.0;
      /* the record enters a brick: first of the ray, after a gap, or in another slot */
      bst[i] = live[i] && (i == 0 || !live[i - 1] || (k->rb.rec[2 * (off + i - 1)] >> 9) != (vox >> 9));
    }
    double sg[3], bi[3];
    cell_polys(c, rr[5] & ~RAN_OUT, sg, bi);
    orc_ray_messages_ex(n, nin, k->rb.len + off, k->rb.zex + off, live, c->p.brick_local_vis ? bst : NULL,
                        (double)d, sg, bi, ln, ro, 1.0, 0, pout);
    for (int i = 0; i < n; ++i)
      if (live[i]) St[k->rpair[off + i]] += pout[i] * k->rb.len[off + i];
  }
}

/* BP iterations from the current beliefs and messages (no reset). */
static int bp_iterate(orc_ctx* c, int max_iters, double tol, double* residual, int64_t* updates_per_iter) {
  const int64_t nv = c->nslots * 512;
  const int nt = orc_get_threads();
  int64_t upd = 0;
  for (int ci = 0; ci < c->ncams; ++ci)
    for (int64_t i = 0; i < c->cams[ci].rb.n; ++i) upd += c->cams[ci].rb.rec[2 * i] != GAP;
  if (updates_per_iter) *updates_per_iter = upd;
  for (int64_t v = 0; v < nv; ++v) c->A[v] = marginal(c->B[v]);
  /* per thread: the pair sums and the per-ray work arrays */
  int64_t np = 1;
  for (int ci = 0; ci < c->ncams; ++ci)
    if (c->cams[ci].npairs > np) np = c->cams[ci].npairs;
  double* Sall = (double*)calloc((size_t)np * (size_t)nt, sizeof(double));
  int maxn = 1;
  for (int ci = 0; ci < c->ncams; ++ci)
    for (int64_t r = 0; r < c->cams[ci].nrays; ++r)
      if ((int)c->cams[ci].rays[6 * r + 1] > maxn) maxn = (int)c->cams[ci].rays[6 * r + 1];
  double* dbuf = (double*)malloc(sizeof(double) * 2 * (size_t)maxn * (size_t)nt);
  int* ibuf = (int*)malloc(sizeof(int) * 2 * (size_t)maxn * (size_t)nt);
  double res = 0.0;
  int it;
  for (it = 0; it < max_iters; ++it) {
    for (int ci = 0; ci < c->ncams; ++ci) {
      ocam* k = &c->cams[ci];
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
      {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        const int64_t r0 = k->nrays * t / nt, r1 = k->nrays * (t + 1) / nt;
        double* nin = dbuf + (size_t)t * 2 * maxn;
        int* live = ibuf + (size_t)t * 2 * maxn;
        camera_rays(c, k, r0, r1, Sall + (size_t)t * np, nin, nin + maxn, live, live + maxn);
      }
      /* a camera's pairs are distinct voxels: its updates are independent */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
      for (int64_t q = 0; q < k->npairs; ++q) {
        const uint32_t v = k->pair_vox[q];
        const double C = k->pair_c[q];
        double s = 0.0;
        for (int t = 0; t < nt; ++t) {  /* thread order */
          s += Sall[(size_t)t * np + q];
          Sall[(size_t)t * np + q] = 0.0;
        }
        if (C > CUM_LEN_GATE) {
          double a = clampd(s / C);
          double m = log(a) - log(1.0 - a);
          c->B[v] = (c->B[v] + m) - k->M[v];
          k->M[v] = m;
          c->upd[v] = 1;
        }
      }
    }
    res = 0.0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static) reduction(max : res)
#endif
    for (int64_t v = 0; v < nv; ++v) {
      double a = marginal(c->B[v]);
      double dl = fabs(a - c->A[v]);
      if (dl > res) res = dl;
      c->A[v] = a;
    }
    if (tol > 0 && res < tol) {
      ++it;
      break;
    }
  }
  free(Sall);
  free(dbuf);
  free(ibuf);
  if (residual) *residual = res;
  return it;
}

/* ------------------------------------------------------------ warm start
 * Features of the learned warm start, as mrfmap_amd/csrc/warm.hip defines
 * them (no counterpart in the reference), from the current state in FP64:
 * one row of 32 per voxel with C_c > 0.1 for some camera, in pool order. */
#define ORC_LF 9.2102404

static int64_t pool_index(const orc_ctx* c, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= c->N[0] || y >= c->N[1] || z >= c->N[2]) return -1;
  const int s = c->slot[((x >> 3) * c->B3[1] + (y >> 3)) * c->B3[2] + (z >> 3)];
  if (s < 0) return -1;
  return (int64_t)s * 512 + ((x & 7) * 64 + (y & 7) * 8 + (z & 7));
}

int orc_features(orc_ctx* c, float* X, uint32_t* vox, int64_t cap, int64_t* n_out) {
  if (!c || orc_prepare(c)) return -1;
  const int64_t nv = c->nslots * 512;
  double* pass = (double*)calloc((size_t)nv, sizeof(double));
  double* msum = (double*)calloc((size_t)nv, sizeof(double));
  double* mmax = (double*)calloc((size_t)nv, sizeof(double));
  uint32_t* nc = (uint32_t*)calloc((size_t)nv, sizeof(uint32_t));
  uint32_t* hits = (uint32_t*)calloc((size_t)nv, sizeof(uint32_t));
  for (int ci = 0; ci < c->ncams; ++ci) {
    ocam* k = &c->cams[ci];
    for (int64_t q = 0; q < k->npairs; ++q) {
      const uint32_t v = k->pair_vox[q];
      const double C = k->pair_c[q];
      pass[v] += C;
      if (C > CUM_LEN_GATE) {
        const double m = k->M[v];
        msum[v] += m;
        mmax[v] = nc[v] ? (m > mmax[v] ? m : mmax[v]) : m;
        nc[v] += 1;
      }
    }
    /* depth endpoints: the activation's endpoint formula and map-box gate */
    for (int y = 0; y < k->rows; ++y)
      for (int x = 0; x < k->cols; ++x) {
        float d = k->depth[(size_t)y * k->cols + x];
        int64_t Oi[3], Di[3];
        double norm;
        if (!pixel_ok(c, k, x, y, d) || !setup_ray(c, k, x, y, Oi, Di, &norm)) continue;
        int v3[3], ok = 1;
        for (int a = 0; a < 3; ++a) {
          double P = (double)d * (double)Di[a];
          P = (double)Oi[a] + P;
          P = P / FIXS;
          if (!(P >= 0.0 && P < (double)c->Mx[a])) {
            ok = 0;
            break;
          }
          v3[a] = (int)floor(P);
        }
        if (!ok) continue;
        const int64_t pv = pool_index(c, v3[0], v3[1], v3[2]);
        if (pv >= 0) hits[pv] += 1;
      }
  }
  int64_t n = 0;
  for (int64_t v = 0; v < nv; ++v) {
    if (nc[v] == 0) continue;
    if (n < cap) {
      const int lin = c->slot_lin[v / 512], t = (int)(v % 512);
      const int x = (lin / (c->B3[1] * c->B3[2])) * 8 + (t >> 6);
      const int y = ((lin / c->B3[2]) % c->B3[1]) * 8 + ((t >> 3) & 7);
      const int z = (lin % c->B3[2]) * 8 + (t & 7);
      float* r = X + n * 32;
      r[0] = (float)log1p((double)hits[v]);
      r[1] = (float)log1p(pass[v]);
      r[2] = (float)((msum[v] / nc[v]) / ORC_LF);
      r[3] = (float)(mmax[v] / ORC_LF);
      int best = 16;
      for (int a = -3; a <= 3; ++a)
        for (int b = -3; b <= 3; ++b)
          for (int e = -3; e <= 3; ++e) {
            const int d2 = a * a + b * b + e * e;
            if (d2 >= best) continue;
            const int64_t u = pool_index(c, x + a, y + b, z + e);
            if (u >= 0 && hits[u] > 0) best = d2;
          }
      r[4] = (float)(sqrt((double)best) / 4.0);
      int f = 5;
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int e = -1; e <= 1; ++e) {
            const int64_t u = pool_index(c, x + a, y + b, z + e);
            double q = (u >= 0 ? c->B[u] : c->prior_lo) / ORC_LF;
            r[f++] = (float)(q < -3.0 ? -3.0 : (q > 3.0 ? 3.0 : q));
          }
      vox[n] = (uint32_t)v;
    }
    ++n;
  }
  free(pass);
  free(msum);
  free(mmax);
  free(nc);
  free(hits);
  *n_out = n;
  return n <= cap ? 0 : -7;
}

/* Warm start from predicted log-odds pred[i] of voxel vox[i]: the belief
 * moves to the prediction (kept where prediction and belief are saturated on
 * the same side and the belief is stronger), and every camera message with
 * C_c > 0.1 moves by the same share (B' - B) / n_c, so B = L0 + sum_c M_c
 * keeps holding; the marginals of the next residual start from B'. */
int orc_warm_start(orc_ctx* c, const uint32_t* vox, const double* pred, int64_t n) {
  if (!c || orc_prepare(c)) return -1;
  const int64_t nv = c->nslots * 512;
  uint32_t* nc = (uint32_t*)calloc((size_t)nv, sizeof(uint32_t));
  double* delta = (double*)calloc((size_t)nv, sizeof(double));
  for (int ci = 0; ci < c->ncams; ++ci)
    for (int64_t q = 0; q < c->cams[ci].npairs; ++q)
      if (c->cams[ci].pair_c[q] > CUM_LEN_GATE) nc[c->cams[ci].pair_vox[q]] += 1;
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t v = vox[i];
    if ((int64_t)v >= nv || nc[v] == 0) {
      free(nc);
      free(delta);
      return -1;
    }
    const double b = c->B[v], p = pred[i];
    double bn = p;
    if ((p <= -ORC_LF && b <= p) || (p >= ORC_LF && b >= p)) bn = b;
    delta[v] = (bn - b) / nc[v];
    c->B[v] = bn;
  }
  for (int ci = 0; ci < c->ncams; ++ci) {
    ocam* k = &c->cams[ci];
    for (int64_t q = 0; q < k->npairs; ++q)
      if (k->pair_c[q] > CUM_LEN_GATE) k->M[k->pair_vox[q]] += delta[k->pair_vox[q]];
  }
  free(nc);
  free(delta);
  return 0;
}

/* Map evaluation from a camera, FP64 (the device path is FP32): every pixel
 * ray walked through the current marginals to the end of the grid.
 * Restates rayLikelihoodBrick / render_likelihood_kernel
 * (gvdbkernels.cu:183-293, 1246-1274), rayExpectedDepthBrick /
 * render_expected_depth_kernel (:295-354, 1332-1363), render_accuracy_kernel
 * (:1066-1104) and raySimpleAccuracyBrick / render_simple_accuracy_kernel
 * (:135-181, 1106-1134).  Images rows x cols, row-major, any may be NULL. */
#define ORC_Z_MIN 0.1
#define ORC_Z_MAX 10.0
#define ORC_LIKE_EPS 1e-3

static double psi_d(double s, double d, const double sig[3], const double bia[3], int ln) {
  double sg3 = 3.0 * poly(sig, ln, s);
  double u = (d - s - poly(bia, ln, s)) / sg3;
  return 0.398942 * exp(-0.5 * u * u) / sg3;
}

int orc_evaluate(orc_ctx* c, const float pose[16], const float K[9], const float* depth, int rows, int cols,
                 double acc_thresh, double* ll, double* ed, double* acc, double* sacc) {
  if (!c || !pose || !K || rows <= 0 || cols <= 0 || orc_prepare(c)) return -1;
  ocam k;
  memset(&k, 0, sizeof(k));
  set_pose(&k, pose);
  k.fx = K[0];
  k.fy = K[4];
  k.cx = K[2];
  k.cy = K[5];
  k.rows = rows;
  k.cols = cols;
  k.lookup_u = (cols + LOOKUP_PX - 1) / LOOKUP_PX;
  const int ln = c->p.lookup_n;
  brickmap bm = {c->B3[0], c->B3[1], c->B3[2], c->slot, NULL};
  segv sv = {0};
  for (int y = 0; y < rows; ++y)
    for (int x = 0; x < cols; ++x) {
      const size_t o = (size_t)y * cols + x;
      const double d = depth ? (double)depth[o] : 0.0;
      const int valid = depth && pixel_ok(c, &k, x, y, (float)d);
      double prev_vis = 1.0, lam = 0.0, sum = 0.0, prev_s = 0.0, norm_x = 0.0, peak = 0.0, esum = 0.0;
      double sa = -1.0;
      int peak_ok = 0, first = 1, sdone = 0;
      int64_t Oi[3], Di[3];
      double norm;
      double csig[3], cbia[3];
      cell_polys(c, cell_offset(&k, ln, x, y), csig, cbia);
      if (setup_ray(c, &k, x, y, Oi, Di, &norm)) {
        ray_segments(Oi, Di, c->N, &sv);
        for (int j = 0; j < sv.n; ++j) {
          const seg_t* e = &sv.s[j];
          const int lin = brick_lin(&bm, e->v);
          if (c->slot[lin] < 0) continue;
          const double z0 = rdbl(e->z0), z1 = rdbl(e->z1);
          const double len = (double)(float)((z1 - z0) * norm);
          const uint32_t local = (uint32_t)(((e->v[0] & 7) * 8 + (e->v[1] & 7)) * 8 + (e->v[2] & 7));
          const int64_t v = (int64_t)c->slot[lin] * 512 + local;
          const double p = c->upd[v] ? marginal(c->B[v]) : 0.0;
          const double s = z1;
          if (first) {
            prev_s = z0;
            first = 0;
          }
          lam += log(1.0 - p) * len;
          const double vis = exp(lam);
          const double w = (prev_vis - vis) / len;
          if (valid) {
            const double pdm = s + poly(cbia, ln, s) - d;
            const double sg = poly(csig, ln, s);
            double prob;
            if (s < ORC_Z_MIN)
              prob = 0.0;
            else if (fabs(pdm) <= 3.0 * sg)
              prob = (1.0 - ORC_LIKE_EPS) * 0.5 * (psi_d(s, d, csig, cbia, ln) + psi_d(prev_s, d, csig, cbia, ln));
            else
              prob = ORC_LIKE_EPS;
            if (w > peak) {
              peak = w;
              peak_ok = fabs(pdm) <= sg;
            }
            sum += (prev_vis - vis) * prob;
            norm_x = pdm + d;
            if (!sdone && p >= acc_thresh) {
              sa = fabs(s - d) <= c->res ? 1.0 : 2.0;
              sdone = 1;
            }
          }
          double zc = 0.0;
          for (int a = 0; a < 3; ++a) zc += k.R[a * 3 + 2] * ((((double)e->v[a] + 0.5) - c->off[a]) * c->res - k.t[a]);
          esum += (prev_vis - vis) * zc;
          prev_vis = vis;
          prev_s = s;
          const int brick_exit = (j + 1 == sv.n) || brick_lin(&bm, sv.s[j + 1].v) != lin;
          if (brick_exit && valid) {
            prev_s = norm_x;
            if (!sdone && s >= d) {
              sa = 0.0;
              sdone = 1;
            }
          }
        }
      }
      if (ed) ed[o] = esum / (1.0 - prev_vis + F_EPS);
      double l = 0.0, a = -1.0;
      if (valid) {
        l = log(sum + prev_vis * (d >= ORC_Z_MAX ? 1.0 - ORC_LIKE_EPS : ORC_LIKE_EPS));
        if (peak > 0.0 && prev_vis >= 0.9)
          a = norm_x <= d ? -1.0 : 0.0;
        else if (peak > 0.0)
          a = peak_ok;
      } else {
        sa = -1.0;
      }
      if (ll) ll[o] = l;
      if (acc) acc[o] = a;
      if (sacc) sacc[o] = sa;
    }
  free(sv.s);
  return 0;
}

/* load_map (GVDBMapLikelihoodEstimator.cpp:168-240, set_channel_kernel
 * gvdbkernels.cu:1040-1064): bricks within one voxel of every point inside
 * the map box become active (new slots in ascending linear brick index);
 * the voxel holding a point gets belief logit(clamp(occupancy)), the last
 * point of a voxel winning. */
int orc_load_map(orc_ctx* c, const float* pts, const float* occ, int64_t n) {
  if (!c || n < 0 || orc_prepare(c)) return -1;
  uint8_t* mark = (uint8_t*)calloc((size_t)c->nbricks, 1);
  for (int64_t i = 0; i < n; ++i) {
    int lo[3], hi[3], ok = 1;
    for (int a = 0; a < 3 && ok; ++a) {
      double P = (double)pts[3 * i + a];
      if (!(P >= 0.0 && P < (double)c->Mx[a])) {
        ok = 0;
        break;
      }
      int l = (int)floor(P - 1.0), h = (int)floor(P + 1.0);
      if (l < 0) l = 0;
      if (h > c->N[a] - 1) h = c->N[a] - 1;
      lo[a] = l >> 3;
      hi[a] = h >> 3;
    }
    if (!ok) continue;
    for (int x = lo[0]; x <= hi[0]; ++x)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int z = lo[2]; z <= hi[2]; ++z) mark[(x * c->B3[1] + y) * c->B3[2] + z] = 1;
  }
  int grew = 0;
  for (int lin = 0; lin < c->nbricks; ++lin)
    if (mark[lin] && c->slot[lin] < 0) {
      c->slot[lin] = (int32_t)c->nslots;
      c->slot_lin[c->nslots++] = lin;
      grew = 1;
    }
  free(mark);
  ensure_pool(c);
  for (int64_t i = 0; i < n; ++i) {
    int v[3], ok = 1;
    for (int a = 0; a < 3 && ok; ++a) {
      double P = (double)pts[3 * i + a];
      if (!(P >= 0.0 && P < (double)c->Mx[a])) ok = 0;
      v[a] = (int)floor(P);
    }
    if (!ok) continue;
    const int lin = ((v[0] >> 3) * c->B3[1] + (v[1] >> 3)) * c->B3[2] + (v[2] >> 3);
    const int64_t vox = (int64_t)c->slot[lin] * 512 + ((v[0] & 7) * 8 + (v[1] & 7)) * 8 + (v[2] & 7);
    double o = (double)occ[i];
    o = o < F_EPS ? F_EPS : (o > 1.0 - F_EPS ? 1.0 - F_EPS : o);
    c->B[vox] = log(o) - log(1.0 - o);
    c->upd[vox] = 1;
  }
  if (grew) c->dirty = 1;
  return 0;
}

/* Test hook: overwrite the beliefs and update flags of the pool (nslots x 512). */
int orc_set_beliefs(orc_ctx* c, const double* B, const uint8_t* upd) {
  if (!c || orc_prepare(c)) return -1;
  memcpy(c->B, B, sizeof(double) * (size_t)(c->nslots * 512));
  memcpy(c->upd, upd, (size_t)(c->nslots * 512));
  return 0;
}

int orc_get_alphas(orc_ctx* c, double* out) {
  if (!c || orc_prepare(c)) return -1;
  for (int64_t v = 0; v < c->nslots * 512; ++v) out[v] = c->upd[v] ? marginal(c->B[v]) : 0.0;
  return 0;
}

int orc_get_beliefs(orc_ctx* c, double* out) {
  if (!c || orc_prepare(c)) return -1;
  memcpy(out, c->B, sizeof(double) * (size_t)(c->nslots * 512));
  return 0;
}

/* gvdbkernels.cu:519-571, 1276-1302: the CUM_LENS channel after camera
 * cam's cumulative-length pass, the FP64 sum of its chord lengths per voxel */
int orc_get_cum_lengths(orc_ctx* c, int cam, double* out) {
  if (!c || orc_prepare(c)) return -1;
  if (cam < 0 || cam >= c->ncams) return -5;
  memset(out, 0, sizeof(double) * (size_t)(c->nslots * 512));
  const ocam* k = &c->cams[cam];
  for (int64_t q = 0; q < k->npairs; ++q) out[k->pair_vox[q]] = k->pair_c[q];
  return 0;
}

int orc_walk_ray(const double O[3], const double D[3], const int N[3], double depth, int trunc, const double sig[3],
                 const double bia[3], int lookup_n, int cap, int32_t* vox, double* len, double* zexit,
                 int* ran_out) {
  int64_t Oi[3], Di[3];
  double n2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    Oi[a] = (int64_t)rint(O[a] * FIXS);
    Di[a] = (int64_t)rint(D[a] * FIXS);
    double dd = (double)Di[a] / FIXS;
    n2 = n2 + dd * dd;
  }
  double norm = sqrt(n2);
  segv sv = {0};
  ray_segments(Oi, Di, N, &sv);
  int m = 0, ro = 1;
  int B3[3] = {(N[0] + 7) >> 3, (N[1] + 7) >> 3, (N[2] + 7) >> 3};
  for (int j = 0; j < sv.n && m < cap; ++j) {
    const seg_t* e = &sv.s[j];
    double zc = rdbl(e->z0), zn = rdbl(e->z1);
    memcpy(&vox[3 * m], e->v, 12);
    len[m] = (zn - zc) * norm;
    zexit[m] = zn;
    ++m;
    if (trunc) {
      int lin = ((e->v[0] >> 3) * B3[1] + (e->v[1] >> 3)) * B3[2] + (e->v[2] >> 3);
      int brick_exit = (j + 1 == sv.n);
      if (!brick_exit) {
        const int32_t* w = sv.s[j + 1].v;
        brick_exit = (((w[0] >> 3) * B3[1] + (w[1] >> 3)) * B3[2] + (w[2] >> 3)) != lin;
      }
      if (brick_exit && zn + poly(bia, lookup_n, zn) >= depth + 3.0 * poly(sig, lookup_n, zn)) {
        ro = 0;
        break;
      }
    }
  }
  free(sv.s);
  if (ran_out) *ran_out = ro;
  return m;
}
