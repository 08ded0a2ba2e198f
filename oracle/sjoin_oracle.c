/*
 * sjoin_oracle.c -- CPU restatement of the reference CPU spatial join.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product path.  Every function below cites the reference file:line whose
 * behaviour it restates.  Build with -ffp-contract=off: the relation refine
 * must evaluate the same IEEE double expression sequence as the reference
 * (x86-64 build of tg.c, which has no FMA contraction).
 */
#include "sjoin_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* WKB decode                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
  int64_t n, cap;
  int64_t* v;
} vec64;

static int v64_push(vec64* a, int64_t x) {
  if (a->n == a->cap) {
    int64_t nc = a->cap ? a->cap * 2 : 1024;
    int64_t* nv = (int64_t*)realloc(a->v, (size_t)nc * sizeof(int64_t));
    if (!nv) return -1;
    a->v = nv;
    a->cap = nc;
  }
  a->v[a->n++] = x;
  return 0;
}

typedef struct {
  int64_t n, cap;
  double* v;
} vecd;

static int vd_push2(vecd* a, double x, double y) {
  if (a->n + 2 > a->cap) {
    int64_t nc = a->cap ? a->cap * 2 : 4096;
    double* nv = (double*)realloc(a->v, (size_t)nc * sizeof(double));
    if (!nv) return -1;
    a->v = nv;
    a->cap = nc;
  }
  a->v[a->n++] = x;
  a->v[a->n++] = y;
  return 0;
}

typedef struct {
  const uint8_t* p;
  const uint8_t* end;
} rd_t;

static int rd_u8(rd_t* r, uint8_t* out) {
  if (r->p + 1 > r->end) return -1;
  *out = *r->p++;
  return 0;
}

static int rd_u32(rd_t* r, int le, uint32_t* out) {
  if (r->p + 4 > r->end) return -1;
  uint32_t v;
  if (le) v = (uint32_t)r->p[0] | ((uint32_t)r->p[1] << 8) | ((uint32_t)r->p[2] << 16) | ((uint32_t)r->p[3] << 24);
  else v = (uint32_t)r->p[3] | ((uint32_t)r->p[2] << 8) | ((uint32_t)r->p[1] << 16) | ((uint32_t)r->p[0] << 24);
  r->p += 4;
  *out = v;
  return 0;
}

static int rd_f64(rd_t* r, int le, double* out) {
  if (r->p + 8 > r->end) return -1;
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) {
    int k = le ? i : 7 - i;
    v |= (uint64_t)r->p[k] << (8 * i);
  }
  r->p += 8;
  memcpy(out, &v, 8);
  return 0;
}

typedef struct {
  vec64 part_off, ring_off, pt_off;
  vecd xy;
} builder_t;

/* Reads one coordinate run of n points into the store as one ring. */
static int rd_ring(rd_t* r, int le, uint32_t n, builder_t* b) {
  for (uint32_t i = 0; i < n; i++) {
    double x, y;
    if (rd_f64(r, le, &x) || rd_f64(r, le, &y)) return -1;
    if (vd_push2(&b->xy, x, y)) return -1;
  }
  return v64_push(&b->pt_off, b->xy.n / 2);
}

/* Parses one WKB geometry; `expect` restricts nested types (0 = any). */
static int rd_geom(rd_t* r, builder_t* b, int nested_expect, uint8_t* type_out) {
  uint8_t bo;
  uint32_t t;
  if (rd_u8(r, &bo)) return -1;
  if (bo > 1) return -2;
  int le = bo == 1;
  if (rd_u32(r, le, &t)) return -1;
  if (t < 1 || t > 6) return -3; /* 2-D Point..MultiPolygon only */
  if (nested_expect && (int)t != nested_expect) return -3;
  *type_out = (uint8_t)t;
  uint32_t n, nr;
  switch (t) {
    case ORA_POINT: {
      double x, y;
      if (rd_f64(r, le, &x) || rd_f64(r, le, &y)) return -1;
      if (isnan(x) && isnan(y)) return 0; /* empty point: no part */
      if (vd_push2(&b->xy, x, y)) return -1;
      if (v64_push(&b->pt_off, b->xy.n / 2)) return -1;
      return v64_push(&b->ring_off, b->pt_off.n - 1);
    }
    case ORA_LINESTRING:
      /* c/sedona-tg/src/tg/tg.c:8380-8387: lines need two positions */
      if (rd_u32(r, le, &n)) return -1;
      if (n == 0) return 0;
      if (n < 2) return -4;
      if (rd_ring(r, le, n, b)) return -1;
      return v64_push(&b->ring_off, b->pt_off.n - 1);
    case ORA_POLYGON:
      if (rd_u32(r, le, &nr)) return -1;
      if (nr == 0) return 0;
      for (uint32_t i = 0; i < nr; i++) {
        if (rd_u32(r, le, &n)) return -1;
        if (rd_ring(r, le, n, b)) return -1;
        if (n == 0) continue;
        /* tg.c:8388-8399: rings need three positions, first == last */
        const double* q = b->xy.v + b->xy.n - 2 * (int64_t)n;
        const double* z = b->xy.v + b->xy.n - 2;
        if (n < 3) return -4;
        if (!(q[0] == z[0] && q[1] == z[1])) return -5;
      }
      return v64_push(&b->ring_off, b->pt_off.n - 1);
    default: {
      if (rd_u32(r, le, &n)) return -1;
      int child = (int)t - 3;
      for (uint32_t i = 0; i < n; i++) {
        uint8_t ct;
        int rc = rd_geom(r, b, child, &ct);
        if (rc) return rc;
      }
      return 0;
    }
  }
}

int ora_store_from_wkb(const int64_t* offsets, const uint8_t* data, const uint8_t* valid,
                       int64_t n, ora_store** out, int64_t* bad_row) {
  builder_t b;
  memset(&b, 0, sizeof(b));
  ora_store* s = (ora_store*)calloc(1, sizeof(ora_store));
  s->n_geoms = n;
  s->type = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
  v64_push(&b.part_off, 0);
  v64_push(&b.ring_off, 0);
  v64_push(&b.pt_off, 0);
  /* part_off/ring_off/pt_off hold prefix offsets: part_off[g] = first part. */
  for (int64_t g = 0; g < n; g++) {
    if (!valid || valid[g]) {
      rd_t r = {data + offsets[g], data + offsets[g + 1]};
      uint8_t t = 0;
      int rc = rd_geom(&r, &b, 0, &t);
      if (rc) {
        if (bad_row) *bad_row = g;
        free(b.part_off.v); free(b.ring_off.v); free(b.pt_off.v); free(b.xy.v);
        free(s->type); free(s);
        return rc;
      }
      s->type[g] = t;
    }
    /* number of parts so far == ring_off.n - 1 */
    v64_push(&b.part_off, b.ring_off.n - 1);
  }
  s->part_off = b.part_off.v;
  s->ring_off = b.ring_off.v;
  s->pt_off = b.pt_off.v;
  s->xy = b.xy.v ? b.xy.v : (double*)calloc(2, sizeof(double));
  s->n_parts = b.ring_off.n - 1;
  s->n_rings = b.pt_off.n - 1;
  s->n_pts = b.xy.n / 2;
  *out = s;
  return 0;
}

void ora_store_free(ora_store* s) {
  if (!s) return;
  free(s->type); free(s->part_off); free(s->ring_off); free(s->pt_off); free(s->xy);
  free(s);
}

void ora_store_sizes(const ora_store* s, int64_t* o) {
  o[0] = s->n_geoms; o[1] = s->n_parts; o[2] = s->n_rings; o[3] = s->n_pts;
}

void ora_store_export(const ora_store* s, uint8_t* type, int64_t* part_off, int64_t* ring_off,
                      int64_t* pt_off, double* xy) {
  memcpy(type, s->type, (size_t)s->n_geoms);
  memcpy(part_off, s->part_off, (size_t)(s->n_geoms + 1) * 8);
  memcpy(ring_off, s->ring_off, (size_t)(s->n_parts + 1) * 8);
  memcpy(pt_off, s->pt_off, (size_t)(s->n_rings + 1) * 8);
  memcpy(xy, s->xy, (size_t)s->n_pts * 16);
}

/* ------------------------------------------------------------------ */
/* Bounds and f32 rects                                                */
/* ------------------------------------------------------------------ */

/* rust/sedona-spatial-join/src/utils/bounds.rs:147-162: cast to nearest,
 * then step one f32 ulp outward when the cast moved inward. */
void ora_f64_interval_to_f32(double lo, double hi, float* out_lo, float* out_hi) {
  float l = (float)lo, h = (float)hi;
  if ((double)l > lo) l = nextafterf(l, -INFINITY);
  if ((double)h < hi) h = nextafterf(h, INFINITY);
  *out_lo = l;
  *out_hi = h;
}

static int geom_bounds(const ora_store* s, int64_t g, double* bx) {
  int64_t p0 = s->part_off[g], p1 = s->part_off[g + 1];
  if (p0 == p1) return 0;
  int64_t r0 = s->ring_off[p0], r1 = s->ring_off[p1];
  int64_t v0 = s->pt_off[r0], v1 = s->pt_off[r1];
  if (v0 == v1) return 0;
  double xmin = INFINITY, ymin = INFINITY, xmax = -INFINITY, ymax = -INFINITY;
  for (int64_t v = v0; v < v1; v++) {
    double x = s->xy[2 * v], y = s->xy[2 * v + 1];
    if (x < xmin) xmin = x;
    if (x > xmax) xmax = x;
    if (y < ymin) ymin = y;
    if (y > ymax) ymax = y;
  }
  bx[0] = xmin; bx[1] = ymin; bx[2] = xmax; bx[3] = ymax;
  return 1;
}

/* rust/sedona-spatial-join-gpu/src/join_provider.rs:108-140 (points: nearest
 * cast) and rust/sedona-spatial-join/src/utils/bounds.rs:49-63 (others). */
void ora_rects(const ora_store* s, int point_nearest, float* out) {
  for (int64_t g = 0; g < s->n_geoms; g++) {
    double b[4];
    float* o = out + 4 * g;
    if (!geom_bounds(s, g, b)) {
      o[0] = o[1] = o[2] = o[3] = NAN;
      continue;
    }
    if (point_nearest && s->type[g] == ORA_POINT) {
      o[0] = o[2] = (float)b[0];
      o[1] = o[3] = (float)b[1];
    } else {
      ora_f64_interval_to_f32(b[0], b[2], &o[0], &o[2]);
      ora_f64_interval_to_f32(b[1], b[3], &o[1], &o[3]);
    }
  }
}

/* rust/sedona-spatial-join/src/operand_evaluator.rs:605-607 with
 * rust/sedona-geometry/src/interval.rs:269-275. */
void ora_expand_rects(float* rects, int64_t n, double d) {
  if (isnan(d) || d < 0) return;
  for (int64_t i = 0; i < n; i++) {
    float* o = rects + 4 * i;
    if (isnan(o[0]) || o[0] > o[2] || o[1] > o[3]) continue;
    double xl = (double)o[0] - d, xh = (double)o[2] + d;
    double yl = (double)o[1] - d, yh = (double)o[3] + d;
    ora_f64_interval_to_f32(xl, xh, &o[0], &o[2]);
    ora_f64_interval_to_f32(yl, yh, &o[1], &o[3]);
  }
}

/* ------------------------------------------------------------------ */
/* Filter: uniform grid over build rects                              */
/* ------------------------------------------------------------------ */

static int rect_valid(const float* r) {
  return !(isnan(r[0]) || isnan(r[1]) || isnan(r[2]) || isnan(r[3])) && r[0] <= r[2] && r[1] <= r[3];
}

static int rect_hit(const float* a, const float* b) {
  return !(b[0] > a[2] || b[2] < a[0] || b[1] > a[3] || b[3] < a[1]);
}

typedef struct {
  double x0, y0, cw, ch;
  int64_t nx, ny;
  int64_t* cell_off;
  uint32_t* items;
} grid_t;

static int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

static void grid_cells(const grid_t* G, const float* r, int64_t* c) {
  c[0] = clampi((int64_t)floor(((double)r[0] - G->x0) / G->cw), 0, G->nx - 1);
  c[1] = clampi((int64_t)floor(((double)r[1] - G->y0) / G->ch), 0, G->ny - 1);
  c[2] = clampi((int64_t)floor(((double)r[2] - G->x0) / G->cw), 0, G->nx - 1);
  c[3] = clampi((int64_t)floor(((double)r[3] - G->y0) / G->ch), 0, G->ny - 1);
}

static void grid_build(grid_t* G, const float* rects, int64_t nb) {
  double xmin = INFINITY, ymin = INFINITY, xmax = -INFINITY, ymax = -INFINITY, area = 0;
  int64_t nvalid = 0;
  for (int64_t i = 0; i < nb; i++) {
    const float* r = rects + 4 * i;
    if (!rect_valid(r)) continue;
    nvalid++;
    if (r[0] < xmin) xmin = r[0];
    if (r[1] < ymin) ymin = r[1];
    if (r[2] > xmax) xmax = r[2];
    if (r[3] > ymax) ymax = r[3];
    area += ((double)r[2] - r[0]) * ((double)r[3] - r[1]);
  }
  if (nvalid == 0) { xmin = ymin = 0; xmax = ymax = 1; }
  double W = xmax - xmin, H = ymax - ymin;
  if (!(W > 0)) W = 1;
  if (!(H > 0)) H = 1;
  /* cell edge ~ mean rect edge, at most ~4 cells per valid rect */
  double mean_edge = nvalid ? sqrt(area / (double)nvalid) : 1;
  double cells_target = (double)(nvalid > 0 ? nvalid : 1) * 2.0;
  double edge = sqrt(W * H / cells_target);
  if (mean_edge > edge) edge = mean_edge;
  if (!(edge > 0)) edge = 1;
  G->nx = (int64_t)ceil(W / edge);
  G->ny = (int64_t)ceil(H / edge);
  if (G->nx < 1) G->nx = 1;
  if (G->ny < 1) G->ny = 1;
  if (G->nx > 8192) G->nx = 8192;
  if (G->ny > 8192) G->ny = 8192;
  G->x0 = xmin; G->y0 = ymin;
  G->cw = W / (double)G->nx; G->ch = H / (double)G->ny;
  if (!(G->cw > 0)) G->cw = 1;
  if (!(G->ch > 0)) G->ch = 1;
  int64_t ncell = G->nx * G->ny;
  G->cell_off = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
  for (int64_t i = 0; i < nb; i++) {
    const float* r = rects + 4 * i;
    if (!rect_valid(r)) continue;
    int64_t c[4];
    grid_cells(G, r, c);
    for (int64_t y = c[1]; y <= c[3]; y++)
      for (int64_t x = c[0]; x <= c[2]; x++) G->cell_off[y * G->nx + x + 1]++;
  }
  for (int64_t i = 0; i < ncell; i++) G->cell_off[i + 1] += G->cell_off[i];
  G->items = (uint32_t*)malloc((size_t)(G->cell_off[ncell] > 0 ? G->cell_off[ncell] : 1) * 4);
  int64_t* fill = (int64_t*)malloc((size_t)ncell * sizeof(int64_t));
  memcpy(fill, G->cell_off, (size_t)ncell * sizeof(int64_t));
  for (int64_t i = 0; i < nb; i++) {
    const float* r = rects + 4 * i;
    if (!rect_valid(r)) continue;
    int64_t c[4];
    grid_cells(G, r, c);
    for (int64_t y = c[1]; y <= c[3]; y++)
      for (int64_t x = c[0]; x <= c[2]; x++) G->items[fill[y * G->nx + x]++] = (uint32_t)i;
  }
  free(fill);
}

static void grid_free(grid_t* G) { free(G->cell_off); free(G->items); }

typedef struct { uint32_t* v; int64_t n, cap; } vecu32;
static void vu_push(vecu32* a, uint32_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? a->cap * 2 : 256;
    a->v = (uint32_t*)realloc(a->v, (size_t)a->cap * 4);
  }
  a->v[a->n++] = x;
}

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : (x > y);
}

/* Candidates of one probe rect, sorted by build id, deduplicated. */
static void grid_query(const grid_t* G, const float* brects, const float* pr, vecu32* out) {
  out->n = 0;
  if (!rect_valid(pr)) return;
  int64_t c[4];
  grid_cells(G, pr, c);
  for (int64_t y = c[1]; y <= c[3]; y++)
    for (int64_t x = c[0]; x <= c[2]; x++) {
      int64_t cell = y * G->nx + x;
      for (int64_t k = G->cell_off[cell]; k < G->cell_off[cell + 1]; k++) {
        uint32_t b = G->items[k];
        if (rect_hit(brects + 4 * (int64_t)b, pr)) vu_push(out, b);
      }
    }
  if (out->n > 1) {
    qsort(out->v, (size_t)out->n, 4, cmp_u32);
    int64_t w = 1;
    for (int64_t i = 1; i < out->n; i++)
      if (out->v[i] != out->v[w - 1]) out->v[w++] = out->v[i];
    out->n = w;
  }
}

/* ------------------------------------------------------------------ */
/* TG relation semantics                                               */
/* ------------------------------------------------------------------ */

enum { RC_OUT = 0, RC_IN = 1, RC_ON = 2 };

/* tg.c:816-823: x86 minsd/maxsd select. */
static inline double fmin0(double x, double y) { return x < y ? x : y; }
static inline double fmax0(double x, double y) { return x > y ? x : y; }
/* tg.c:853-859 */
static inline int feq(double x, double y) { return !((x < y) | (x > y)); }
static inline int eq_zero(double x) { return feq(x, 0); }

/* tg.c:861-901: cross-product test with a one-ulp correction of each
 * product in the direction in which its division check says it lost
 * precision. */
int ora_collinear(double x1, double y1, double x2, double y2, double x3, double y3) {
  int x1x2 = feq(x1, x2), x1x3 = feq(x1, x3), x2x3 = feq(x2, x3);
  int y1y2 = feq(y1, y2), y1y3 = feq(y1, y3), y2y3 = feq(y2, y3);
  if (x1x2) return x1x3;
  if (y1y2) return y1y3;
  if ((x1x2 & y1y2) | (x1x3 & y1y3) | (x2x3 & y2y3)) return 1;
  double cx1 = x3 - x1, cy1 = y3 - y1;
  double cx2 = x2 - x1, cy2 = y2 - y1;
  double s1 = cx1 * cy2;
  double s2 = cy1 * cx2;
  double s3 = (s1 / cy2) - cx1;
  double s4 = (s2 / cx2) - cy1;
  if (s3 < 0) s1 = nextafter(s1, -INFINITY);
  else if (s3 > 0) s1 = nextafter(s1, INFINITY);
  if (s4 < 0) s2 = nextafter(s2, -INFINITY);
  else if (s4 > 0) s2 = nextafter(s2, INFINITY);
  return eq_zero(s1 - s2);
}

/* tg.c:1199-1275: classification of a +x ray from p against segment ab. */
int ora_raycast(double ax, double ay, double bx, double by, double px, double py) {
  double rminx = fmin0(ax, bx), rminy = fmin0(ay, by);
  double rmaxx = fmax0(ax, bx), rmaxy = fmax0(ay, by);
  if (py < rminy || py > rmaxy) return RC_OUT;
  if (px < rminx) {
    if (py != rminy && py != rmaxy) return RC_IN;
  } else if (px > rmaxx) {
    if (rminy != rmaxy && rminx != rmaxx) return RC_OUT;
  }
  if (by < ay) {
    double t;
    t = ax; ax = bx; bx = t;
    t = ay; ay = by; by = t;
  }
  if ((feq(px, ax) && feq(py, ay)) || (feq(px, bx) && feq(py, by))) return RC_ON;
  if (ay == by) {
    if (ax == bx) return RC_OUT;
    if (py == by) {
      if (!(px < rminx || px > rmaxx)) return RC_ON;
    }
  }
  if (ax == bx && px == bx) {
    if (py >= ay && py <= by) return RC_ON;
  }
  if (ora_collinear(ax, ay, bx, by, px, py)) {
    if (px < rminx) {
      if (rminy == rmaxy) return RC_OUT;
    } else if (px > rmaxx) {
      return RC_OUT;
    }
    return RC_ON;
  }
  if (py == ay || py == by) py = nextafter(py, INFINITY);
  if (py < ay || py > by) return RC_OUT;
  if (ax > bx) {
    if (px >= ax) return RC_OUT;
    if (px <= bx) return RC_IN;
  } else {
    if (px >= bx) return RC_OUT;
    if (px <= ax) return RC_IN;
  }
  if ((py - ay) / (px - ax) >= (by - ay) / (bx - ax)) return RC_IN;
  return RC_OUT;
}

typedef struct {
  const double* xy; /* ring points */
  int64_t npts;
} ring_t;

/* tg.c:2126-2134 (closed series). */
static int64_t ring_nsegs(const ring_t* r) {
  if (r->npts < 3) return 0;
  const double* p = r->xy;
  const double* q = r->xy + 2 * (r->npts - 1);
  if (feq(q[0], p[0]) && feq(q[1], p[1])) return r->npts - 1;
  return r->npts;
}

static void ring_seg(const ring_t* r, int64_t i, double* a) {
  int64_t j = i + 1 == r->npts ? 0 : i + 1; /* points[npoints] == points[0] */
  a[0] = r->xy[2 * i]; a[1] = r->xy[2 * i + 1];
  a[2] = r->xy[2 * j]; a[3] = r->xy[2 * j + 1];
}

static void ring_rect(const ring_t* r, double* rc) {
  rc[0] = rc[1] = INFINITY;
  rc[2] = rc[3] = -INFINITY;
  for (int64_t i = 0; i < r->npts; i++) {
    double x = r->xy[2 * i], y = r->xy[2 * i + 1];
    rc[0] = fmin0(rc[0], x); rc[1] = fmin0(rc[1], y);
    rc[2] = fmax0(rc[2], x); rc[3] = fmax0(rc[3], y);
  }
}

static int ring_empty(const ring_t* r) { return r->npts < 3; } /* tg.c:3113-3118 */

/* tg.c:2809-2822 with standard_pip (2741-2760) and pip_eval_seg_slow
 * (2669-2687): any ON segment decides allow_on_edge, otherwise the parity
 * of IN segments decides. */
static int ring_contains_point(const ring_t* r, double px, double py, int allow_on_edge) {
  double rc[4];
  ring_rect(r, rc);
  if (px < rc[0] || py < rc[1] || px > rc[2] || py > rc[3]) return 0;
  int in = 0, on = 0;
  int64_t ns = ring_nsegs(r);
  for (int64_t i = 0; i < ns; i++) {
    double s[4];
    ring_seg(r, i, s);
    int rcst = ora_raycast(s[0], s[1], s[2], s[3], px, py);
    if (rcst == RC_ON) on = 1;
    else if (rcst == RC_IN) in ^= 1;
  }
  return on ? allow_on_edge : in;
}

typedef struct {
  const ora_store* s;
  int64_t part;
} poly_t;

static ring_t poly_ring(const poly_t* p, int64_t k) {
  int64_t r = p->s->ring_off[p->part] + k;
  ring_t o;
  o.xy = p->s->xy + 2 * p->s->pt_off[r];
  o.npts = p->s->pt_off[r + 1] - p->s->pt_off[r];
  return o;
}

static int64_t poly_nrings(const poly_t* p) {
  return p->s->ring_off[p->part + 1] - p->s->ring_off[p->part];
}

/* tg.c:3909-3932 */
static int poly_contains_point(const poly_t* p, double px, double py, int allow_on_edge) {
  int64_t nr = poly_nrings(p);
  if (nr == 0) return 0;
  ring_t ext = poly_ring(p, 0);
  if (nr == 1) return ring_contains_point(&ext, px, py, allow_on_edge);
  if (ring_empty(&ext)) return 0;
  if (!ring_contains_point(&ext, px, py, allow_on_edge)) return 0;
  for (int64_t k = 1; k < nr; k++) {
    ring_t h = poly_ring(p, k);
    if (ring_contains_point(&h, px, py, !allow_on_edge)) return 0;
  }
  return 1;
}

/* tg.c:1186-1197 point_on_segment: rect cover + collinear. */
static int segment_covers_point(const double* s, double px, double py) {
  double minx = fmin0(s[0], s[2]), miny = fmin0(s[1], s[3]);
  double maxx = fmax0(s[0], s[2]), maxy = fmax0(s[1], s[3]);
  if (px < minx || py < miny || px > maxx || py > maxy) return 0;
  return ora_collinear(s[0], s[1], s[2], s[3], px, py);
}

/* tg.c:1392-1443 */
static int seg_intersects_seg(const double* A, const double* B) {
  double ax = A[0], ay = A[1], bx = A[2], by = A[3];
  double cx = B[0], cy = B[1], dx = B[2], dy = B[3];
  double r1[4] = {fmin0(ax, bx), fmin0(ay, by), fmax0(ax, bx), fmax0(ay, by)};
  double r2[4] = {fmin0(cx, dx), fmin0(cy, dy), fmax0(cx, dx), fmax0(cy, dy)};
  if (r2[0] > r1[2] || r2[2] < r1[0] || r2[1] > r1[3] || r2[3] < r1[1]) return 0;
  if ((feq(ax, cx) && feq(ay, cy)) || (feq(ax, dx) && feq(ay, dy)) ||
      (feq(bx, cx) && feq(by, cy)) || (feq(bx, dx) && feq(by, dy)))
    return 1;
  double cmpx = cx - ax, cmpy = cy - ay;
  double rx = bx - ax, ry = by - ay;
  double cmpxr = cmpx * ry - cmpy * rx;
  if (eq_zero(cmpxr)) {
    if (!(((cx - ax <= 0) != (cx - bx <= 0))) || ((cy - ay <= 0) != (cy - by <= 0))) {
      return segment_covers_point(A, cx, cy) || segment_covers_point(A, dx, dy) ||
             segment_covers_point(B, ax, ay) || segment_covers_point(B, bx, by);
    }
    return 1;
  }
  double sx = dx - cx, sy = dy - cy;
  double rxs = rx * sy - ry * sx;
  if (eq_zero(rxs)) return 0;
  double cmpxs = cmpx * sy - cmpy * sx;
  double rxsr = 1 / rxs;
  double t = cmpxs * rxsr;
  double u = cmpxr * rxsr;
  if (!((t >= 0) && (t <= 1) && (u >= 0) && (u <= 1))) return 0;
  return 1;
}

int ora_segment_intersects(double ax, double ay, double bx, double by, double cx, double cy,
                           double dx, double dy) {
  double A[4] = {ax, ay, bx, by}, B[4] = {cx, cy, dx, dy};
  return seg_intersects_seg(A, B);
}

/* tg.c:3089-3111 with intersegiter (3045-3087) for allow_on_edge = true:
 * true when an endpoint is covered or at least two ring segments cross. */
static int ring_intersects_segment(const ring_t* ring, const double* seg) {
  double rr[4];
  ring_rect(ring, rr);
  double sr[4] = {fmin0(seg[0], seg[2]), fmin0(seg[1], seg[3]), fmax0(seg[0], seg[2]), fmax0(seg[1], seg[3])};
  if (rr[0] > sr[2] || rr[2] < sr[0] || rr[1] > sr[3] || rr[3] < sr[1]) return 0;
  if (ring_contains_point(ring, seg[0], seg[1], 1) || ring_contains_point(ring, seg[2], seg[3], 1)) return 1;
  int count = 0;
  int64_t ns = ring_nsegs(ring);
  for (int64_t i = 0; i < ns && count < 2; i++) {
    double s[4];
    ring_seg(ring, i, s);
    if (seg_intersects_seg(seg, s)) count++;
  }
  return count >= 2;
}

static double rect_area(const double* r) { return (r[2] - r[0]) * (r[3] - r[1]); }

/* tg.c:3166-3190 (allow_on_edge = true). */
static int ring_intersects_ring(const ring_t* ring, const ring_t* other) {
  if (ring_empty(ring) || ring_empty(other)) return 0;
  double a[4], b[4];
  ring_rect(ring, a);
  ring_rect(other, b);
  if (b[0] > a[2] || b[2] < a[0] || b[1] > a[3] || b[3] < a[1]) return 0;
  if (rect_area(b) > rect_area(a)) {
    const ring_t* t = ring;
    ring = other;
    other = t;
  }
  int64_t ns = ring_nsegs(other);
  for (int64_t i = 0; i < ns; i++) {
    double s[4];
    ring_seg(other, i, s);
    if (ring_intersects_segment(ring, s)) return 1;
  }
  return 0;
}

/* tg.c:3119-3151 tg_ring_contains_ring(a, b, allow_on_edge = false), the
 * hole test of poly x poly intersects. */
static int ring_contains_ring_strict(const ring_t* a, const ring_t* b);

/* tg.c:4100-4136 */
static int poly_intersects_poly(const poly_t* p, const poly_t* o) {
  int64_t np = poly_nrings(p), no = poly_nrings(o);
  if (np == 0 || no == 0) return 0;
  ring_t pe = poly_ring(p, 0), oe = poly_ring(o, 0);
  if (np == 1 && no == 1) return ring_intersects_ring(&pe, &oe);
  if (ring_empty(&pe) || ring_empty(&oe)) return 0;
  if (!ring_intersects_ring(&oe, &pe)) return 0;
  for (int64_t k = 1; k < np; k++) {
    ring_t h = poly_ring(p, k);
    if (ring_contains_ring_strict(&h, &oe)) return 0;
  }
  for (int64_t k = 1; k < no; k++) {
    ring_t h = poly_ring(o, k);
    if (ring_contains_ring_strict(&h, &pe)) return 0;
  }
  return 1;
}

/* tg.c:1910-2120 (process_points): the convex flag TG keeps per ring.  Every
 * point takes part, the closing one included, with its two successors taken
 * modulo npoints; the first turn's sign sets the direction (a zero turn
 * counts as counter-clockwise), later zero turns are ignored, and the first
 * turn against the direction makes the ring concave. */
static int ring_convex_tg(const ring_t* r) {
  int64_t n = r->npts;
  if (n < 2) return 0;
  int dir = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t j = i + 1 < n ? i + 1 : i + 1 - n;
    int64_t k = i + 2 < n ? i + 2 : i + 2 - n;
    double ax = r->xy[2 * i], ay = r->xy[2 * i + 1];
    double bx = r->xy[2 * j], by = r->xy[2 * j + 1];
    double cx = r->xy[2 * k], cy = r->xy[2 * k + 1];
    double z = (bx - ax) * (cy - by) - (by - ay) * (cx - bx);
    if (dir == 0) {
      dir = z < 0 ? -1 : 1;
    } else if (z < 0) {
      if (dir == 1) return 0;
    } else if (z > 0) {
      if (dir == -1) return 0;
    }
  }
  return 1;
}

/* tg.c:2910-3016 tg_ring_contains_segment with allow_on_edge = false: the
 * ring's rect covers the segment's, both endpoints lie strictly inside (a
 * zero-length segment needs one), and, for a concave ring, no ring segment
 * meets it (contsegiter3, tg.c:2899-2907; the rect search of tg_ring_search
 * drops only segments that tg_segment_intersects_segment rejects anyway). */
static int ring_contains_segment_strict(const ring_t* r, const double* rr, const double* s, int convex) {
  double sr[4] = {fmin0(s[0], s[2]), fmin0(s[1], s[3]), fmax0(s[0], s[2]), fmax0(s[1], s[3])};
  if (sr[0] < rr[0] || sr[2] > rr[2] || sr[1] < rr[1] || sr[3] > rr[3]) return 0;
  if (!ring_contains_point(r, s[0], s[1], 0)) return 0;
  if (feq(s[2], s[0]) && feq(s[3], s[1])) return 1;
  if (!ring_contains_point(r, s[2], s[3], 0)) return 0;
  if (convex) return 1;
  int64_t ns = ring_nsegs(r);
  for (int64_t i = 0; i < ns; i++) {
    double t[4];
    ring_seg(r, i, t);
    if (seg_intersects_seg(s, t)) return 0;
  }
  return 1;
}

/* tg.c:3119-3151: a convex ring contains b when every point of b lies
 * strictly inside it, a concave one when every segment of b does
 * (ring_contains_segment_strict). */
static int ring_contains_ring_strict(const ring_t* a, const ring_t* b) {
  if (ring_empty(a) || ring_empty(b)) return 0;
  double ra[4], rb[4];
  ring_rect(a, ra);
  ring_rect(b, rb);
  if (rb[0] < ra[0] || rb[2] > ra[2] || rb[1] < ra[1] || rb[3] > ra[3]) return 0;
  if (ring_convex_tg(a)) {
    for (int64_t i = 0; i < b->npts; i++)
      if (!ring_contains_point(a, b->xy[2 * i], b->xy[2 * i + 1], 0)) return 0;
    return 1;
  }
  int64_t nb = ring_nsegs(b);
  for (int64_t i = 0; i < nb; i++) {
    double s[4];
    ring_seg(b, i, s);
    if (!ring_contains_segment_strict(a, ra, s, 0)) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------ */
/* TG ring / line relations with boundaries (polygon x polygon covers, */
/* contains, touches, equals; intersects with lines)                   */
/* ------------------------------------------------------------------ */

static int pteq2(double ax, double ay, double bx, double by) { return feq(ax, bx) && feq(ay, by); }

/* tg.c:1910-2120 (process_points): TG's clockwise flag, the shoelace sum
 * over every point (closing one included, successor modulo npoints) in
 * index order, clockwise when positive. */
static int ring_clockwise_tg(const ring_t* r) {
  int64_t n = r->npts;
  if (n < 2) return 0;
  double cwc = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t j = i + 1 < n ? i + 1 : i + 1 - n;
    double ax = r->xy[2 * i], ay = r->xy[2 * i + 1], bx = r->xy[2 * j], by = r->xy[2 * j + 1];
    cwc += (bx - ax) * (by + ay);
  }
  return cwc > 0;
}

/* tg.c:2669-2822: point in ring, with the index of the segment the point
 * is on (-1 when on none).  A segment the point is on sets the result to
 * allow_on_edge and its index; IN toggles count only before the first
 * such segment; segments are visited in index order, so the index is the
 * last one. */
typedef struct { int hit; int64_t idx; } ring_hit;

static ring_hit ring_contains_point_ix(const ring_t* r, double px, double py, int allow_on_edge) {
  ring_hit out = {0, -1};
  double rc[4];
  ring_rect(r, rc);
  if (px < rc[0] || py < rc[1] || px > rc[2] || py > rc[3]) return out;
  int in = 0;
  int64_t idx = -1;
  int64_t ns = ring_nsegs(r);
  for (int64_t i = 0; i < ns; i++) {
    double s[4];
    ring_seg(r, i, s);
    int rcst = ora_raycast(s[0], s[1], s[2], s[3], px, py);
    if (rcst == RC_ON) {
      in = allow_on_edge;
      idx = i;
    } else if (rcst == RC_IN && idx == -1) {
      in = !in;
    }
  }
  out.hit = in;
  out.idx = idx;
  return out;
}

static void seg_rect4(const double* s, double* r) {
  r[0] = fmin0(s[0], s[2]); r[1] = fmin0(s[1], s[3]);
  r[2] = fmax0(s[0], s[2]); r[3] = fmax0(s[1], s[3]);
}
static int rect_covers_rect4(const double* a, const double* b) {
  return !(b[0] < a[0] || b[2] > a[2] || b[1] < a[1] || b[3] > a[3]);
}
static int rect_isect4(const double* a, const double* b) {
  return !(b[0] > a[2] || b[2] < a[0] || b[1] > a[3] || b[3] < a[1]);
}

/* tg.c:2910-3016 tg_ring_contains_segment.  The rect searches of
 * tg_ring_search (tg.c:2468-2485) visit every segment whose rect meets the
 * query, and each iterator (tg.c:2847-2907) first asks
 * tg_segment_intersects_segment, which rejects disjoint rects: visiting
 * every segment decides the same. */
static int ring_contains_segment_tg(const ring_t* r, const double* s, int allow_on_edge) {
  double rr[4], sr[4];
  ring_rect(r, rr);
  seg_rect4(s, sr);
  if (!rect_covers_rect4(rr, sr)) return 0;
  ring_hit a = ring_contains_point_ix(r, s[0], s[1], allow_on_edge);
  if (!a.hit) return 0;
  if (pteq2(s[2], s[3], s[0], s[1])) return 1;
  ring_hit b = ring_contains_point_ix(r, s[2], s[3], allow_on_edge);
  if (!b.hit) return 0;
  if (ring_convex_tg(r)) return 1;
  int64_t ns = ring_nsegs(r);
  if (allow_on_edge) {
    int mode;
    if (a.idx != -1 && b.idx != -1) {
      if (b.idx == a.idx) return 1;
      double sa[4], sb[4];
      ring_seg(r, a.idx, sa);
      ring_seg(r, b.idx, sb);
      if (pteq2(sa[0], sa[1], s[0], s[1]) || pteq2(sa[2], sa[3], s[0], s[1]) || pteq2(sb[0], sb[1], s[0], s[1]) ||
          pteq2(sb[2], sb[3], s[0], s[1]) || pteq2(sa[0], sa[1], s[2], s[3]) || pteq2(sa[2], sa[3], s[2], s[3]) ||
          pteq2(sb[0], sb[1], s[2], s[3]) || pteq2(sb[2], sb[3], s[2], s[3]))
        return 1;
      if (b.idx < a.idx) {
        double t[4];
        memcpy(t, sa, sizeof t); memcpy(sa, sb, sizeof t); memcpy(sb, t, sizeof t);
      }
      double pts[10] = {sa[0], sa[1], sa[2], sa[3], sb[0], sb[1], sb[2], sb[3], sa[0], sa[1]};
      double cwc = 0.0;
      for (int i = 0; i < 4; i++) cwc += (pts[2 * i + 2] - pts[2 * i]) * (pts[2 * i + 3] + pts[2 * i + 1]);
      if ((cwc > 0) != ring_clockwise_tg(r)) return 0;
      mode = 1; /* contsegiter1 */
    } else if (a.idx != -1) {
      mode = 4; /* contsegiter4 */
    } else if (b.idx != -1) {
      mode = 2; /* contsegiter2 */
    } else {
      mode = 5; /* contsegiter5 */
    }
    for (int64_t i = 0; i < ns; i++) {
      double t[4];
      ring_seg(r, i, t);
      if (!seg_intersects_seg(s, t)) continue;
      int hit;
      switch (mode) {
        case 1: hit = ora_raycast(t[0], t[1], t[2], t[3], s[0], s[1]) != RC_ON &&
                      ora_raycast(t[0], t[1], t[2], t[3], s[2], s[3]) != RC_ON; break;
        case 4: hit = ora_raycast(t[0], t[1], t[2], t[3], s[0], s[1]) != RC_ON; break;
        case 2: hit = ora_raycast(t[0], t[1], t[2], t[3], s[2], s[3]) != RC_ON; break;
        default: hit = ora_raycast(s[0], s[1], s[2], s[3], t[0], t[1]) != RC_ON &&
                       ora_raycast(s[0], s[1], s[2], s[3], t[2], t[3]) != RC_ON;
      }
      if (hit) return 0;
    }
    return 1;
  }
  for (int64_t i = 0; i < ns; i++) { /* contsegiter3 */
    double t[4];
    ring_seg(r, i, t);
    if (seg_intersects_seg(s, t)) return 0;
  }
  return 1;
}

/* tg.c:3119-3151 with allow_on_edge. */
static int ring_contains_ring_tg(const ring_t* a, const ring_t* b, int allow_on_edge) {
  if (ring_empty(a) || ring_empty(b)) return 0;
  double ra[4], rb[4];
  ring_rect(a, ra);
  ring_rect(b, rb);
  if (!rect_covers_rect4(ra, rb)) return 0;
  if (ring_convex_tg(a)) {
    for (int64_t i = 0; i < b->npts; i++)
      if (!ring_contains_point_ix(a, b->xy[2 * i], b->xy[2 * i + 1], allow_on_edge).hit) return 0;
    return 1;
  }
  int64_t nb = ring_nsegs(b);
  for (int64_t i = 0; i < nb; i++) {
    double s[4];
    ring_seg(b, i, s);
    if (!ring_contains_segment_tg(a, s, allow_on_edge)) return 0;
  }
  return 1;
}

/* tg.c:3019-3111 tg_ring_intersects_segment with intersegiter: segments
 * visited in index order (tg_ring_search), the count reset by a collinear
 * overlap and the visit stopped at two crossings. */
static int ring_intersects_segment_tg(const ring_t* r, const double* s, int allow_on_edge) {
  double rr[4], sr[4];
  ring_rect(r, rr);
  seg_rect4(s, sr);
  if (!rect_isect4(sr, rr)) return 0;
  if (ring_contains_point_ix(r, s[0], s[1], allow_on_edge).hit ||
      ring_contains_point_ix(r, s[2], s[3], allow_on_edge).hit)
    return 1;
  int count = 0, a_on = 0, b_on = 0;
  int64_t ns = ring_nsegs(r);
  for (int64_t i = 0; i < ns; i++) {
    double t[4], tr[4];
    ring_seg(r, i, t);
    seg_rect4(t, tr);
    if (!rect_isect4(sr, tr)) continue;
    if (!seg_intersects_seg(s, t)) continue;
    if (allow_on_edge) {
      count++;
      if (count >= 2) break;
      continue;
    }
    int ccol = ora_collinear(s[0], s[1], s[2], s[3], t[0], t[1]);
    int dcol = ora_collinear(s[0], s[1], s[2], s[3], t[2], t[3]);
    if (ccol && dcol) {
      count = 0;
    } else {
      if (!a_on && (pteq2(s[0], s[1], t[0], t[1]) || pteq2(s[0], s[1], t[2], t[3]))) {
        a_on = 1;
        continue;
      }
      if (!b_on && (pteq2(s[2], s[3], t[0], t[1]) || pteq2(s[2], s[3], t[2], t[3]))) {
        b_on = 1;
        continue;
      }
      count++;
    }
    if (count >= 2) break;
  }
  return count >= 2;
}

/* tg.c:3166-3190 with allow_on_edge. */
static int ring_intersects_ring_tg(const ring_t* ring, const ring_t* other, int allow_on_edge) {
  if (ring_empty(ring) || ring_empty(other)) return 0;
  double a[4], b[4];
  ring_rect(ring, a);
  ring_rect(other, b);
  if (!rect_isect4(a, b)) return 0;
  if (rect_area(b) > rect_area(a)) {
    const ring_t* t = ring;
    ring = other;
    other = t;
  }
  int64_t ns = ring_nsegs(other);
  for (int64_t i = 0; i < ns; i++) {
    double s[4];
    ring_seg(other, i, s);
    if (ring_intersects_segment_tg(ring, s, allow_on_edge)) return 1;
  }
  return 0;
}

/* tg.c:3436-3443 tg_line_covers_point over a ring's segments. */
static int ring_boundary_covers_point(const ring_t* r, double px, double py) {
  int64_t ns = ring_nsegs(r);
  for (int64_t i = 0; i < ns; i++) {
    double s[4];
    ring_seg(r, i, s);
    if (segment_covers_point(s, px, py)) return 1;
  }
  return 0;
}

/* tg.c:4040-4092 tg_poly_covers_poly (tg_poly_contains_poly is the same,
 * tg.c:4094-4096). */
static int poly_covers_poly_tg(const poly_t* a, const poly_t* b) {
  int64_t na = poly_nrings(a), nb = poly_nrings(b);
  if (na == 0 || nb == 0) return 0;
  ring_t ae = poly_ring(a, 0), be = poly_ring(b, 0);
  if (ring_empty(&ae) || ring_empty(&be)) return 0;
  if (!ring_contains_ring_tg(&ae, &be, 1)) return 0;
  for (int64_t i = 1; i < na; i++) {
    ring_t h = poly_ring(a, i);
    if (ring_contains_ring_tg(&h, &be, 1) || ring_intersects_ring_tg(&h, &be, 0)) {
      int covers = 0;
      for (int64_t j = 1; j < nb; j++) {
        ring_t bh = poly_ring(b, j);
        if (ring_contains_ring_tg(&bh, &h, 1)) {
          covers = 1;
          break;
        }
      }
      if (!covers) return 0;
    }
  }
  return 1;
}

/* tg.c:4184-4249 tg_poly_touches_poly. */
static int poly_touches_poly_tg(const poly_t* a, const poly_t* b) {
  if (poly_nrings(a) == 0 || poly_nrings(b) == 0) return 0;
  ring_t aext = poly_ring(a, 0), bext = poly_ring(b, 0);
  double ra[4], rb[4];
  ring_rect(&aext, ra);
  ring_rect(&bext, rb);
  if (!rect_isect4(ra, rb)) return 0;
  for (int ii = 0; ii < 2; ii++) {
    const poly_t* poly = ii == 0 ? a : b;
    const ring_t* ring = ii == 0 ? &bext : &aext;
    for (int64_t i = 1; i < poly_nrings(poly); i++) {
      ring_t hole = poly_ring(poly, i);
      if (ring_contains_ring_tg(&hole, ring, 1)) {
        for (int64_t j = 0; j < ring->npts; j++)
          if (ring_boundary_covers_point(&hole, ring->xy[2 * j], ring->xy[2 * j + 1])) return 1;
        return 0;
      }
    }
  }
  int64_t ansegs = ring_nsegs(&aext), bnsegs = ring_nsegs(&bext);
  int64_t atouches = 0, btouches = 0;
  for (int ii = 0; ii < 2; ii++) {
    const ring_t* ra_ = ii == 0 ? &aext : &bext;
    const ring_t* rb_ = ii == 0 ? &bext : &aext;
    int64_t nsegs = ii == 0 ? bnsegs : ansegs;
    int64_t touches = 0;
    for (int64_t i = 0; i < nsegs; i++) {
      double s[4];
      ring_seg(rb_, i, s);
      int isects0 = ring_intersects_segment_tg(ra_, s, 1);
      int isects1 = ring_intersects_segment_tg(ra_, s, 0);
      if (isects0 && !isects1) touches++;
      else if (isects0 || isects1) return 0;
    }
    if (ii == 0) btouches = touches;
    else atouches = touches;
  }
  if (atouches > 0 || btouches > 0) return !(atouches == ansegs && btouches == bnsegs);
  return 0;
}

/* A (multi)polygon row's part q is a polygon; an empty part (no exterior,
 * or an exterior of fewer than 3 points) is TG's empty polygon. */
static int part_empty(const ora_store* s, int64_t q) {
  poly_t p = {s, q};
  if (poly_nrings(&p) == 0) return 1;
  ring_t e = poly_ring(&p, 0);
  return ring_empty(&e);
}

/* tg_geom_covers(A, B) for (multi)polygon rows (tg.c:5925-6045):
 * polygon A covers B when it covers B's polygon, or every part of a
 * multipolygon B (none empty, at least one); a multipolygon A covers B
 * when every non-empty part of B (at least one) is covered by some part of
 * A. */
static int polyish_covers_polyish(const ora_store* A, int64_t a, const ora_store* B, int64_t b) {
  int64_t a0 = A->part_off[a], a1 = A->part_off[a + 1], b0 = B->part_off[b], b1 = B->part_off[b + 1];
  if (a0 == a1 || b0 == b1) return 0;
  if (A->type[a] == ORA_POLYGON) {
    poly_t pa = {A, a0};
    for (int64_t q = b0; q < b1; q++) {
      poly_t pb = {B, q};
      if (!poly_covers_poly_tg(&pa, &pb)) return 0;
    }
    return 1;
  }
  int any = 0;
  for (int64_t q = b0; q < b1; q++) {
    if (part_empty(B, q)) continue;
    poly_t pb = {B, q};
    int found = 0;
    for (int64_t r = a0; r < a1 && !found; r++) {
      poly_t pa = {A, r};
      found = poly_covers_poly_tg(&pa, &pb);
    }
    if (!found) return 0;
    any = 1;
  }
  return any;
}

/* tg_geom_contains(A, B) for (multi)polygon rows (tg.c:6060-6106,
 * 6418-6520): a polygon A contains exactly what it covers; a multipolygon A
 * must cover B and some non-empty part of B must meet some part of A
 * without merely touching it. */
static int polyish_contains_polyish(const ora_store* A, int64_t a, const ora_store* B, int64_t b) {
  if (A->type[a] == ORA_POLYGON) return polyish_covers_polyish(A, a, B, b);
  if (!polyish_covers_polyish(A, a, B, b)) return 0;
  for (int64_t q = B->part_off[b]; q < B->part_off[b + 1]; q++) {
    if (part_empty(B, q)) continue;
    poly_t pb = {B, q};
    for (int64_t r = A->part_off[a]; r < A->part_off[a + 1]; r++) {
      if (part_empty(A, r)) continue;
      poly_t pa = {A, r};
      if (poly_intersects_poly(&pa, &pb) && !poly_touches_poly_tg(&pa, &pb)) return 1;
    }
  }
  return 0;
}

/* poly_touches_geom(part, B) (tg.c:6660-6700): B a polygon, or every part
 * of a multipolygon B touching or missing the polygon and one touching. */
static int part_touches_polyish(const poly_t* pa, const ora_store* B, int64_t b) {
  int64_t b0 = B->part_off[b], b1 = B->part_off[b + 1];
  if (b0 == b1) return 0;
  if (B->type[b] == ORA_POLYGON) {
    poly_t pb = {B, b0};
    return poly_touches_poly_tg(pa, &pb);
  }
  int touches = 0;
  for (int64_t q = b0; q < b1; q++) {
    poly_t pb = {B, q};
    if (poly_touches_poly_tg(pa, &pb)) touches = 1;
    else if (poly_intersects_poly(pa, &pb)) return 0;
  }
  return touches;
}

static int part_intersects_polyish(const poly_t* pa, const ora_store* B, int64_t b) {
  for (int64_t q = B->part_off[b]; q < B->part_off[b + 1]; q++) {
    poly_t pb = {B, q};
    if (poly_intersects_poly(pa, &pb)) return 1;
  }
  return 0;
}

/* tg_geom_touches(A, B) for (multi)polygon rows (tg.c:6660-6770). */
static int polyish_touches_polyish(const ora_store* A, int64_t a, const ora_store* B, int64_t b) {
  int64_t a0 = A->part_off[a], a1 = A->part_off[a + 1];
  if (a0 == a1) return 0;
  if (A->type[a] == ORA_POLYGON) {
    poly_t pa = {A, a0};
    return part_touches_polyish(&pa, B, b);
  }
  int touches = 0;
  for (int64_t r = a0; r < a1; r++) {
    poly_t pa = {A, r};
    if (part_touches_polyish(&pa, B, b)) touches = 1;
    else if (part_intersects_polyish(&pa, B, b)) return 0;
  }
  return touches;
}

/* Lines: a row's part holds one run of points; TG lines are open series
 * (tg.c:2126-2135: npoints - 1 segments, empty below two points). */
static int64_t line_nsegs(const ring_t* l) { return l->npts < 2 ? 0 : l->npts - 1; }
static void line_seg(const ring_t* l, int64_t i, double* a) {
  a[0] = l->xy[2 * i]; a[1] = l->xy[2 * i + 1];
  a[2] = l->xy[2 * i + 2]; a[3] = l->xy[2 * i + 3];
}
static void line_rect(const ring_t* l, double* rc) { ring_rect(l, rc); }

/* tg.c:3252-3275 tg_ring_intersects_line. */
static int ring_intersects_line_tg(const ring_t* r, const ring_t* l, int allow_on_edge) {
  if (ring_empty(r) || l->npts < 2) return 0;
  double rr[4], lr[4];
  ring_rect(r, rr);
  line_rect(l, lr);
  if (!rect_isect4(rr, lr)) return 0;
  for (int64_t i = 0; i < line_nsegs(l); i++) {
    double s[4];
    line_seg(l, i, s);
    if (ring_intersects_segment_tg(r, s, allow_on_edge)) return 1;
  }
  return 0;
}

/* tg.c:3194-3248 tg_ring_contains_line. */
static int ring_contains_line_tg(const ring_t* a, const ring_t* l, int allow_on_edge, int respect_boundaries) {
  if (ring_empty(a) || l->npts < 2) return 0;
  double ra[4], lr[4];
  ring_rect(a, ra);
  line_rect(l, lr);
  if (!rect_covers_rect4(ra, lr)) return 0;
  for (int64_t i = 0; i < line_nsegs(l); i++) {
    double s[4];
    line_seg(l, i, s);
    if (!allow_on_edge && respect_boundaries) {
      if (!ring_contains_segment_tg(a, s, 1)) return 0;
      if (!ring_intersects_segment_tg(a, s, 0)) return 0;
    } else if (!ring_contains_segment_tg(a, s, allow_on_edge)) {
      return 0;
    }
  }
  return 1;
}

/* tg.c:4016-4038 tg_poly_intersects_line. */
static int poly_intersects_line_tg(const poly_t* p, const ring_t* l) {
  int64_t nr = poly_nrings(p);
  if (nr == 0) return 0;
  ring_t e = poly_ring(p, 0);
  if (ring_empty(&e) || l->npts < 2) return 0;
  if (!ring_intersects_line_tg(&e, l, 1)) return 0;
  for (int64_t k = 1; k < nr; k++) {
    ring_t h = poly_ring(p, k);
    if (ring_contains_line_tg(&h, l, 0, 0)) return 0;
  }
  return 1;
}

/* tg.c:3595-3611 line_intersects_line (SI_INTERSECTS) through
 * tg_ring_ring_search (tg.c:2596-2647): both non-empty, rects meeting,
 * some pair of segments intersecting. */
static int line_intersects_line_tg(const ring_t* a, const ring_t* b) {
  if (a->npts < 2 || b->npts < 2) return 0;
  double ra[4], rb[4];
  line_rect(a, ra);
  line_rect(b, rb);
  if (!rect_isect4(ra, rb)) return 0;
  for (int64_t i = 0; i < line_nsegs(a); i++) {
    double s[4];
    line_seg(a, i, s);
    for (int64_t j = 0; j < line_nsegs(b); j++) {
      double t[4];
      line_seg(b, j, t);
      if (seg_intersects_seg(s, t)) return 1;
    }
  }
  return 0;
}

/* tg.c:3436-3462 tg_line_covers_point (= intersects). */
static int line_covers_point_tg(const ring_t* l, double px, double py) {
  for (int64_t i = 0; i < line_nsegs(l); i++) {
    double s[4];
    line_seg(l, i, s);
    if (segment_covers_point(s, px, py)) return 1;
  }
  return 0;
}

static ring_t part_line(const ora_store* s, int64_t q) {
  int64_t r = s->ring_off[q];
  ring_t o;
  o.xy = s->xy + 2 * s->pt_off[r];
  o.npts = s->pt_off[r + 1] - s->pt_off[r];
  return o;
}

static int is_lineish(int t) { return t == ORA_LINESTRING || t == ORA_MULTILINESTRING; }

/* tg_geom_intersects between rows where one side is a (multi)linestring
 * and the other a point, (multi)linestring or (multi)polygon
 * (tg.c:5544-5740: some pair of parts intersects). */
static int lineish_intersects(const ora_store* L, int64_t l, const ora_store* O, int64_t o) {
  int to = O->type[o];
  for (int64_t q = L->part_off[l]; q < L->part_off[l + 1]; q++) {
    ring_t ln = part_line(L, q);
    for (int64_t r = O->part_off[o]; r < O->part_off[o + 1]; r++) {
      int hit;
      if (to == ORA_POINT || to == ORA_MULTIPOINT) {
        ring_t pt = part_line(O, r);
        hit = line_covers_point_tg(&ln, pt.xy[0], pt.xy[1]);
      } else if (is_lineish(to)) {
        ring_t ol = part_line(O, r);
        hit = line_intersects_line_tg(&ln, &ol);
      } else {
        poly_t pp = {O, r};
        hit = poly_intersects_line_tg(&pp, &ln);
      }
      if (hit) return 1;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Exact orientation (Shewchuk-style expansion arithmetic)            */
/* ------------------------------------------------------------------ */

static inline void two_sum(double a, double b, double* x, double* y) {
  double s = a + b;
  double bv = s - a;
  double av = s - bv;
  *x = s;
  *y = (a - av) + (b - bv);
}

static inline void two_prod(double a, double b, double* x, double* y) {
  double p = a * b;
  *x = p;
  *y = fma(a, b, -p);
}

/* grow an expansion e (n terms, increasing magnitude, nonoverlapping) by b. */
static int grow(const double* e, int n, double b, double* h) {
  double q = b;
  int k = 0;
  for (int i = 0; i < n; i++) {
    double s, t;
    two_sum(q, e[i], &s, &t);
    if (t != 0) h[k++] = t;
    q = s;
  }
  if (q != 0 || k == 0) h[k++] = q;
  return k;
}

/* Sign of (ax-cx)(by-cy) - (ay-cy)(bx-cx), exactly. */
int ora_orient2d_sign(double ax, double ay, double bx, double by, double cx, double cy) {
  /* Shewchuk's first-stage filter (errboundA): the plain evaluation's sign
   * is exact whenever |det| exceeds this bound. */
  double detl = (ax - cx) * (by - cy), detr = (ay - cy) * (bx - cx);
  double det = detl - detr, detsum = fabs(detl) + fabs(detr);
  if (fabs(det) > 3.3306690738754716e-16 * detsum) return det > 0 ? 1 : -1;
  double acx, acx_e, bcy, bcy_e, acy, acy_e, bcx, bcx_e;
  two_sum(ax, -cx, &acx, &acx_e);
  two_sum(by, -cy, &bcy, &bcy_e);
  two_sum(ay, -cy, &acy, &acy_e);
  two_sum(bx, -cx, &bcx, &bcx_e);
  double u[2] = {acx_e, acx}, v[2] = {bcy_e, bcy}, w[2] = {acy_e, acy}, z[2] = {bcx_e, bcx};
  double e[64], tmp[64];
  int n = 0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      double p, pe;
      two_prod(u[i], v[j], &p, &pe);
      n = grow(e, n, p, tmp); memcpy(e, tmp, (size_t)n * 8);
      n = grow(e, n, pe, tmp); memcpy(e, tmp, (size_t)n * 8);
      two_prod(w[i], z[j], &p, &pe);
      n = grow(e, n, -p, tmp); memcpy(e, tmp, (size_t)n * 8);
      n = grow(e, n, -pe, tmp); memcpy(e, tmp, (size_t)n * 8);
    }
  for (int i = n - 1; i >= 0; i--) {
    if (e[i] > 0) return 1;
    if (e[i] < 0) return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Geo distance semantics                                              */
/* ------------------------------------------------------------------ */

/* utils.rs:127-166 */
static double coord_to_line(double px, double py, double sx, double sy, double ex, double ey) {
  if (sx == ex && sy == ey) return hypot(px - sx, py - sy);
  double dx = ex - sx, dy = ey - sy;
  double d2 = dx * dx + dy * dy;
  double r = ((px - sx) * dx + (py - sy) * dy) / d2;
  if (r <= 0) return hypot(px - sx, py - sy);
  if (r >= 1) return hypot(px - ex, py - ey);
  double s = ((sy - py) * dx - (sx - px) * dy) / d2;
  return fabs(s) * hypot(dx, dy);
}

/* intersects/mod.rs:118-138 */
static int in_between(double v, double b1, double b2) {
  if (b1 < b2) return v >= b1 && v <= b2;
  return v >= b2 && v <= b1;
}

/* intersects/line.rs:27-43 */
static int line_intersects_coord(double sx, double sy, double ex, double ey, double px, double py) {
  return ora_orient2d_sign(sx, sy, ex, ey, px, py) == 0 && in_between(px, sx, ex) && in_between(py, sy, ey);
}

/* Rust f64::min: NaN-ignoring minimum. */
static double rmin(double a, double b) {
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  return a < b ? a : b;
}

/* utils.rs:184-210 over a run of n points (open line). */
static double point_to_linestring(double px, double py, const double* xy, int64_t n) {
  if (n == 0) return 0;
  /* bbox check of line_string.rs:14-19 is implied by the per-line check */
  for (int64_t i = 0; i + 1 < n; i++)
    if (line_intersects_coord(xy[2 * i], xy[2 * i + 1], xy[2 * i + 2], xy[2 * i + 3], px, py)) return 0;
  if (n < 2) return 0;
  double m = coord_to_line(px, py, xy[0], xy[1], xy[2], xy[3]);
  for (int64_t i = 1; i + 1 < n; i++)
    m = rmin(m, coord_to_line(px, py, xy[2 * i], xy[2 * i + 1], xy[2 * i + 2], xy[2 * i + 3]));
  return m;
}

/* Exact point-in-ring status for geo's polygon-intersects-coord:
 * 2 = boundary, 1 = inside, 0 = outside (crossing parity, exact signs). */
static int exact_ring_position(const double* xy, int64_t n, double px, double py) {
  if (n < 2) return 0;
  int inside = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t j = i + 1 == n ? 0 : i + 1;
    double ax = xy[2 * i], ay = xy[2 * i + 1], bx = xy[2 * j], by = xy[2 * j + 1];
    if (line_intersects_coord(ax, ay, bx, by, px, py)) return 2;
    if ((ay > py) != (by > py)) {
      int o = ora_orient2d_sign(ax, ay, bx, by, px, py);
      /* edge upward (a below b): point left of edge -> crossing to the right */
      if ((by > ay) ? (o > 0) : (o < 0)) inside ^= 1;
    }
  }
  return inside;
}

/* utils.rs:231-271 */
static double point_to_polygon(const poly_t* p, double px, double py) {
  int64_t nr = poly_nrings(p);
  if (nr == 0) return 0;
  ring_t ext = poly_ring(p, 0);
  int st = exact_ring_position(ext.xy, ext.npts, px, py);
  if (st == 2) return 0;
  if (st == 1) {
    int in_hole = 0;
    for (int64_t k = 1; k < nr; k++) {
      ring_t h = poly_ring(p, k);
      int hs = exact_ring_position(h.xy, h.npts, px, py);
      if (hs == 2) return 0;
      if (hs == 1) in_hole = 1;
    }
    if (!in_hole) return 0;
  }
  double m = INFINITY;
  for (int64_t k = 0; k < nr; k++) {
    ring_t r = poly_ring(p, k);
    for (int64_t i = 0; i + 1 < r.npts; i++)
      m = rmin(m, coord_to_line(px, py, r.xy[2 * i], r.xy[2 * i + 1], r.xy[2 * i + 2], r.xy[2 * i + 3]));
  }
  return m;
}

static int geom_point(const ora_store* s, int64_t g, double* px, double* py) {
  if (s->type[g] != ORA_POINT) return 0;
  int64_t p0 = s->part_off[g];
  if (p0 == s->part_off[g + 1]) return 0;
  int64_t v = s->pt_off[s->ring_off[p0]];
  *px = s->xy[2 * v];
  *py = s->xy[2 * v + 1];
  return 1;
}

double ora_distance(const ora_store* B, int64_t b, const ora_store* P, int64_t p) {
  double px, py;
  if (!geom_point(P, p, &px, &py)) return NAN;
  int t = B->type[b];
  int64_t q0 = B->part_off[b], q1 = B->part_off[b + 1];
  if (q0 == q1) return NAN;
  double m = INFINITY;
  for (int64_t q = q0; q < q1; q++) {
    int64_t r0 = B->ring_off[q];
    const double* xy = B->xy + 2 * B->pt_off[r0];
    int64_t n = B->pt_off[r0 + 1] - B->pt_off[r0];
    double d;
    if (t == ORA_POINT || t == ORA_MULTIPOINT) d = hypot(xy[0] - px, xy[1] - py);
    else if (t == ORA_LINESTRING || t == ORA_MULTILINESTRING) d = point_to_linestring(px, py, xy, n);
    else {
      poly_t pp = {B, q};
      d = point_to_polygon(&pp, px, py);
    }
    m = rmin(m, d);
  }
  return m;
}

/* ------------------------------------------------------------------ */
/* Predicate dispatch                                                  */
/* ------------------------------------------------------------------ */

static int is_polyish(int t) { return t == ORA_POLYGON || t == ORA_MULTIPOLYGON; }

/* poly/multipoly (g in s) vs point: covers (allow_on_edge) or contains.
 * Multipolygon: tg.c:6185-6220 (any part decides). */
static int polyish_point(const ora_store* s, int64_t g, double px, double py, int allow_on_edge) {
  for (int64_t q = s->part_off[g]; q < s->part_off[g + 1]; q++) {
    poly_t p = {s, q};
    if (poly_contains_point(&p, px, py, allow_on_edge)) return 1;
  }
  return 0;
}

/* tg_point_covers_poly (tg.c:1150-1154): a degenerate polygon equal to the point. */
static int point_covers_polyish(const ora_store* s, int64_t g, double px, double py) {
  int64_t p0 = s->part_off[g], p1 = s->part_off[g + 1];
  if (p0 == p1) return 0;
  for (int64_t q = p0; q < p1; q++) {
    poly_t p = {s, q};
    if (poly_nrings(&p) == 0) return 0;
    ring_t e = poly_ring(&p, 0);
    if (ring_empty(&e)) return 0;
    double r[4];
    ring_rect(&e, r);
    if (!(feq(r[0], px) && feq(r[1], py) && feq(r[2], px) && feq(r[3], py))) return 0;
  }
  return 1;
}

/* tg.c:1171-1186 tg_point_touches_poly: the point lies on a segment of the
 * exterior ring or of a hole (tg_line_covers_point, tg.c:3417-3443). */
static int poly_boundary_covers_point(const poly_t* p, double px, double py) {
  for (int64_t k = 0; k < poly_nrings(p); k++) {
    ring_t r = poly_ring(p, k);
    for (int64_t i = 0; i < ring_nsegs(&r); i++) {
      double sg[4];
      ring_seg(&r, i, sg);
      if (segment_covers_point(sg, px, py)) return 1;
    }
  }
  return 0;
}

/* tg_geom_touches(poly or multipolygon, point).  Polygon: tg.c:6682-6690
 * -> tg_poly_touches_point (tg.c:3957-3959).  Multipolygon: tg.c:6735-6763,
 * some part touches and no part that does not touch intersects. */
static int polyish_touches_point(const ora_store* s, int64_t g, double px, double py) {
  if (s->type[g] == ORA_POLYGON) {
    poly_t p = {s, s->part_off[g]};
    return poly_boundary_covers_point(&p, px, py);
  }
  int touches = 0;
  for (int64_t q = s->part_off[g]; q < s->part_off[g + 1]; q++) {
    poly_t p = {s, q};
    if (poly_boundary_covers_point(&p, px, py)) touches = 1;
    else if (poly_contains_point(&p, px, py, 1)) return 0;
  }
  return touches;
}

/* tg_geom_touches(point, poly or multipolygon): tg.c:6578-6604, the point
 * touches some part (no test of the other parts' interiors). */
static int point_touches_polyish(const ora_store* s, int64_t g, double px, double py) {
  for (int64_t q = s->part_off[g]; q < s->part_off[g + 1]; q++) {
    poly_t p = {s, q};
    if (poly_boundary_covers_point(&p, px, py)) return 1;
  }
  return 0;
}

static int polyish_polyish_intersects(const ora_store* A, int64_t a, const ora_store* B, int64_t b) {
  for (int64_t q = A->part_off[a]; q < A->part_off[a + 1]; q++)
    for (int64_t r = B->part_off[b]; r < B->part_off[b + 1]; r++) {
      poly_t pa = {A, q}, pb = {B, r};
      if (poly_intersects_poly(&pa, &pb)) return 1;
    }
  return 0;
}

/* Returns 1/0, or -1 for an unsupported combination. */
static int eval_pair(const ora_store* B, int64_t b, const ora_store* P, int64_t p, int pred, double dist) {
  int tb = B->type[b], tp = P->type[p];
  if (tb == ORA_NULL || tp == ORA_NULL) return 0;
  if (B->part_off[b] == B->part_off[b + 1] || P->part_off[p] == P->part_off[p + 1]) return 0;
  if (pred == ORA_PRED_DWITHIN) {
    if (tp != ORA_POINT) return -1;
    double d = ora_distance(B, b, P, p);
    return d <= dist;
  }
  double bx, by, px, py;
  if (is_polyish(tb) && tp == ORA_POINT) {
    geom_point(P, p, &px, &py);
    switch (pred) {
      case ORA_PRED_INTERSECTS:
      case ORA_PRED_COVERS: return polyish_point(B, b, px, py, 1);
      case ORA_PRED_CONTAINS: return polyish_point(B, b, px, py, 0);
      case ORA_PRED_WITHIN: return 0;
      case ORA_PRED_COVEREDBY: return point_covers_polyish(B, b, px, py);
      case ORA_PRED_TOUCHES: return polyish_touches_point(B, b, px, py);
      case ORA_PRED_EQUALS: return 0; /* tg.c:15299-15301: within && contains; within(poly, point) is false */
      default: return -1;
    }
  }
  if (tb == ORA_POINT && is_polyish(tp)) {
    geom_point(B, b, &bx, &by);
    switch (pred) {
      case ORA_PRED_INTERSECTS:
      case ORA_PRED_COVEREDBY: return polyish_point(P, p, bx, by, 1);
      case ORA_PRED_WITHIN: return polyish_point(P, p, bx, by, 0);
      case ORA_PRED_CONTAINS: return 0;
      case ORA_PRED_COVERS: return point_covers_polyish(P, p, bx, by);
      case ORA_PRED_TOUCHES: return point_touches_polyish(P, p, bx, by);
      case ORA_PRED_EQUALS: return 0; /* tg.c:15299-15301: contains(point, poly) is false */
      default: return -1;
    }
  }
  if (tb == ORA_POINT && tp == ORA_POINT) {
    geom_point(B, b, &bx, &by);
    geom_point(P, p, &px, &py);
    switch (pred) {
      case ORA_PRED_INTERSECTS: case ORA_PRED_CONTAINS: case ORA_PRED_WITHIN:
      case ORA_PRED_COVERS: case ORA_PRED_COVEREDBY: case ORA_PRED_EQUALS:
        return feq(bx, px) && feq(by, py);
      case ORA_PRED_TOUCHES: return 0; /* tg.c:1102-1106: points have no boundary */
      default: return -1;
    }
  }
  if (is_polyish(tb) && is_polyish(tp)) {
    switch (pred) {
      case ORA_PRED_INTERSECTS: return polyish_polyish_intersects(B, b, P, p);
      case ORA_PRED_COVERS: return polyish_covers_polyish(B, b, P, p);
      case ORA_PRED_COVEREDBY: return polyish_covers_polyish(P, p, B, b);
      case ORA_PRED_CONTAINS: return polyish_contains_polyish(B, b, P, p);
      case ORA_PRED_WITHIN: return polyish_contains_polyish(P, p, B, b);
      case ORA_PRED_TOUCHES: return polyish_touches_polyish(B, b, P, p);
      /* tg.c:15299-15301: within(a, b) && contains(a, b) */
      case ORA_PRED_EQUALS: return polyish_contains_polyish(P, p, B, b) && polyish_contains_polyish(B, b, P, p);
      default: return -1;
    }
  }
  if (pred == ORA_PRED_INTERSECTS && (is_lineish(tb) || is_lineish(tp))) {
    if (is_lineish(tb)) return lineish_intersects(B, b, P, p);
    return lineish_intersects(P, p, B, b);
  }
  return -1;
}

int ora_refine(const ora_store* B, const ora_store* P, int pred, double dist, const uint32_t* bidx,
               const uint32_t* pidx, int64_t n, uint8_t* mask, int nthreads) {
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : err)
#endif
  for (int64_t i = 0; i < n; i++) {
    int r = eval_pair(B, bidx[i], P, pidx[i], pred, dist);
    if (r < 0) { err |= 1; r = 0; }
    mask[i] = (uint8_t)r;
  }
  (void)nthreads;
  return err ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* Filter and join drivers                                             */
/* ------------------------------------------------------------------ */

int64_t ora_filter(const float* brects, int64_t nb, const float* prects, int64_t np,
                   uint32_t* ob, uint32_t* op, int64_t cap, int nthreads) {
  grid_t G;
  grid_build(&G, brects, nb);
  int64_t* cnt = (int64_t*)calloc((size_t)np + 1, sizeof(int64_t));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    vecu32 tmp = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4096)
#endif
    for (int64_t p = 0; p < np; p++) {
      grid_query(&G, brects, prects + 4 * p, &tmp);
      cnt[p + 1] = tmp.n;
    }
    free(tmp.v);
  }
  for (int64_t p = 0; p < np; p++) cnt[p + 1] += cnt[p];
  int64_t total = cnt[np];
  if (ob && op && total <= cap) {
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
      vecu32 tmp = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4096)
#endif
      for (int64_t p = 0; p < np; p++) {
        grid_query(&G, brects, prects + 4 * p, &tmp);
        for (int64_t k = 0; k < tmp.n; k++) {
          ob[cnt[p] + k] = tmp.v[k];
          op[cnt[p] + k] = (uint32_t)p;
        }
      }
      free(tmp.v);
    }
  }
  (void)nthreads;
  free(cnt);
  grid_free(&G);
  return total;
}

int64_t ora_join(const ora_store* B, const ora_store* P, int pred, double dist, uint32_t* ob,
                 uint32_t* op, int64_t cap, int nthreads) {
  float* br = (float*)malloc((size_t)(B->n_geoms > 0 ? B->n_geoms : 1) * 16);
  float* pr = (float*)malloc((size_t)(P->n_geoms > 0 ? P->n_geoms : 1) * 16);
  ora_rects(B, 0, br);
  ora_rects(P, pred == ORA_PRED_DWITHIN ? 0 : 1, pr);
  if (pred == ORA_PRED_DWITHIN) ora_expand_rects(pr, P->n_geoms, dist);
  grid_t G;
  grid_build(&G, br, B->n_geoms);
  int64_t np = P->n_geoms;
  int64_t* cnt = (int64_t*)calloc((size_t)np + 1, sizeof(int64_t));
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  /* pass 1 counts, pass 2 writes: each probe's refined builds sorted by id */
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int64_t p = 0; p < np; p++) cnt[p + 1] += cnt[p];
      if (!ob || !op || cnt[np] > cap) break;
    }
#ifdef _OPENMP
#pragma omp parallel reduction(| : err)
#endif
    {
      vecu32 tmp = {0, 0, 0};
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1024)
#endif
      for (int64_t p = 0; p < np; p++) {
        grid_query(&G, br, pr + 4 * p, &tmp);
        int64_t w = pass ? cnt[p] : 0;
        int64_t c = 0;
        for (int64_t k = 0; k < tmp.n; k++) {
          int r = eval_pair(B, tmp.v[k], P, p, pred, dist);
          if (r < 0) { err |= 1; continue; }
          if (!r) continue;
          if (pass) { ob[w + c] = tmp.v[k]; op[w + c] = (uint32_t)p; }
          c++;
        }
        if (!pass) cnt[p + 1] = c;
      }
      free(tmp.v);
    }
  }
  int64_t total = cnt[np];
  free(cnt); free(br); free(pr);
  grid_free(&G);
  (void)nthreads;
  return err ? -1 : total;
}

/* ------------------------------------------------------------------ */
/* KNN (exact, ties by id)                                             */
/* ------------------------------------------------------------------ */

typedef struct { double d; uint32_t id; } knn_ent;

static int knn_less(const knn_ent* a, const knn_ent* b) {
  return a->d < b->d || (a->d == b->d && a->id < b->id);
}

/* Insert into a sorted top-k list (ascending). */
static void knn_insert(knn_ent* top, int* cnt, int k, knn_ent e) {
  if (*cnt == k && !knn_less(&e, &top[k - 1])) return;
  int i = *cnt < k ? (*cnt)++ : k - 1;
  while (i > 0 && knn_less(&e, &top[i - 1])) { top[i] = top[i - 1]; i--; }
  top[i] = e;
}

int ora_knn(const ora_store* B, const ora_store* P, int k, uint32_t* out_ids, double* out_d, int nthreads) {
  if (k <= 0) return -1;
  int64_t nb = B->n_geoms, np = P->n_geoms;
  double* bx = (double*)malloc((size_t)(nb > 0 ? nb : 1) * 16);
  uint8_t* ok = (uint8_t*)calloc((size_t)(nb > 0 ? nb : 1), 1);
  for (int64_t b = 0; b < nb; b++) {
    double x, y;
    if (geom_point(B, b, &x, &y)) { bx[2 * b] = x; bx[2 * b + 1] = y; ok[b] = 1; }
    else if (B->type[b] != ORA_NULL && B->part_off[b] != B->part_off[b + 1]) { free(bx); free(ok); return -1; }
  }
  /* grid over object points for pruning */
  double xmin = INFINITY, ymin = INFINITY, xmax = -INFINITY, ymax = -INFINITY;
  int64_t nvalid = 0;
  for (int64_t b = 0; b < nb; b++) if (ok[b]) {
    nvalid++;
    xmin = fmin(xmin, bx[2 * b]); xmax = fmax(xmax, bx[2 * b]);
    ymin = fmin(ymin, bx[2 * b + 1]); ymax = fmax(ymax, bx[2 * b + 1]);
  }
  if (!nvalid) { xmin = ymin = 0; xmax = ymax = 1; }
  double W = fmax(xmax - xmin, 1e-300), H = fmax(ymax - ymin, 1e-300);
  int64_t side = (int64_t)ceil(sqrt((double)(nvalid > 0 ? nvalid : 1) / 4.0));
  if (side < 1) side = 1;
  if (side > 4096) side = 4096;
  double cw = W / side, ch = H / side;
  int64_t ncell = side * side;
  int64_t* off = (int64_t*)calloc((size_t)ncell + 1, 8);
  uint32_t* items = (uint32_t*)malloc((size_t)(nvalid > 0 ? nvalid : 1) * 4);
#define CELLX(x) clampi((int64_t)floor(((x) - xmin) / cw), 0, side - 1)
#define CELLY(y) clampi((int64_t)floor(((y) - ymin) / ch), 0, side - 1)
  for (int64_t b = 0; b < nb; b++) if (ok[b]) off[CELLY(bx[2 * b + 1]) * side + CELLX(bx[2 * b]) + 1]++;
  for (int64_t c = 0; c < ncell; c++) off[c + 1] += off[c];
  int64_t* fill = (int64_t*)malloc((size_t)ncell * 8);
  memcpy(fill, off, (size_t)ncell * 8);
  for (int64_t b = 0; b < nb; b++) if (ok[b]) items[fill[CELLY(bx[2 * b + 1]) * side + CELLX(bx[2 * b])]++] = (uint32_t)b;
  free(fill);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    knn_ent* top = (knn_ent*)malloc((size_t)k * sizeof(knn_ent));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (int64_t p = 0; p < np; p++) {
      int cnt = 0;
      double qx, qy;
      if (geom_point(P, p, &qx, &qy) && nvalid > 0) {
        int64_t cx = CELLX(qx), cy = CELLY(qy);
        for (int64_t ring = 0;; ring++) {
          /* cells at Chebyshev distance `ring` from (cx, cy) */
          int any = 0;
          for (int64_t y = cy - ring; y <= cy + ring; y++) {
            if (y < 0 || y >= side) continue;
            for (int64_t x = cx - ring; x <= cx + ring; x++) {
              if (x < 0 || x >= side) continue;
              if (y != cy - ring && y != cy + ring && x != cx - ring && x != cx + ring) continue;
              any = 1;
              int64_t c = y * side + x;
              for (int64_t t = off[c]; t < off[c + 1]; t++) {
                uint32_t b = items[t];
                knn_ent e = {hypot(qx - bx[2 * b], qy - bx[2 * b + 1]), b};
                knn_insert(top, &cnt, k, e);
              }
            }
          }
          if (!any) break;
          if (cnt == k) {
            /* lower bound of any point beyond this ring */
            double gx = fmin(qx - (xmin + (cx - ring) * cw), (xmin + (cx + ring + 1) * cw) - qx);
            double gy = fmin(qy - (ymin + (cy - ring) * ch), (ymin + (cy + ring + 1) * ch) - qy);
            double g = fmin(gx, gy) * (1 - 1e-12);
            if (g > top[k - 1].d) break;
          }
        }
      }
      for (int i = 0; i < k; i++) {
        out_ids[p * k + i] = i < cnt ? top[i].id : 0xFFFFFFFFu;
        out_d[p * k + i] = i < cnt ? top[i].d : INFINITY;
      }
    }
    free(top);
  }
#undef CELLX
#undef CELLY
  (void)nthreads;
  free(off); free(items); free(bx); free(ok);
  return 0;
}

/* ------------------------------------------------------------------ */
/* KNN by brute force over any geometry (one side points)              */
/* ------------------------------------------------------------------ */

static int knn_ent_cmp(const void* a, const void* b) {
  const knn_ent* x = (const knn_ent*)a;
  const knn_ent* y = (const knn_ent*)b;
  if (x->d < y->d) return -1;
  if (x->d > y->d) return 1;
  return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

/* The reference's KNN (default_spatial_index.rs:306-502, ground truth of
 * spatial_join_integration.rs:1373-1414): per query every object's geo
 * distance (ora_distance; symmetric when the query is not a point and the
 * objects are), sorted by (distance, id), the first k kept, and with
 * tol >= 0 every further object within tol of the k-th distance
 * (include_tie_breakers, DISTANCE_TOLERANCE).  Rows of variable length:
 * out_off[q] .. out_off[q+1]; returns the total (outputs written when it
 * fits cap), or -1 when neither side is points. */
int64_t ora_knn_brute(const ora_store* B, const ora_store* P, int k, double tol, int64_t* out_off,
                      uint32_t* out_ids, double* out_d, int64_t cap, int nthreads) {
  int64_t nb = B->n_geoms, np = P->n_geoms;
  int bpts = 1, qpts = 1;
  for (int64_t b = 0; b < nb; b++) if (B->type[b] != ORA_NULL && B->type[b] != ORA_POINT) bpts = 0;
  for (int64_t q = 0; q < np; q++) if (P->type[q] != ORA_NULL && P->type[q] != ORA_POINT) qpts = 0;
  if (!bpts && !qpts) return -1;
  int64_t* cnt = (int64_t*)calloc((size_t)np + 1, 8);
  knn_ent** rows = (knn_ent**)calloc((size_t)(np > 0 ? np : 1), sizeof(knn_ent*));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int64_t q = 0; q < np; q++) {
    knn_ent* all = (knn_ent*)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(knn_ent));
    int64_t m = 0;
    double qx, qy;
    int qpoint = geom_point(P, q, &qx, &qy);
    for (int64_t b = 0; b < nb; b++) {
      double d;
      if (qpoint) d = ora_distance(B, b, P, q);
      else if (P->type[q] != ORA_NULL && P->part_off[q] != P->part_off[q + 1]) d = ora_distance(P, q, B, b);
      else d = NAN;
      if (d == d) { all[m].d = d; all[m].id = (uint32_t)b; m++; }
    }
    qsort(all, (size_t)m, sizeof(knn_ent), knn_ent_cmp);
    int64_t keep = m < k ? m : k;
    if (tol >= 0 && m >= k && k > 0) {
      double thr = all[k - 1].d + tol;
      while (keep < m && all[keep].d <= thr) keep++;
    }
    cnt[q + 1] = keep;
    rows[q] = all;
  }
  for (int64_t q = 0; q < np; q++) cnt[q + 1] += cnt[q];
  int64_t total = cnt[np];
  if (out_off) memcpy(out_off, cnt, (size_t)(np + 1) * 8);
  for (int64_t q = 0; q < np; q++) {
    if (total <= cap && out_ids && out_d)
      for (int64_t j = cnt[q]; j < cnt[q + 1]; j++) {
        out_ids[j] = rows[q][j - cnt[q]].id;
        out_d[j] = rows[q][j - cnt[q]].d;
      }
    free(rows[q]);
  }
  free(rows);
  free(cnt);
  (void)nthreads;
  return total;
}

/* ora_distance over n (build, probe) pairs. */
void ora_pair_distances(const ora_store* B, const ora_store* P, const uint32_t* bidx, const uint32_t* pidx,
                        int64_t n, double* out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4096)
#endif
  for (int64_t i = 0; i < n; i++) out[i] = ora_distance(B, bidx[i], P, pidx[i]);
  (void)nthreads;
}
