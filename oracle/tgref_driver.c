/*
 * tgref_driver.c -- drives the reference's own relation refine library
 * (c/sedona-tg/src/tg/tg.c, compiled from the reference tree by
 * oracle/Makefile into oracle/_ref/, never committed) over Arrow-style WKB
 * columns.
 *
 * TEST INFRASTRUCTURE ONLY: used to mint golden vectors, to pin the C
 * restatement in sjoin_oracle.c, and as the "reference" CPU baseline.
 *
 * Mirrors the reference CPU refine: rust/sedona-spatial-join/src/refine/tg.rs:
 * 200-268 (parse build geometry once, PrepareBuild; parse probe geometry;
 * Op::evaluate(build, probe)) with the predicate mapping of
 * c/sedona-tg/src/tg.rs:165-215.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "sjoin_oracle.h"
#include "tg.h"

typedef struct {
  int64_t n;
  struct tg_geom** g;
} tgcol;

static void* tgcol_parse(const int64_t* off, const uint8_t* data, const uint8_t* valid, int64_t n,
                         int ix, int nthreads) {
  tgcol* c = (tgcol*)calloc(1, sizeof(tgcol));
  c->n = n;
  c->g = (struct tg_geom**)calloc((size_t)(n > 0 ? n : 1), sizeof(void*));
  enum tg_index idx = ix == 0 ? TG_NONE : (ix == 1 ? TG_NATURAL : TG_YSTRIPES);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t i = 0; i < n; i++) {
    if (valid && !valid[i]) continue;
    struct tg_geom* g = tg_parse_wkb_ix(data + off[i], (size_t)(off[i + 1] - off[i]), idx);
    if (g && tg_geom_error(g)) {
      tg_geom_free(g);
      g = NULL;
    }
    c->g[i] = g;
  }
  (void)nthreads;
  return c;
}

void* tgref_parse(const int64_t* off, const uint8_t* data, const uint8_t* valid, int64_t n, int ix,
                  int nthreads) {
  return tgcol_parse(off, data, valid, n, ix, nthreads);
}

void tgref_free(void* p) {
  tgcol* c = (tgcol*)p;
  if (!c) return;
  for (int64_t i = 0; i < c->n; i++)
    if (c->g[i]) tg_geom_free(c->g[i]);
  free(c->g);
  free(c);
}

static int tg_eval(const struct tg_geom* b, const struct tg_geom* p, int pred) {
  switch (pred) {
    case ORA_PRED_EQUALS: return tg_geom_equals(b, p);
    case ORA_PRED_INTERSECTS: return tg_geom_intersects(b, p);
    case ORA_PRED_CONTAINS: return tg_geom_contains(b, p);
    case ORA_PRED_WITHIN: return tg_geom_within(b, p);
    case ORA_PRED_COVERS: return tg_geom_covers(b, p);
    case ORA_PRED_COVEREDBY: return tg_geom_coveredby(b, p);
    case ORA_PRED_TOUCHES: return tg_geom_touches(b, p);
    default: return -1;
  }
}

/* mask[i] = pred(build[bidx[i]], probe[pidx[i]]) evaluated by TG. */
int tgref_refine(void* build, void* probe, int pred, const uint32_t* bidx, const uint32_t* pidx,
                 int64_t n, uint8_t* mask, int nthreads) {
  tgcol* B = (tgcol*)build;
  tgcol* P = (tgcol*)probe;
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : err)
#endif
  for (int64_t i = 0; i < n; i++) {
    const struct tg_geom* b = B->g[bidx[i]];
    const struct tg_geom* p = P->g[pidx[i]];
    int r = (b && p) ? tg_eval(b, p, pred) : 0;
    if (r < 0) { err |= 1; r = 0; }
    mask[i] = (uint8_t)r;
  }
  (void)nthreads;
  return err ? -1 : 0;
}

/* Full reference-refined join: f32 rect filter (reference semantics, see
 * sjoin_oracle.c) + TG refine.  Pairs sorted by (probe, build). */
int64_t tgref_join(void* build, void* probe, const float* brects, const float* prects, int pred,
                   uint32_t* ob, uint32_t* op, int64_t cap, int nthreads) {
  tgcol* B = (tgcol*)build;
  tgcol* P = (tgcol*)probe;
  int64_t nc = ora_filter(brects, B->n, prects, P->n, NULL, NULL, 0, nthreads);
  uint32_t* cb = (uint32_t*)malloc((size_t)(nc > 0 ? nc : 1) * 4);
  uint32_t* cp = (uint32_t*)malloc((size_t)(nc > 0 ? nc : 1) * 4);
  uint8_t* m = (uint8_t*)malloc((size_t)(nc > 0 ? nc : 1));
  ora_filter(brects, B->n, prects, P->n, cb, cp, nc, nthreads);
  int rc = tgref_refine(build, probe, pred, cb, cp, nc, m, nthreads);
  int64_t w = 0;
  for (int64_t i = 0; i < nc; i++) {
    if (!m[i]) continue;
    if (ob && op && w < cap) { ob[w] = cb[i]; op[w] = cp[i]; }
    w++;
  }
  free(cb); free(cp); free(m);
  return rc ? -1 : w;
}
