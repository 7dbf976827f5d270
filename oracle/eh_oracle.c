/*
 * eh_oracle.c -- CPU restatement of EmptyHeaded's set-intersection pattern engine.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, and the CPU baseline
 * timed beside it.  Nothing in the product path loads this file.
 *
 * Reference files restated (paths relative to the reference repository root):
 *   src/MutableGraph.hpp:20-32,100-123,268-272,298-360   binary I/O, degree relabel
 *   src/SparseMatrix.hpp:156-183,249-266,406-494          row filter, pack, row images
 *   src/set/layouts/{uinteger,pshort,bitset,hybrid}.hpp   encodings, type rule
 *   src/set/ops/intersection.hpp                          intersection kernels + dispatch
 *   apps/undirected_triangle_counting.cpp:27-65           triangle driver
 *   apps/n_clique.cpp:30-101                              k-clique driver
 *   src/common.hpp:213-309                                dynamic chunk scheduler
 */
#define _GNU_SOURCE
#include "eh_oracle.h"

#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------- */
/* small helpers                                                             */
/* ------------------------------------------------------------------------- */

static inline uint64_t ld64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static inline uint32_t ld32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline void st64(uint8_t *p, uint64_t v) { memcpy(p, &v, 8); }
static inline void st32(uint8_t *p, uint32_t v) { memcpy(p, &v, 4); }
static inline uint16_t ld16(const uint8_t *p) { uint16_t v; memcpy(&v, p, 2); return v; }

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------- */
/* dynamic chunk scheduler: common::par_for_range (src/common.hpp:213-276).   */
/* Threads pull [start, start+block) ranges from one atomic counter.          */
/* ------------------------------------------------------------------------- */

typedef void (*orc_body_fn)(void *ctx, int tid, uint64_t i);

typedef struct {
  void *ctx;
  orc_body_fn body;
  uint64_t from, len, block;
  uint64_t next; /* shared atomic counter */
  int tid;
} par_arg;

typedef struct {
  par_arg *shared;
  int tid;
} par_thread;

static void *par_worker(void *p) {
  par_thread *t = (par_thread *)p;
  par_arg *a = t->shared;
  for (;;) {
    uint64_t s = __atomic_fetch_add(&a->next, a->block, __ATOMIC_RELAXED);
    if (s >= a->len) break;
    uint64_t e = s + a->block < a->len ? s + a->block : a->len;
    for (uint64_t j = s; j < e; j++) a->body(a->ctx, t->tid, a->from + j);
  }
  return NULL;
}

static int par_for_range(int threads, uint64_t from, uint64_t to, uint64_t block, orc_body_fn body,
                         void *ctx) {
  uint64_t len = to - from;
  uint64_t want = len / block + 1;
  int real = (int)(want < (uint64_t)threads ? want : (uint64_t)threads);
  if (real < 1) real = 1;
  if (real == 1) {
    for (uint64_t i = from; i < to; i++) body(ctx, 0, i);
    return 1;
  }
  par_arg a = {ctx, body, from, len, block, 0, 0};
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)real);
  par_thread *pt = (par_thread *)malloc(sizeof(par_thread) * (size_t)real);
  for (int k = 0; k < real; k++) {
    pt[k].shared = &a;
    pt[k].tid = k;
    pthread_create(&th[k], NULL, par_worker, &pt[k]);
  }
  for (int k = 0; k < real; k++) pthread_join(th[k], NULL);
  free(th);
  free(pt);
  return real;
}

/* ------------------------------------------------------------------------- */
/* graph I/O: MutableGraph::undirectedFromBinary / writeUndirectedToBinary    */
/* (src/MutableGraph.hpp:298-360).  Format: size_t n, then per node          */
/* { uint64 external id; size_t degree; uint32 nbr[degree] }.                */
/* ------------------------------------------------------------------------- */

int orc_load_binary(const char *path, uint32_t *n_out, uint64_t **off_out, uint32_t **nbr_out,
                    uint64_t **ext_out) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  uint64_t n = 0;
  if (fread(&n, 8, 1, f) != 1 || n > 0xFFFFFFFFull) { fclose(f); return -2; }
  /* first pass: sizes */
  long start = ftell(f);
  uint64_t total = 0;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t ext, d;
    if (fread(&ext, 8, 1, f) != 1 || fread(&d, 8, 1, f) != 1) { fclose(f); return -3; }
    total += d;
    if (fseek(f, (long)(d * 4), SEEK_CUR) != 0) { fclose(f); return -3; }
  }
  fseek(f, start, SEEK_SET);
  uint64_t *off = (uint64_t *)malloc(8 * (n + 1));
  uint32_t *nbr = (uint32_t *)malloc(4 * (total ? total : 1));
  uint64_t *ext = (uint64_t *)malloc(8 * (n ? n : 1));
  off[0] = 0;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t d;
    if (fread(&ext[i], 8, 1, f) != 1 || fread(&d, 8, 1, f) != 1) goto bad;
    off[i + 1] = off[i] + d;
    if (d && fread(nbr + off[i], 4, d, f) != d) goto bad;
  }
  fclose(f);
  *n_out = (uint32_t)n;
  *off_out = off;
  *nbr_out = nbr;
  if (ext_out) *ext_out = ext; else free(ext);
  return 0;
bad:
  fclose(f);
  free(off); free(nbr); free(ext);
  return -3;
}

int orc_write_binary(const char *path, uint32_t n, const uint64_t *off, const uint32_t *nbr,
                     const uint64_t *ext) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  uint64_t nn = n;
  fwrite(&nn, 8, 1, f);
  for (uint32_t i = 0; i < n; i++) {
    uint64_t e = ext ? ext[i] : i;
    uint64_t d = off[i + 1] - off[i];
    fwrite(&e, 8, 1, f);
    fwrite(&d, 8, 1, f);
    if (d) fwrite(nbr + off[i], 4, d, f);
  }
  fclose(f);
  return 0;
}

/* LSD radix sort of 64-bit keys, 16 bits per pass. */
static void radix_sort_u64(uint64_t *a, uint64_t n) {
  if (n < 2) return;
  uint64_t *tmp = (uint64_t *)malloc(8 * n);
  uint64_t *cnt = (uint64_t *)malloc(8 * 65536);
  uint64_t *src = a, *dst = tmp;
  for (int pass = 0; pass < 4; pass++) {
    int sh = 16 * pass;
    memset(cnt, 0, 8 * 65536);
    for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> sh) & 0xFFFF]++;
    uint64_t s = 0;
    for (int b = 0; b < 65536; b++) { uint64_t c = cnt[b]; cnt[b] = s; s += c; }
    for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
    uint64_t *t = src; src = dst; dst = t;
  }
  /* after 4 passes src == a */
  free(tmp);
  free(cnt);
}

/* Canonical undirected graph from an edge list: the project's generator
 * canonicalisation (self-loops dropped, (min,max) pairs deduplicated, both
 * directions stored; rows come out sorted). */
int orc_edges_to_csr(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                     uint64_t **off_out, uint32_t **nbr_out) {
  uint64_t *key = (uint64_t *)malloc(8 * (m ? m : 1));
  uint64_t k = 0;
  for (uint64_t e = 0; e < m; e++) {
    uint32_t u = src[e], v = dst[e];
    if (u == v || u >= n || v >= n) continue;
    uint32_t a = u < v ? u : v, b = u < v ? v : u;
    key[k++] = ((uint64_t)a << 32) | b;
  }
  radix_sort_u64(key, k);
  uint64_t u = 0;
  for (uint64_t i = 0; i < k; i++)
    if (i == 0 || key[i] != key[i - 1]) key[u++] = key[i];
  uint64_t *off = (uint64_t *)calloc((size_t)n + 1, 8);
  for (uint64_t i = 0; i < u; i++) {
    off[(key[i] >> 32) + 1]++;
    off[(uint32_t)key[i] + 1]++;
  }
  for (uint32_t i = 0; i < n; i++) off[i + 1] += off[i];
  uint32_t *nbr = (uint32_t *)malloc(4 * (off[n] ? off[n] : 1));
  uint64_t *pos = (uint64_t *)malloc(8 * ((size_t)n + 1));
  memcpy(pos, off, 8 * ((size_t)n + 1));
  for (uint64_t i = 0; i < u; i++) {
    uint32_t a = (uint32_t)(key[i] >> 32), b = (uint32_t)key[i];
    nbr[pos[a]++] = b;
    nbr[pos[b]++] = a;
  }
  free(pos);
  free(key);
  *off_out = off;
  *nbr_out = nbr;
  return 0;
}

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : x > y;
}

typedef struct {
  const uint64_t *off;
  const uint32_t *nbr;
  const uint32_t *new2old, *old2new;
  const uint64_t *off_out;
  uint32_t *nbr_out;
} relabel_ctx;

static void relabel_row(void *c, int tid, uint64_t r) {
  (void)tid;
  relabel_ctx *x = (relabel_ctx *)c;
  uint32_t o = x->new2old[r];
  uint64_t b = x->off[o], e = x->off[o + 1];
  uint32_t *dst = x->nbr_out + x->off_out[r];
  for (uint64_t k = b; k < e; k++) dst[k - b] = x->old2new[x->nbr[k]];
  qsort(dst, (size_t)(e - b), 4, cmp_u32);
}

/* MutableGraph::reorder_by_degree + OrderNeighborhoodByDegree + reassign_ids
 * (src/MutableGraph.hpp:20-32,100-123,268-272): new ids ordered by degree
 * descending, ties by old id ascending; every row remapped and re-sorted. */
int orc_relabel_by_degree(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t *new2old,
                          uint64_t *off_out, uint32_t *nbr_out) {
  uint64_t maxd = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint64_t d = off[i + 1] - off[i];
    if (d > maxd) maxd = d;
  }
  /* stable counting sort by degree descending keeps old ids ascending */
  uint64_t *bucket = (uint64_t *)calloc(maxd + 2, 8);
  for (uint32_t i = 0; i < n; i++) bucket[maxd - (off[i + 1] - off[i])]++;
  uint64_t s = 0;
  for (uint64_t b = 0; b <= maxd; b++) { uint64_t c = bucket[b]; bucket[b] = s; s += c; }
  for (uint32_t i = 0; i < n; i++) new2old[bucket[maxd - (off[i + 1] - off[i])]++] = i;
  free(bucket);
  uint32_t *old2new = (uint32_t *)malloc(4 * ((size_t)n + 1));
  for (uint32_t r = 0; r < n; r++) old2new[new2old[r]] = r;
  off_out[0] = 0;
  for (uint32_t r = 0; r < n; r++) off_out[r + 1] = off_out[r] + (off[new2old[r] + 1] - off[new2old[r]]);
  relabel_ctx c = {off, nbr, new2old, old2new, off_out, nbr_out};
  par_for_range(8, 0, n, 1024, relabel_row, &c);
  free(old2new);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Vertex orderings of tools/undirectedEdgeListToBinary.cpp:26-32             */
/* (MutableGraph::reorder_*, src/MutableGraph.hpp:100-285), sequential.  The  */
/* reference's std::random_shuffle is replaced by counter-based permutations: */
/* the ids sorted by the high 32 bits of draw(stream_key(seed, stream), id),  */
/* ties by id.  Orders are returned as new2old.                               */
/* ------------------------------------------------------------------------- */

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

void orc_perm(uint64_t seed, uint64_t stream, uint32_t n, uint32_t *new2old) {
  uint64_t key = orc_stream_key(seed, stream);
  uint64_t *k = (uint64_t *)malloc(8 * ((size_t)n + 1));
  for (uint32_t i = 0; i < n; i++) k[i] = ((orc_draw(key, i) >> 32) << 32) | i;
  qsort(k, n, 8, cmp_u64);
  for (uint32_t r = 0; r < n; r++) new2old[r] = (uint32_t)k[r];
  free(k);
}

/* stable order of ids by a 64-bit key per id */
static void order_by_key(uint32_t n, const uint64_t *key, uint32_t *new2old) {
  typedef struct { uint64_t k; uint32_t id; } kv;
  kv *a = (kv *)malloc(sizeof(kv) * ((size_t)n + 1));
  for (uint32_t i = 0; i < n; i++) { a[i].k = key[i]; a[i].id = i; }
  /* bottom-up merge sort: equal keys keep id order */
  kv *t = (kv *)malloc(sizeof(kv) * ((size_t)n + 1));
  for (uint64_t w = 1; w < n; w *= 2) {
    for (uint64_t lo = 0; lo < n; lo += 2 * w) {
      uint64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      uint64_t i = lo, j = mid, o = lo;
      while (i < mid && j < hi) t[o++] = (a[j].k < a[i].k) ? a[j++] : a[i++];
      while (i < mid) t[o++] = a[i++];
      while (j < hi) t[o++] = a[j++];
    }
    kv *x = a; a = t; t = x;
  }
  for (uint32_t r = 0; r < n; r++) new2old[r] = a[r].id;
  free(a);
  free(t);
}

static void order_degree(uint32_t n, const uint64_t *off, uint32_t *new2old, int descending) {
  uint64_t *key = (uint64_t *)calloc((size_t)n + 1, 8);
  for (uint32_t i = 0; i < n; i++) {
    uint64_t d = off[i + 1] - off[i];
    key[i] = descending ? ~d : d;
  }
  order_by_key(n, key, new2old);
  free(key);
}

/* reorder_strong_run (src/MutableGraph.hpp:179-201) */
static void order_strong_run(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t *new2old) {
  uint32_t *rows = (uint32_t *)malloc(4 * ((size_t)n + 1));
  order_degree(n, off, rows, 1);
  uint8_t *seen = (uint8_t *)calloc((size_t)n + 1, 1);
  uint32_t c = 0;
  for (uint32_t r = 0; r < n; r++) {
    uint32_t v = rows[r];
    for (uint64_t k = off[v]; k < off[v + 1]; k++)
      if (!seen[nbr[k]]) { seen[nbr[k]] = 1; new2old[c++] = nbr[k]; }
  }
  for (uint32_t u = 0; u < n; u++)
    if (!seen[u]) new2old[c++] = u;
  free(seen);
  free(rows);
}

/* reorder_bfs (src/MutableGraph.hpp:126-172) */
static void order_bfs(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint64_t seed, uint32_t *new2old) {
  uint32_t *tmp = (uint32_t *)malloc(4 * ((size_t)n + 1));
  orc_perm(seed, 102, n, tmp);
  uint8_t *visited = (uint8_t *)calloc((size_t)n + 1, 1);
  uint32_t *cur = (uint32_t *)malloc(4 * ((size_t)n + 1)), *next = (uint32_t *)malloc(4 * ((size_t)n + 1));
  size_t ncur = 0, nnext = 0, c = 0, bfs_i = 0;
  cur[ncur++] = tmp[bfs_i];
  new2old[c++] = tmp[bfs_i];
  visited[tmp[bfs_i]] = 1;
  bfs_i++;
  while (c != n) {
    for (size_t i = 0; i < ncur; i++) {
      uint32_t v = cur[i];
      for (uint64_t k = off[v]; k < off[v + 1]; k++)
        if (!visited[nbr[k]]) {
          next[nnext++] = nbr[k];
          new2old[c++] = nbr[k];
          visited[nbr[k]] = 1;
        }
    }
    if (nnext == 0 && c < n) {
      while (visited[tmp[bfs_i]]) bfs_i++;
      next[nnext++] = tmp[bfs_i];
      new2old[c++] = tmp[bfs_i];
      visited[tmp[bfs_i]] = 1;
      bfs_i++;
    }
    uint32_t *x = cur; cur = next; next = x;
    ncur = nnext;
    nnext = 0;
  }
  free(cur);
  free(next);
  free(visited);
  free(tmp);
}

/* reorder_by_shingles (src/MutableGraph.hpp:214-266) */
static void order_shingles(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint64_t seed, uint32_t *new2old) {
  uint32_t *ord = (uint32_t *)malloc(4 * ((size_t)n + 1));
  uint64_t *key = (uint64_t *)calloc((size_t)n + 1, 8);
  for (int s = 0; s < 2; s++) {
    orc_perm(seed, 103 + (uint64_t)s, n, ord);
    for (uint32_t j = 0; j < n; j++) {
      uint32_t sh = 0;
      if (off[j + 1] > off[j]) {
        sh = nbr[off[j]];
        uint32_t so = ord[sh];
        for (uint64_t k = off[j] + 1; k < off[j + 1]; k++)
          if (ord[nbr[k]] < so) { sh = nbr[k]; so = ord[sh]; }
      }
      key[j] |= (uint64_t)sh << (s == 0 ? 32 : 0);
    }
  }
  order_by_key(n, key, new2old);
  free(key);
  free(ord);
}

int orc_order(uint32_t n, const uint64_t *off, const uint32_t *nbr, int ordering, uint64_t seed, uint32_t *new2old) {
  switch (ordering) {
    case 0: order_degree(n, off, new2old, 1); return 0;
    case 1: order_degree(n, off, new2old, 0); return 0;
    case 2: order_strong_run(n, off, nbr, new2old); return 0;
    case 3: order_bfs(n, off, nbr, seed, new2old); return 0;
    case 4: order_shingles(n, off, nbr, seed, new2old); return 0;
    case 5: orc_perm(seed, 101, n, new2old); return 0;
    case 6: {
      /* reorder_by_the_game (src/MutableGraph.hpp:282-285): BFS, renumber, then degree */
      uint32_t *b = (uint32_t *)malloc(4 * ((size_t)n + 1)), *d = (uint32_t *)malloc(4 * ((size_t)n + 1));
      uint32_t *o2n = (uint32_t *)malloc(4 * ((size_t)n + 1));
      uint64_t *off2 = (uint64_t *)malloc(8 * ((size_t)n + 1));
      order_bfs(n, off, nbr, seed, b);
      for (uint32_t r = 0; r < n; r++) o2n[b[r]] = r;
      off2[0] = 0;
      for (uint32_t r = 0; r < n; r++) off2[r + 1] = off2[r] + (off[b[r] + 1] - off[b[r]]);
      order_degree(n, off2, d, 1);
      for (uint32_t r = 0; r < n; r++) new2old[r] = b[d[r]];
      free(b); free(d); free(o2n); free(off2);
      return 0;
    }
    default: return -1;
  }
}

/* Orientation filter nbr < node (apps/undirected_triangle_counting.cpp:20-23
 * applied by pack_data, src/SparseMatrix.hpp:249-266).  Rows are sorted, so
 * the kept entries are a prefix of each row. */
uint64_t orc_orient_lower(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint64_t *off_out,
                          uint32_t *nbr_out) {
  off_out[0] = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint64_t b = off[i], e = off[i + 1], k = b;
    while (k < e && nbr[k] < i) k++;
    if (nbr_out) memcpy(nbr_out + off_out[i], nbr + b, 4 * (k - b));
    off_out[i + 1] = off_out[i] + (k - b);
  }
  return off_out[n];
}

/* ------------------------------------------------------------------------- */
/* layouts                                                                   */
/* ------------------------------------------------------------------------- */

/* hybrid::get_type (src/set/layouts/hybrid.hpp:73-100) for the live two-way
 * rule; the HYBRID3 policy adds compute_type's pshort band
 * (src/set/layouts/hybrid.hpp:55-63) on the same density len/(last-first). */
int orc_row_type(const uint32_t *data, size_t len, int policy) {
  switch (policy) {
    case ORC_POLICY_UINT: return ORC_UINTEGER;
    case ORC_POLICY_PSHORT: return ORC_PSHORT;
    case ORC_POLICY_BITSET: return ORC_BITSET;
    default: break;
  }
  switch (policy) {
    case ORC_POLICY_BITSET_NEW: return ORC_BITSET_NEW;
    case ORC_POLICY_NEW_TYPE: return ORC_NEW_TYPE;
    default: break;
  }
  if (len <= 1) return ORC_UINTEGER;
  uint32_t range = data[len - 1] - data[0];
  if (range == 0) return ORC_UINTEGER;
  double density = (double)len / (double)range;
  if (policy == ORC_POLICY_HYBRID)
    return density > (1.0 / 256.0) ? ORC_BITSET : ORC_UINTEGER;
  if (policy == ORC_POLICY_HYBRID_BLOCK)
    return density > (1.0 / 256.0) ? ORC_BITSET_NEW : ORC_UINTEGER;
  if (density > 1.0 / 32.0) return ORC_BITSET;
  if (density > 1.0 / 256.0) return ORC_PSHORT;
  return ORC_UINTEGER;
}

/* uinteger::build (src/set/layouts/uinteger.hpp:49-53): raw copy. */
size_t orc_uint_build(uint8_t *out, const uint32_t *data, size_t len) {
  memcpy(out, data, 4 * len);
  return 4 * len;
}

/* pshort::build (src/set/layouts/pshort.hpp:46-78): partitions by the upper
 * 16 bits, each written as [prefix, count-1, low16 x count]. */
size_t orc_pshort_build(uint8_t *out, const uint32_t *data, size_t len) {
  if (len == 0) return 0;
  uint16_t *R = (uint16_t *)out;
  size_t w = 0, hdr = 0;
  uint32_t cur_hi = 0xFFFFFFFFu;
  size_t cnt = 0;
  for (size_t p = 0; p < len; p++) {
    uint32_t hi = data[p] >> 16;
    if (hi != cur_hi) {
      if (cur_hi != 0xFFFFFFFFu) R[hdr + 1] = (uint16_t)(cnt - 1);
      hdr = w;
      R[w++] = (uint16_t)hi;
      R[w++] = 0;
      cur_hi = hi;
      cnt = 0;
    }
    R[w++] = (uint16_t)(data[p] & 0xFFFF);
    cnt++;
  }
  R[hdr + 1] = (uint16_t)(cnt - 1);
  return 2 * w;
}

/* bitset::build (src/set/layouts/bitset.hpp:65-96): [u64 first word index]
 * followed by the 64-bit words from word(first) to word(last). */
size_t orc_bitset_build(uint8_t *out, const uint32_t *data, size_t len) {
  if (len == 0) return 0;
  uint64_t w0 = data[0] >> 6, w1 = data[len - 1] >> 6;
  st64(out, w0);
  uint8_t *W = out + 8;
  memset(W, 0, 8 * (size_t)(w1 - w0 + 1));
  for (size_t i = 0; i < len; i++) {
    uint64_t wi = (data[i] >> 6) - w0;
    uint64_t v = ld64(W + 8 * wi) | (1ull << (data[i] & 63));
    st64(W + 8 * wi, v);
  }
  return 8 + 8 * (size_t)(w1 - w0 + 1);
}

/* bitset_new::build (src/set/layouts/bitset_new.hpp:66-108), BLOCK_SIZE 128
 * (src/common.hpp:29): the u32 ids (value >> 7) of the non-empty blocks, then
 * each block's two 64-bit words (bit v & 127 of the block; word (v >> 6) & 1). */
size_t orc_bitset_new_build(uint8_t *out, const uint32_t *data, size_t len) {
  if (len == 0) return 0;
  size_t nb = 0;
  for (size_t i = 0; i < len; i++)
    if (i == 0 || (data[i] >> 7) != (data[i - 1] >> 7)) nb++;
  uint8_t *W = out + 4 * nb;
  memset(W, 0, 16 * nb);
  size_t b = 0;
  for (size_t i = 0; i < len; i++) {
    if (i > 0 && (data[i] >> 7) != (data[i - 1] >> 7)) b++;
    st32(out + 4 * b, data[i] >> 7);
    uint8_t *w = W + 16 * b + 8 * ((data[i] >> 6) & 1u);
    st64(w, ld64(w) | (1ull << (data[i] & 63)));
  }
  return 20 * nb;
}

/* new_type::build (src/set/layouts/new_type.hpp:50-88): a block whose element
 * count c has density (c < 2 ? 0 : c / 128) above BITSET_THRESHOLD = 1/16
 * (src/common.hpp:30) goes to a bitset_new part, every other element to a uint
 * part: [u64 uint bytes][uint part][bitset_new part]. */
size_t orc_new_type_build(uint8_t *out, const uint32_t *data, size_t len) {
  if (len == 0) return 0;
  uint32_t *dense = (uint32_t *)malloc(4 * len);
  size_t nu = 0, nd = 0;
  uint8_t *U = out + 8;
  for (size_t i = 0; i < len;) {
    size_t j = i;
    while (j < len && (data[j] >> 7) == (data[i] >> 7)) j++;
    const double density = (j - i) < 2 ? 0.0 : (double)(j - i) / 128.0;
    for (size_t t = i; t < j; t++) {
      if (density > 1.0 / 16.0) dense[nd++] = data[t];
      else st32(U + 4 * nu++, data[t]);
    }
    i = j;
  }
  st64(out, 4 * (uint64_t)nu);
  size_t bytes = 8 + 4 * nu + orc_bitset_new_build(out + 8 + 4 * nu, dense, nd);
  free(dense);
  return bytes;
}

size_t orc_row_image_bound(const uint32_t *data, size_t len) {
  size_t b = 8 + 4 + 8 + 20 * len + 16;
  if (len) b += 8 * (size_t)((data[len - 1] >> 6) - (data[0] >> 6) + 1) + 8;
  return b;
}

/* Bound for the layout the policy picks for this row only: the arena of a
 * large sparse matrix must not reserve a bitset span for every uint row. */
size_t orc_row_image_bound_policy(const uint32_t *data, size_t len, int policy) {
  size_t b = 8 + 4 + 16; /* type tag, sub-layout byte count, slack */
  int t = orc_row_type(data, len, policy);
  if (t == ORC_UINTEGER) b += 4 * len;
  else if (t == ORC_PSHORT) b += 6 * len; /* worst case: one partition per element */
  else if (t == ORC_BITSET_NEW || t == ORC_NEW_TYPE) b += 8 + 20 * len; /* a block per element */
  else if (len) b += 8 * (size_t)((data[len - 1] >> 6) - (data[0] >> 6) + 1) + 8;
  return b;
}

/* The flattened sub-layout of one type: uint is raw, bitset/pshort carry a
 * u32 byte count and are empty for empty rows (bitset.hpp:99-109,
 * pshort.hpp:81-91, uinteger.hpp:56-59). */
static size_t flat_build(uint8_t *out, int type, const uint32_t *data, size_t len) {
  if (type == ORC_UINTEGER) return orc_uint_build(out, data, len);
  if (len == 0) return 0;
  size_t nb = type == ORC_BITSET       ? orc_bitset_build(out + 4, data, len)
              : type == ORC_PSHORT     ? orc_pshort_build(out + 4, data, len)
              : type == ORC_BITSET_NEW ? orc_bitset_new_build(out + 4, data, len)
                                       : orc_new_type_build(out + 4, data, len);
  st32(out, (uint32_t)nb);
  return nb + 4;
}

/* Set<T>::flatten_from_array (src/set/Set.hpp:182-185); for the hybrid
 * layouts a u64 type tag precedes the sub-layout (hybrid.hpp:132-166). */
size_t orc_row_image(uint8_t *out, const uint32_t *data, size_t len, int policy) {
  int t = orc_row_type(data, len, policy);
  if (policy == ORC_POLICY_HYBRID || policy == ORC_POLICY_HYBRID3 || policy == ORC_POLICY_HYBRID_BLOCK) {
    st64(out, (uint64_t)t);
    return 8 + flat_build(out + 8, t, data, len);
  }
  return flat_build(out, t, data, len);
}

/* Set<T>::from_flattened + get_flattened_data (src/set/Set.hpp:172-176,
 * hybrid.hpp:168-201): locate the payload of a row image. */
static void view_image(const uint8_t *img, uint64_t card, int policy, orc_set *s) {
  int t;
  const uint8_t *p = img;
  if (policy == ORC_POLICY_HYBRID || policy == ORC_POLICY_HYBRID3 || policy == ORC_POLICY_HYBRID_BLOCK) {
    t = (int)ld64(p);
    p += 8;
  } else {
    t = policy == ORC_POLICY_UINT         ? ORC_UINTEGER
        : policy == ORC_POLICY_PSHORT     ? ORC_PSHORT
        : policy == ORC_POLICY_BITSET_NEW ? ORC_BITSET_NEW
        : policy == ORC_POLICY_NEW_TYPE   ? ORC_NEW_TYPE
                                          : ORC_BITSET;
  }
  s->type = t;
  s->cardinality = card;
  if (t == ORC_UINTEGER) {
    s->data = (uint8_t *)p;
    s->number_of_bytes = 4 * card;
  } else if (card > 0) {
    s->number_of_bytes = ld32(p);
    s->data = (uint8_t *)p + 4;
  } else {
    s->number_of_bytes = 0;
    s->data = (uint8_t *)p;
  }
}

/* foreach over a typed set payload (uinteger.hpp:67-80, bitset.hpp:147-173,
 * pshort.hpp:133-158). */
size_t orc_decode(int type, const uint8_t *payload, size_t card, size_t nbytes, uint32_t *out) {
  size_t k = 0;
  if (type == ORC_UINTEGER) {
    memcpy(out, payload, 4 * card);
    return card;
  }
  if (type == ORC_BITSET) {
    if (nbytes == 0) return 0;
    uint64_t w0 = ld64(payload);
    size_t nw = (nbytes - 8) / 8;
    for (size_t i = 0; i < nw; i++) {
      uint64_t w = ld64(payload + 8 + 8 * i);
      while (w) {
        int b = __builtin_ctzll(w);
        out[k++] = (uint32_t)(64 * (w0 + i) + (uint64_t)b);
        w &= w - 1;
      }
    }
    return k;
  }
  if (type == ORC_BITSET_NEW) {
    /* bitset_new::foreach (src/set/layouts/bitset_new.hpp:159-191) */
    size_t nb = nbytes / 20;
    for (size_t b = 0; b < nb; b++) {
      uint64_t base = 128ull * ld32(payload + 4 * b);
      for (int h = 0; h < 2; h++) {
        uint64_t w = ld64(payload + 4 * nb + 16 * b + 8 * h);
        while (w) {
          out[k++] = (uint32_t)(base + 64 * h + (uint64_t)__builtin_ctzll(w));
          w &= w - 1;
        }
      }
    }
    return k;
  }
  if (type == ORC_NEW_TYPE) {
    /* new_type::foreach (src/set/layouts/new_type.hpp:135-152): uint part, then blocks */
    if (nbytes == 0) return 0;
    size_t ub = (size_t)ld64(payload);
    memcpy(out, payload + 8, ub);
    k = ub / 4;
    return k + orc_decode(ORC_BITSET_NEW, payload + 8 + ub, card - k, nbytes - 8 - ub, out + k);
  }
  /* pshort */
  size_t i = 0;
  while (k < card) {
    uint32_t pre = (uint32_t)ld16(payload + 2 * i) << 16;
    size_t c = (size_t)ld16(payload + 2 * i + 2) + 1;
    i += 2;
    for (size_t t = 0; t < c; t++) out[k++] = pre | ld16(payload + 2 * (i + t));
    i += c;
  }
  return k;
}

/* ------------------------------------------------------------------------- */
/* intersection kernels                                                      */
/* ------------------------------------------------------------------------- */

/* Two-pointer merge (ops::scalar, src/set/ops/intersection.hpp:46-79). */
static size_t isect_scalar(const uint32_t *A, size_t la, const uint32_t *B, size_t lb,
                           uint32_t *out) {
  size_t i = 0, j = 0, c = 0;
  while (i < la && j < lb) {
    uint32_t a = A[i], b = B[j];
    if (a < b) i++;
    else if (a > b) j++;
    else {
      if (out) out[c] = a;
      c++; i++; j++;
    }
  }
  return c;
}

/* 4x4 block all-pairs compare (set_intersect_standard,
 * src/set/ops/intersection.hpp:696-771): compare a block of four from each
 * side under all four rotations, count the hit mask, advance the block(s)
 * whose maximum is smaller; the tail goes to the scalar merge. */
static size_t isect_standard(const uint32_t *A, size_t la, const uint32_t *B, size_t lb,
                             uint32_t *out) {
  size_t ia = 0, ib = 0, c = 0;
  size_t qa = la & ~(size_t)3, qb = lb & ~(size_t)3;
  while (ia < qa && ib < qb) {
    __m128i va = _mm_loadu_si128((const __m128i *)(A + ia));
    __m128i vb = _mm_loadu_si128((const __m128i *)(B + ib));
    uint32_t amax = A[ia + 3], bmax = B[ib + 3];
    __m128i r1 = _mm_shuffle_epi32(vb, 0x39);
    __m128i r2 = _mm_shuffle_epi32(vb, 0x4E);
    __m128i r3 = _mm_shuffle_epi32(vb, 0x93);
    __m128i hit = _mm_or_si128(_mm_or_si128(_mm_cmpeq_epi32(va, vb), _mm_cmpeq_epi32(va, r1)),
                               _mm_or_si128(_mm_cmpeq_epi32(va, r2), _mm_cmpeq_epi32(va, r3)));
    int mask = _mm_movemask_ps(_mm_castsi128_ps(hit));
    if (out) {
      for (int t = 0; t < 4; t++)
        if (mask & (1 << t)) out[c + (size_t)__builtin_popcount(mask & ((1 << t) - 1))] = A[ia + t];
    }
    c += (size_t)__builtin_popcount(mask);
    ia += (amax <= bmax) ? 4 : 0;
    ib += (amax >= bmax) ? 4 : 0;
  }
  return c + isect_scalar(A + ia, la - ia, B + ib, lb - ib, out ? out + c : NULL);
}

/* Galloping intersection (set_intersect_galloping,
 * src/set/ops/intersection.hpp:393-584).  For every rare element, the frequent
 * side moves forward in blocks of 128 by exponential then binary search over
 * block maxima; the element is then compared against the 32-element quarter of
 * the block that can hold it.  The last rare element and what remains of the
 * frequent side finish in the scalar merge. */
static size_t isect_galloping(const uint32_t *rare, size_t lr, const uint32_t *freq, size_t lf,
                              uint32_t *out) {
  if (lr == 0 || lf == 0) return 0;
  if (lf < 128) return isect_scalar(freq, lf, rare, lr, out);
  const uint32_t *stop_freq = freq + lf - 128;
  const uint32_t *stop_rare = rare + lr - 1;
  size_t c = 0;
  for (; rare < stop_rare; ++rare) {
    const uint32_t x = *rare;
    if (freq[127] < x) {
      size_t hi = 1;
      if (freq + 128 > stop_freq) { freq += 128; goto tail; }
      while (freq[hi * 128 + 127] < x) {
        if (freq + 2 * hi * 128 <= stop_freq) {
          hi *= 2;
        } else if (freq + (hi + 1) * 128 <= stop_freq) {
          hi = (size_t)(stop_freq - freq) / 128;
          if (freq[hi * 128 + 127] < x) { freq += hi * 128; goto tail; }
          break;
        } else {
          freq += hi * 128;
          goto tail;
        }
      }
      size_t lo = hi / 2;
      while (lo + 1 != hi) {
        size_t mid = (lo + hi) / 2;
        if (freq[mid * 128 + 127] < x) lo = mid; else hi = mid;
      }
      freq += hi * 128;
    }
    {
      const uint32_t *q;
      if (freq[63] >= x) q = (freq[31] < x) ? freq + 32 : freq;
      else q = (freq[95] < x) ? freq + 96 : freq + 64;
      __m128i m = _mm_set1_epi32((int)x);
      __m128i acc = _mm_setzero_si128();
      for (int v = 0; v < 8; v++)
        acc = _mm_or_si128(acc, _mm_cmpeq_epi32(_mm_loadu_si128((const __m128i *)(q + 4 * v)), m));
      if (!_mm_testz_si128(acc, acc)) {
        if (out) out[c] = x;
        c++;
      }
    }
  }
tail:
  return c + isect_scalar(freq, (size_t)(stop_freq + 128 - freq), rare, (size_t)(stop_rare + 1 - rare),
                          out ? out + c : NULL);
}

/* uint x uint selector (src/set/ops/intersection.hpp:1325-1336): the smaller
 * side is rare; galloping when max/max(1,min) > 32 (integer division),
 * otherwise the 4x4 block compare. */
static size_t isect_select(const uint32_t *A, size_t la, const uint32_t *B, size_t lb,
                           uint32_t *out) {
  const uint32_t *rare = la > lb ? B : A, *freq = la > lb ? A : B;
  size_t lr = la > lb ? lb : la, lf = la > lb ? la : lb;
  size_t mn = lr > 1 ? lr : 1;
  if (lf / mn > 32) return isect_galloping(rare, lr, freq, lf, out);
  return isect_standard(rare, lr, freq, lf, out);
}

size_t orc_isect_uint(int alg, const uint32_t *a, size_t la, const uint32_t *b, size_t lb,
                      uint32_t *out) {
  switch (alg) {
    case ORC_ALG_SCALAR: return isect_scalar(a, la, b, lb, out);
    case ORC_ALG_STANDARD: return isect_standard(a, la, b, lb, out);
    case ORC_ALG_GALLOPING:
      return la <= lb ? isect_galloping(a, la, b, lb, out) : isect_galloping(b, lb, a, la, out);
    default: return isect_select(a, la, b, lb, out);
  }
}

/* uint x bitset probe (src/set/ops/intersection.hpp:1023-1058): each uint
 * element tests its bit when its word lies inside the bitset's word span; the
 * scan stops at the first element past the span. */
static size_t isect_u_bs(const uint32_t *A, size_t la, const uint8_t *bs, size_t bs_bytes,
                         uint32_t *out) {
  if (bs_bytes == 0) return 0;
  uint64_t w0 = ld64(bs);
  uint64_t nw = (bs_bytes - 8) / 8;
  const uint8_t *W = bs + 8;
  size_t c = 0;
  for (size_t i = 0; i < la; i++) {
    uint64_t w = A[i] >> 6;
    if (w >= w0 + nw) break;
    if (w >= w0 && ((ld64(W + 8 * (w - w0)) >> (A[i] & 63)) & 1)) {
      if (out) out[c] = A[i];
      c++;
    }
  }
  return c;
}

/* bitset x bitset (src/set/ops/intersection.hpp:880-955): AND plus popcount
 * over the overlapping word span; a materialised result is a bitset starting
 * at the overlap's first word. */
static size_t isect_bs_bs(const uint8_t *a, size_t a_bytes, const uint8_t *b, size_t b_bytes,
                          uint8_t *out, size_t *out_bytes) {
  if (out_bytes) *out_bytes = 0;
  if (a_bytes == 0 || b_bytes == 0) return 0;
  uint64_t a0 = ld64(a), b0 = ld64(b);
  uint64_t sa = (a_bytes - 8) / 8, sb = (b_bytes - 8) / 8;
  uint64_t start = a0 > b0 ? a0 : b0;
  uint64_t end = (a0 + sa) < (b0 + sb) ? (a0 + sa) : (b0 + sb);
  uint64_t total = start > end ? 0 : end - start;
  const uint8_t *A = a + 8 + 8 * (start - a0), *B = b + 8 + 8 * (start - b0);
  size_t c = 0;
  uint64_t i = 0;
  if (out) st64(out, start);
  for (; i + 4 <= total; i += 4) {
    __m256i x = _mm256_loadu_si256((const __m256i *)(A + 8 * i));
    __m256i y = _mm256_loadu_si256((const __m256i *)(B + 8 * i));
    __m256i r = _mm256_and_si256(x, y);
    if (out) _mm256_storeu_si256((__m256i *)(out + 8 + 8 * i), r);
    c += (size_t)__builtin_popcountll((uint64_t)_mm256_extract_epi64(r, 0));
    c += (size_t)__builtin_popcountll((uint64_t)_mm256_extract_epi64(r, 1));
    c += (size_t)__builtin_popcountll((uint64_t)_mm256_extract_epi64(r, 2));
    c += (size_t)__builtin_popcountll((uint64_t)_mm256_extract_epi64(r, 3));
  }
  for (; i < total; i++) {
    uint64_t r = ld64(A + 8 * i) & ld64(B + 8 * i);
    if (out) st64(out + 8 + 8 * i, r);
    c += (size_t)__builtin_popcountll(r);
  }
  if (out_bytes) *out_bytes = 8 * total + 8;
  return c;
}

/* 16-bit intersection inside one matched pshort partition
 * (simd_intersect_vector16, src/set/ops/intersection.hpp:773-828): 8x8 block
 * equal-any compare with SSE4.2, then a scalar tail. */
static size_t isect_u16(const uint16_t *A, size_t la, const uint16_t *B, size_t lb, uint16_t *out) {
  size_t ia = 0, ib = 0, c = 0;
  size_t qa = la & ~(size_t)7, qb = lb & ~(size_t)7;
  while (ia < qa && ib < qb) {
    __m128i va = _mm_loadu_si128((const __m128i *)(A + ia));
    __m128i vb = _mm_loadu_si128((const __m128i *)(B + ib));
    uint16_t amax = A[ia + 7], bmax = B[ib + 7];
    __m128i r = _mm_cmpestrm(vb, 8, va, 8, _SIDD_UWORD_OPS | _SIDD_CMP_EQUAL_ANY | _SIDD_BIT_MASK);
    uint32_t mask = (uint32_t)_mm_extract_epi32(r, 0);
    if (out)
      for (int t = 0; t < 8; t++)
        if (mask & (1u << t)) out[c + (size_t)__builtin_popcount(mask & ((1u << t) - 1))] = A[ia + t];
    c += (size_t)__builtin_popcount(mask);
    ia += (amax <= bmax) ? 8 : 0;
    ib += (amax >= bmax) ? 8 : 0;
  }
  while (ia < la && ib < lb) {
    if (A[ia] < B[ib]) ia++;
    else if (A[ia] > B[ib]) ib++;
    else {
      if (out) out[c] = A[ia];
      c++; ia++; ib++;
    }
  }
  return c;
}

/* pshort x pshort (src/set/ops/intersection.hpp:829-878): merge on partition
 * prefixes; equal prefixes intersect their 16-bit runs.  A materialised result
 * is again a pshort stream (empty partitions are dropped). */
static size_t isect_ps_ps(const uint8_t *a, size_t a_bytes, const uint8_t *b, size_t b_bytes,
                          uint16_t *out, size_t *out_bytes) {
  const uint16_t *A = (const uint16_t *)a, *B = (const uint16_t *)b;
  size_t sa = a_bytes / 2, sb = b_bytes / 2, ia = 0, ib = 0, w = 0, c = 0;
  while (ia < sa && ib < sb) {
    if (A[ia] < B[ib]) ia += (size_t)A[ia + 1] + 3;
    else if (B[ib] < A[ia]) ib += (size_t)B[ib + 1] + 3;
    else {
      size_t la = (size_t)A[ia + 1] + 1, lb = (size_t)B[ib + 1] + 1;
      size_t pc = isect_u16(A + ia + 2, la, B + ib + 2, lb, out ? out + w + 2 : NULL);
      if (out && pc) {
        out[w] = A[ia];
        out[w + 1] = (uint16_t)(pc - 1);
        w += 2 + pc;
      }
      c += pc;
      ia += la + 2;
      ib += lb + 2;
    }
  }
  if (out_bytes) *out_bytes = 2 * w;
  return c;
}

/* uint x pshort (src/set/ops/intersection.hpp:1062-1173): walk the pshort
 * partitions; uint elements below a partition's prefix are skipped, the ones
 * sharing it are merged against the partition's 16-bit run. */
static size_t isect_u_ps(const uint32_t *A, size_t la, const uint8_t *b, size_t b_bytes,
                         uint32_t *out) {
  const uint16_t *B = (const uint16_t *)b;
  size_t sb = b_bytes / 2, ia = 0, ib = 0, c = 0;
  while (ia < la && ib < sb) {
    uint32_t pre = (uint32_t)B[ib] << 16;
    size_t cnt = (size_t)B[ib + 1] + 1;
    const uint16_t *run = B + ib + 2;
    if (pre < (A[ia] & 0xFFFF0000u)) { ib += cnt + 2; continue; }
    if (pre > A[ia]) { ia++; continue; }
    size_t t = 0;
    while (ia < la && t < cnt && (A[ia] & 0xFFFF0000u) == pre) {
      uint32_t v = pre | run[t];
      if (v < A[ia]) t++;
      else if (v > A[ia]) ia++;
      else {
        if (out) out[c] = A[ia];
        c++; ia++; t++;
      }
    }
    ib += cnt + 2;
  }
  return c;
}

/* pshort x bitset (src/set/ops/intersection.hpp:956-1017): partitions whose
 * 1024-word window overlaps the bitset probe every element.  A materialised
 * result is written as a well-formed pshort stream. */
static size_t isect_ps_bs(const uint8_t *a, size_t a_bytes, const uint8_t *bs, size_t bs_bytes,
                          uint16_t *out, size_t *out_bytes) {
  const uint16_t *A = (const uint16_t *)a;
  size_t sa = a_bytes / 2, i = 0, c = 0, w = 0;
  uint64_t w0 = bs_bytes ? ld64(bs) : 0, nw = bs_bytes ? (bs_bytes - 8) / 8 : 0;
  const uint8_t *W = bs + 8;
  while (i < sa) {
    uint32_t pre = (uint32_t)A[i] << 16;
    size_t cnt = (size_t)A[i + 1] + 1;
    uint64_t pw = pre >> 6;
    if (pw >= w0 + nw) break;
    if (pw + 1024 >= w0) {
      size_t hdr = w, pc = 0;
      if (out) w += 2;
      for (size_t t = 0; t < cnt; t++) {
        uint32_t v = pre | A[i + 2 + t];
        uint64_t vw = v >> 6;
        if (vw < w0 + nw && vw >= w0 && ((ld64(W + 8 * (vw - w0)) >> (v & 63)) & 1)) {
          if (out) out[w++] = A[i + 2 + t];
          pc++;
        }
      }
      if (out) {
        if (pc) { out[hdr] = A[i]; out[hdr + 1] = (uint16_t)(pc - 1); }
        else w = hdr;
      }
      c += pc;
    }
    i += cnt + 2;
  }
  if (out_bytes) *out_bytes = 2 * w;
  return c;
}

/* bitset_new x bitset_new (src/set/ops/intersection.hpp:1338-1413): the blocks
 * present on both sides, in id order (simd_intersect_offsets, :1174-1320), are
 * AND-ed word by word (BLOCK_SIZE 128 < 256: two 64-bit words, :1354-1362).  The
 * result is a bitset_new set of every matched block, empty ANDs included. */
static size_t isect_bn_bn(const uint8_t *a, size_t a_bytes, const uint8_t *b, size_t b_bytes, uint8_t *out,
                          size_t *out_bytes) {
  const size_t na = a_bytes / 20, nb = b_bytes / 20;
  const uint8_t *AW = a + 4 * na, *BW = b + 4 * nb;
  size_t k = 0;
  for (size_t i = 0, j = 0; i < na && j < nb;) {
    uint32_t x = ld32(a + 4 * i), y = ld32(b + 4 * j);
    if (x < y) i++;
    else if (x > y) j++;
    else { k++; i++; j++; }
  }
  size_t c = 0, t = 0;
  for (size_t i = 0, j = 0; i < na && j < nb;) {
    uint32_t x = ld32(a + 4 * i), y = ld32(b + 4 * j);
    if (x < y) { i++; continue; }
    if (x > y) { j++; continue; }
    for (int h = 0; h < 2; h++) {
      uint64_t r = ld64(AW + 16 * i + 8 * h) & ld64(BW + 16 * j + 8 * h);
      if (out) st64(out + 4 * k + 16 * t + 8 * h, r);
      c += (size_t)__builtin_popcountll(r);
    }
    if (out) st32(out + 4 * t, x);
    t++; i++; j++;
  }
  *out_bytes = 20 * k;
  return c;
}

/* uint x bitset_new (src/set/ops/intersection.hpp:1425-1509): each uint value
 * whose block (value >> 7) is present probes its bit; the hits form a uint set. */
static size_t isect_u_bn(const uint32_t *A, size_t la, const uint8_t *b, size_t b_bytes, uint32_t *out) {
  const size_t nb = b_bytes / 20;
  const uint8_t *BW = b + 4 * nb;
  size_t c = 0, j = 0;
  for (size_t i = 0; i < la && j < nb; i++) {
    const uint32_t blk = A[i] >> 7;
    while (j < nb && ld32(b + 4 * j) < blk) j++;
    if (j < nb && ld32(b + 4 * j) == blk) {
      const uint32_t v = A[i] & 127u;
      if ((ld64(BW + 16 * j + 8 * (v >> 6)) >> (v & 63)) & 1) {
        if (out) out[c] = A[i];
        c++;
      }
    }
  }
  return c;
}

/* merge of disjoint sorted lists (distinct_merge_three_way, :1511-1553) */
static void merge3(uint32_t *out, const uint32_t *x, size_t nx, const uint32_t *y, size_t ny, const uint32_t *z,
                   size_t nz) {
  size_t i = 0, j = 0, k = 0, w = 0;
  while (i < nx || j < ny || k < nz) {
    uint32_t vx = i < nx ? x[i] : 0xFFFFFFFFu, vy = j < ny ? y[j] : 0xFFFFFFFFu, vz = k < nz ? z[k] : 0xFFFFFFFFu;
    if (i < nx && vx <= vy && vx <= vz) out[w++] = x[i++];
    else if (j < ny && vy <= vz) out[w++] = y[j++];
    else out[w++] = z[k++];
  }
}

/* new_type x new_type (src/set/ops/intersection.hpp:1555-1649): uint parts
 * against each other and against the other side's blocks, merged into the
 * result's uint part; the block parts AND-ed into its bitset_new part.  When
 * neither side has blocks the result is the uint intersection alone. */
static size_t isect_nt_nt(const uint8_t *a, uint64_t a_card, size_t a_bytes, const uint8_t *b, uint64_t b_card,
                          size_t b_bytes, uint8_t *out, size_t *out_bytes) {
  (void)a_card; (void)b_card;
  const size_t aub = (size_t)ld64(a), bub = (size_t)ld64(b);
  const uint32_t *AU = (const uint32_t *)(a + 8), *BU = (const uint32_t *)(b + 8);
  const size_t anu = aub / 4, bnu = bub / 4;
  const uint8_t *AB = a + 8 + aub, *BB = b + 8 + bub;
  const size_t abb = a_bytes - 8 - aub, bbb = b_bytes - 8 - bub;
  size_t c = 0, nu = 0;
  uint32_t *U = out ? (uint32_t *)(out + 8) : NULL;
  if ((aub != 0 || bub != 0) && (abb != 0 || bbb != 0)) {
    uint32_t *s1 = NULL, *s2 = NULL, *s3 = NULL;
    if (out) {
      s1 = (uint32_t *)malloc(4 * (anu + 1));
      s2 = (uint32_t *)malloc(4 * (anu + 1));
      s3 = (uint32_t *)malloc(4 * (bnu + 1));
    }
    size_t c1 = isect_select(AU, anu, BU, bnu, s1);
    size_t c2 = bbb ? isect_u_bn(AU, anu, BB, bbb, s2) : 0;
    size_t c3 = abb ? isect_u_bn(BU, bnu, AB, abb, s3) : 0;
    nu = c1 + c2 + c3;
    if (out) {
      merge3(U, s1, c1, s2, c2, s3, c3);
      free(s1); free(s2); free(s3);
    }
  } else if (abb == 0 && bbb == 0) {
    nu = isect_select(AU, anu, BU, bnu, U);
    if (out) st64(out, 4 * (uint64_t)nu);
    *out_bytes = 8 + 4 * nu;
    return nu;
  }
  c = nu;
  if (out) st64(out, 4 * (uint64_t)nu);
  size_t bsb = 0;
  if (abb != 0 && bbb != 0) c += isect_bn_bn(AB, abb, BB, bbb, out ? out + 8 + 4 * nu : NULL, &bsb);
  *out_bytes = 8 + 4 * nu + bsb;
  return c;
}

/* Typed-set dispatcher (ops::set_intersect(Set<hybrid>*...),
 * src/set/ops/intersection.hpp:2066-2118, widened with the pshort pairs of
 * :829-1021,1062-1173,1321-1323).  Result types follow the reference: U for
 * any pair with a uint side, BS for BS x BS, PS for PS x PS and PS x BS. */
uint64_t orc_set_intersect(orc_set *C, const orc_set *A, const orc_set *B) {
  uint8_t *o = C ? C->data : NULL;
  size_t ob = 0;
  uint64_t c = 0;
  int rt = ORC_UINTEGER;
  if (A->cardinality == 0 || B->cardinality == 0) {
    if (C) { C->cardinality = 0; C->number_of_bytes = 0; }
    return 0;
  }
  int ta = A->type, tb = B->type;
  if (ta == ORC_UINTEGER && tb == ORC_UINTEGER) {
    c = isect_select((const uint32_t *)A->data, A->cardinality, (const uint32_t *)B->data,
                     B->cardinality, (uint32_t *)o);
    ob = 4 * c;
  } else if (ta == ORC_UINTEGER && tb == ORC_BITSET) {
    c = isect_u_bs((const uint32_t *)A->data, A->cardinality, B->data, B->number_of_bytes, (uint32_t *)o);
    ob = 4 * c;
  } else if (ta == ORC_BITSET && tb == ORC_UINTEGER) {
    c = isect_u_bs((const uint32_t *)B->data, B->cardinality, A->data, A->number_of_bytes, (uint32_t *)o);
    ob = 4 * c;
  } else if (ta == ORC_BITSET && tb == ORC_BITSET) {
    c = isect_bs_bs(A->data, A->number_of_bytes, B->data, B->number_of_bytes, o, &ob);
    rt = ORC_BITSET;
  } else if (ta == ORC_PSHORT && tb == ORC_PSHORT) {
    c = isect_ps_ps(A->data, A->number_of_bytes, B->data, B->number_of_bytes, (uint16_t *)o, &ob);
    rt = ORC_PSHORT;
  } else if (ta == ORC_UINTEGER && tb == ORC_PSHORT) {
    c = isect_u_ps((const uint32_t *)A->data, A->cardinality, B->data, B->number_of_bytes, (uint32_t *)o);
    ob = 4 * c;
  } else if (ta == ORC_PSHORT && tb == ORC_UINTEGER) {
    c = isect_u_ps((const uint32_t *)B->data, B->cardinality, A->data, A->number_of_bytes, (uint32_t *)o);
    ob = 4 * c;
  } else if (ta == ORC_PSHORT && tb == ORC_BITSET) {
    c = isect_ps_bs(A->data, A->number_of_bytes, B->data, B->number_of_bytes, (uint16_t *)o, &ob);
    rt = ORC_PSHORT;
  } else if (ta == ORC_BITSET && tb == ORC_PSHORT) {
    c = isect_ps_bs(B->data, B->number_of_bytes, A->data, A->number_of_bytes, (uint16_t *)o, &ob);
    rt = ORC_PSHORT;
  } else if (ta == ORC_BITSET_NEW && tb == ORC_BITSET_NEW) {
    c = isect_bn_bn(A->data, A->number_of_bytes, B->data, B->number_of_bytes, o, &ob);
    rt = ORC_BITSET_NEW;
  } else if (ta == ORC_UINTEGER && tb == ORC_BITSET_NEW) {
    c = isect_u_bn((const uint32_t *)A->data, A->cardinality, B->data, B->number_of_bytes, (uint32_t *)o);
    ob = 4 * c;
  } else if (ta == ORC_BITSET_NEW && tb == ORC_UINTEGER) {
    c = isect_u_bn((const uint32_t *)B->data, B->cardinality, A->data, A->number_of_bytes, (uint32_t *)o);
    ob = 4 * c;
  } else if (ta == ORC_NEW_TYPE && tb == ORC_NEW_TYPE) {
    c = isect_nt_nt(A->data, A->cardinality, A->number_of_bytes, B->data, B->cardinality, B->number_of_bytes, o,
                    &ob);
    rt = ORC_NEW_TYPE;
  } else {
    fprintf(stderr, "orc_set_intersect: unknown type pair %d/%d\n", ta, tb);
    return 0;
  }
  if (C) {
    C->cardinality = c;
    C->number_of_bytes = ob;
    C->type = rt;
  }
  return c;
}

/* ------------------------------------------------------------------------- */
/* SparseMatrix (src/SparseMatrix.hpp:406-494): one contiguous arena of row   */
/* images; row i's image starts at row_pos[i] and holds row_lengths[i]        */
/* elements.                                                                  */
/* ------------------------------------------------------------------------- */

struct orc_matrix {
  uint32_t n;
  int policy;
  uint64_t card;
  uint32_t max_row;
  uint32_t *row_lengths;
  uint64_t *row_pos;
  uint8_t *arena;
  uint64_t bytes;
};

typedef struct {
  const uint64_t *off;
  const uint32_t *nbr;
  int lower;
  int policy;
  uint32_t *len;
  uint64_t *size;
  uint8_t *arena;
  const uint64_t *pos;
  uint32_t **scratch;
} build_ctx;

static size_t select_row(const build_ctx *b, uint64_t i, uint32_t *buf) {
  uint64_t s = b->off[i], e = b->off[i + 1];
  size_t k = 0;
  for (uint64_t t = s; t < e; t++)
    if (!b->lower || b->nbr[t] < (uint32_t)i) buf[k++] = b->nbr[t];
  return k;
}

static void build_size(void *c, int tid, uint64_t i) {
  build_ctx *b = (build_ctx *)c;
  uint32_t *buf = b->scratch[tid];
  size_t k = select_row(b, i, buf);
  b->len[i] = (uint32_t)k;
  b->size[i] = orc_row_image_bound_policy(buf, k, b->policy);
}

static void build_fill(void *c, int tid, uint64_t i) {
  build_ctx *b = (build_ctx *)c;
  uint32_t *buf = b->scratch[tid];
  size_t k = select_row(b, i, buf);
  size_t w = orc_row_image(b->arena + b->pos[i], buf, k, b->policy);
  b->size[i] = w;
}

orc_matrix *orc_matrix_build(uint32_t n, const uint64_t *off, const uint32_t *nbr, int lower_only,
                             int policy, int threads) {
  if (threads < 1) threads = 1;
  orc_matrix *M = (orc_matrix *)calloc(1, sizeof(orc_matrix));
  M->n = n;
  M->policy = policy;
  M->row_lengths = (uint32_t *)malloc(4 * ((size_t)n + 1));
  M->row_pos = (uint64_t *)malloc(8 * ((size_t)n + 1));
  uint64_t *size = (uint64_t *)malloc(8 * ((size_t)n + 1));
  uint64_t maxd = 0;
  for (uint32_t i = 0; i < n; i++)
    if (off[i + 1] - off[i] > maxd) maxd = off[i + 1] - off[i];
  uint32_t **scratch = (uint32_t **)malloc(sizeof(uint32_t *) * (size_t)threads);
  for (int t = 0; t < threads; t++) scratch[t] = (uint32_t *)malloc(4 * (maxd + 1));
  build_ctx b = {off, nbr, lower_only, policy, M->row_lengths, size, NULL, NULL, scratch};
  par_for_range(threads, 0, n, 256, build_size, &b);
  uint64_t p = 0;
  for (uint32_t i = 0; i < n; i++) { M->row_pos[i] = p; p += (size[i] + 7) & ~7ull; }
  M->row_pos[n] = p;
  M->arena = (uint8_t *)malloc(p + 64);
  if (!M->arena) {
    fprintf(stderr, "orc_matrix_build: cannot allocate %llu bytes of row images\n", (unsigned long long)(p + 64));
    for (int t = 0; t < threads; t++) free(scratch[t]);
    free(scratch);
    free(size);
    free(M->row_lengths);
    free(M->row_pos);
    free(M);
    return NULL;
  }
  b.arena = M->arena;
  b.pos = M->row_pos;
  par_for_range(threads, 0, n, 256, build_fill, &b);
  M->bytes = 0;
  M->card = 0;
  M->max_row = 0;
  for (uint32_t i = 0; i < n; i++) {
    M->bytes += size[i];
    M->card += M->row_lengths[i];
    if (M->row_lengths[i] > M->max_row) M->max_row = M->row_lengths[i];
  }
  for (int t = 0; t < threads; t++) free(scratch[t]);
  free(scratch);
  free(size);
  return M;
}

void orc_matrix_free(orc_matrix *M) {
  if (!M) return;
  free(M->row_lengths);
  free(M->row_pos);
  free(M->arena);
  free(M);
}

uint32_t orc_matrix_rows(const orc_matrix *M) { return M->n; }
uint64_t orc_matrix_cardinality(const orc_matrix *M) { return M->card; }
uint32_t orc_matrix_max_row(const orc_matrix *M) { return M->max_row; }
uint64_t orc_matrix_bytes(const orc_matrix *M) { return M->bytes; }

/* SparseMatrix::get_row (src/SparseMatrix.hpp:156-160). */
int orc_matrix_get_row(const orc_matrix *M, uint32_t row, orc_set *out) {
  if (row >= M->n) return -1;
  view_image(M->arena + M->row_pos[row], M->row_lengths[row], M->policy, out);
  return 0;
}

void orc_matrix_type_counts(const orc_matrix *M, uint64_t counts[3]) {
  counts[0] = counts[1] = counts[2] = 0;
  for (uint32_t i = 0; i < M->n; i++) {
    orc_set s;
    orc_matrix_get_row(M, i, &s);
    if (s.cardinality == 0) continue;
    if (s.type >= 0 && s.type <= 2) counts[s.type]++;
    else if (s.type == ORC_BITSET_NEW || s.type == ORC_NEW_TYPE) counts[ORC_BITSET]++;
  }
}

/* iterate the elements of a typed set */
typedef void (*elem_fn)(void *ctx, uint32_t x);
static void set_foreach(const orc_set *s, elem_fn fn, void *ctx) {
  if (s->cardinality == 0) return;
  if (s->type == ORC_UINTEGER) {
    const uint32_t *d = (const uint32_t *)s->data;
    for (uint64_t i = 0; i < s->cardinality; i++) fn(ctx, d[i]);
  } else if (s->type == ORC_BITSET) {
    if (s->number_of_bytes == 0) return;
    uint64_t w0 = ld64(s->data);
    uint64_t nw = (s->number_of_bytes - 8) / 8;
    for (uint64_t i = 0; i < nw; i++) {
      uint64_t w = ld64(s->data + 8 + 8 * i);
      while (w) {
        int b = __builtin_ctzll(w);
        fn(ctx, (uint32_t)(64 * (w0 + i) + (uint64_t)b));
        w &= w - 1;
      }
    }
  } else if (s->type == ORC_BITSET_NEW || s->type == ORC_NEW_TYPE) {
    /* bitset_new::foreach, new_type::foreach (bitset_new.hpp:159-191, new_type.hpp:135-152) */
    const uint8_t *p = s->data;
    uint64_t nbytes = s->number_of_bytes;
    if (s->type == ORC_NEW_TYPE) {
      if (nbytes == 0) return;
      uint64_t ub = ld64(p);
      for (uint64_t i = 0; i < ub / 4; i++) fn(ctx, ld32(p + 8 + 4 * i));
      p += 8 + ub;
      nbytes -= 8 + ub;
    }
    uint64_t nb = nbytes / 20;
    for (uint64_t b = 0; b < nb; b++) {
      uint64_t base = 128ull * ld32(p + 4 * b);
      for (int h = 0; h < 2; h++) {
        uint64_t w = ld64(p + 4 * nb + 16 * b + 8 * h);
        while (w) {
          fn(ctx, (uint32_t)(base + 64 * h + (uint64_t)__builtin_ctzll(w)));
          w &= w - 1;
        }
      }
    }
  } else {
    const uint16_t *d = (const uint16_t *)s->data;
    uint64_t k = 0, i = 0;
    while (k < s->cardinality) {
      uint32_t pre = (uint32_t)d[i] << 16;
      uint64_t cnt = (uint64_t)d[i + 1] + 1;
      for (uint64_t t = 0; t < cnt; t++) fn(ctx, pre | d[i + 2 + t]);
      k += cnt;
      i += cnt + 2;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* triangle counting (apps/undirected_triangle_counting.cpp:27-65)            */
/* ------------------------------------------------------------------------- */

#define ORC_PAD 16
typedef struct {
  const orc_matrix *M;
  uint64_t *count; /* per thread, padded */
  uint64_t *edges;
  uint32_t stride, phase;
  int k;
  uint8_t **bufs;
  size_t buf_bytes;
  uint64_t *per_row; /* optional: count closed at each source row */
} tc_ctx;

typedef struct {
  const orc_matrix *M;
  const orc_set *A;
  uint64_t acc;
} tc_inner;

static void tc_visit(void *c, uint32_t j) {
  tc_inner *t = (tc_inner *)c;
  orc_set B;
  orc_matrix_get_row(t->M, j, &B);
  t->acc += orc_set_intersect(NULL, t->A, &B);
}

static void tc_row(void *c, int tid, uint64_t i) {
  tc_ctx *x = (tc_ctx *)c;
  if (x->stride > 1 && (i % x->stride) != x->phase) return;
  orc_set A;
  orc_matrix_get_row(x->M, (uint32_t)i, &A);
  tc_inner t = {x->M, &A, 0};
  set_foreach(&A, tc_visit, &t);
  if (x->per_row) x->per_row[i] = t.acc;
  x->count[(size_t)tid * ORC_PAD] += t.acc;
  x->edges[(size_t)tid * ORC_PAD] += A.cardinality;
}

uint64_t orc_triangle_count_sample(const orc_matrix *M, int threads, uint32_t stride,
                                   uint32_t phase, uint64_t *edges_done, double *seconds) {
  if (threads < 1) threads = 1;
  uint64_t *cnt = (uint64_t *)calloc((size_t)threads * ORC_PAD, 8);
  uint64_t *edg = (uint64_t *)calloc((size_t)threads * ORC_PAD, 8);
  tc_ctx x = {M, cnt, edg, stride, phase, 0, NULL, 0, NULL};
  double t0 = now_s();
  par_for_range(threads, 0, M->n, 100, tc_row, &x);
  double t1 = now_s();
  uint64_t total = 0, e = 0;
  for (int t = 0; t < threads; t++) { total += cnt[(size_t)t * ORC_PAD]; e += edg[(size_t)t * ORC_PAD]; }
  free(cnt);
  free(edg);
  if (seconds) *seconds = t1 - t0;
  if (edges_done) *edges_done = e;
  return total;
}

uint64_t orc_triangle_count(const orc_matrix *M, int threads, double *seconds) {
  return orc_triangle_count_sample(M, threads, 1, 0, NULL, seconds);
}

/* ------------------------------------------------------------------------- */
/* k-clique (apps/n_clique.cpp:30-101): level sets are materialised; the last */
/* level only counts.                                                         */
/* ------------------------------------------------------------------------- */

typedef struct {
  const orc_matrix *M;
  int k;
  uint8_t **level; /* level[d] buffer for depth d */
  orc_set *sets;   /* sets[d] */
  uint64_t acc;
} clq_state;

static uint64_t clq_apply(clq_state *s, uint32_t node, int depth);

typedef struct {
  clq_state *s;
  int depth;
  uint64_t acc;
} clq_visit_ctx;

static void clq_visit(void *c, uint32_t x) {
  clq_visit_ctx *v = (clq_visit_ctx *)c;
  v->acc += clq_apply(v->s, x, v->depth);
}

/* n_clique::apply_function (apps/n_clique.cpp:30-50) */
static uint64_t clq_apply(clq_state *s, uint32_t node, int depth) {
  orc_set B;
  orc_matrix_get_row(s->M, node, &B);
  orc_set *A = &s->sets[depth - 1];
  if (depth + 1 == s->k) return orc_set_intersect(NULL, A, &B);
  orc_set *C = &s->sets[depth];
  C->data = s->level[depth];
  orc_set_intersect(C, A, &B);
  orc_set Cc = *C;
  clq_visit_ctx v = {s, depth + 1, 0};
  set_foreach(&Cc, clq_visit, &v);
  return v.acc;
}

static void clq_row(void *c, int tid, uint64_t i) {
  tc_ctx *x = (tc_ctx *)c;
  if (x->stride > 1 && (i % x->stride) != x->phase) return;
  int k = x->k;
  uint8_t *base = x->bufs[tid];
  uint8_t *lv[32];
  for (int d = 0; d < k && d < 32; d++) lv[d] = base + (size_t)d * x->buf_bytes;
  orc_set sets[32];
  clq_state s = {x->M, k, lv, sets, 0};
  orc_matrix_get_row(x->M, (uint32_t)i, &sets[1]);
  clq_visit_ctx v = {&s, 2, 0};
  set_foreach(&sets[1], clq_visit, &v);
  if (x->per_row) x->per_row[i] = v.acc;
  x->count[(size_t)tid * ORC_PAD] += v.acc;
  x->edges[(size_t)tid * ORC_PAD] += sets[1].cardinality;
}

uint64_t orc_clique_count_sample(const orc_matrix *M, int k, int threads, uint32_t stride,
                                 uint32_t phase, uint64_t *edges_done, double *seconds) {
  if (k < 3 || k > 31) return 0;
  if (threads < 1) threads = 1;
  size_t bb = 8 * ((size_t)M->n / 64 + 4) + 24 * ((size_t)M->max_row + 16) + 64; /* new_type: 8 + 20 per element */
  uint8_t **bufs = (uint8_t **)malloc(sizeof(uint8_t *) * (size_t)threads);
  for (int t = 0; t < threads; t++) bufs[t] = (uint8_t *)malloc(bb * (size_t)k);
  uint64_t *cnt = (uint64_t *)calloc((size_t)threads * ORC_PAD, 8);
  uint64_t *edg = (uint64_t *)calloc((size_t)threads * ORC_PAD, 8);
  tc_ctx x = {M, cnt, edg, stride, phase, k, bufs, bb, NULL};
  double t0 = now_s();
  par_for_range(threads, 0, M->n, 100, clq_row, &x);
  double t1 = now_s();
  uint64_t total = 0, e = 0;
  for (int t = 0; t < threads; t++) { total += cnt[(size_t)t * ORC_PAD]; e += edg[(size_t)t * ORC_PAD]; }
  for (int t = 0; t < threads; t++) free(bufs[t]);
  free(bufs);
  free(cnt);
  free(edg);
  if (seconds) *seconds = t1 - t0;
  if (edges_done) *edges_done = e;
  return total;
}

uint64_t orc_clique_count(const orc_matrix *M, int k, int threads, double *seconds) {
  return orc_clique_count_sample(M, k, threads, 1, 0, NULL, seconds);
}

/* Pair mix of the triangle workload: counts of (i->j) pairs by the type pair
 * the dispatcher sees: [U x U, U x BS (either order), BS x BS, other]. */
uint64_t orc_pair_mix(const orc_matrix *M, uint64_t mix[4]) {
  mix[0] = mix[1] = mix[2] = mix[3] = 0;
  uint64_t tot = 0;
  uint32_t *buf = (uint32_t *)malloc(4 * ((size_t)M->max_row + 1));
  for (uint32_t i = 0; i < M->n; i++) {
    orc_set A;
    orc_matrix_get_row(M, i, &A);
    size_t k = orc_decode(A.type, A.data, A.cardinality, A.number_of_bytes, buf);
    for (size_t t = 0; t < k; t++) {
      orc_set B;
      orc_matrix_get_row(M, buf[t], &B);
      tot++;
      if (B.cardinality == 0) { mix[3]++; continue; }
      int a = A.type == ORC_BITSET_NEW ? ORC_BITSET : A.type, b = B.type == ORC_BITSET_NEW ? ORC_BITSET : B.type;
      if (a == ORC_UINTEGER && b == ORC_UINTEGER) mix[0]++;
      else if ((a == ORC_UINTEGER && b == ORC_BITSET) || (a == ORC_BITSET && b == ORC_UINTEGER)) mix[1]++;
      else if (a == ORC_BITSET && b == ORC_BITSET) mix[2]++;
      else mix[3]++;
    }
  }
  free(buf);
  return tot;
}

/* ------------------------------------------------------------------------- */
/* BFS path length (apps/n_path.cpp:41-85)                                    */
/* ------------------------------------------------------------------------- */

uint32_t orc_bfs_path(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t start, uint32_t max_depth,
                      uint32_t *rounds, uint64_t *level_sizes, uint32_t cap, uint64_t *reached) {
  const size_t words = (size_t)n / 64 + 2;
  uint64_t *visited = (uint64_t *)calloc(words, 8), *old = (uint64_t *)calloc(words, 8);
  uint32_t *frontier = (uint32_t *)malloc(4 * ((size_t)n + 1));
  size_t nf = 1;
  uint32_t path = 0, r = 0;
  frontier[0] = start;
  visited[start >> 6] |= 1ull << (start & 63);
  for (;;) {
    memcpy(old, visited, 8 * words);
    /* set_union of every frontier row into the visited bitset */
    for (size_t f = 0; f < nf; f++)
      for (uint64_t k = off[frontier[f]]; k < off[frontier[f] + 1]; k++)
        visited[nbr[k] >> 6] |= 1ull << (nbr[k] & 63);
    /* set_difference(visited, old): the new frontier in id order */
    nf = 0;
    for (size_t w = 0; w < words; w++) {
      uint64_t d = visited[w] & ~old[w];
      while (d) {
        frontier[nf++] = (uint32_t)(64 * w + (size_t)__builtin_ctzll(d));
        d &= d - 1;
      }
    }
    if (r < cap && level_sizes) level_sizes[r] = nf;
    r++;
    if (nf == 0 || path >= max_depth) break;
    path++;
  }
  if (rounds) *rounds = r;
  if (reached) {
    uint64_t c = 0;
    for (size_t w = 0; w < words; w++) c += (uint64_t)__builtin_popcountll(visited[w]);
    *reached = c;
  }
  free(visited);
  free(old);
  free(frontier);
  return path;
}

/* The same search with the frontier's set_union spread over threads as n_path does it
 * (apps/n_path.cpp:48-77: frontier.par_foreach into per-thread visited bitsets, then their
 * union): chunks of 100 frontier rows (par_for_range), each thread ORs its rows into its own
 * bitset, the bitsets are merged into the visited set, and the new frontier is the
 * difference with the old visited set, in id order.  Same results as orc_bfs_path. */
typedef struct {
  const uint64_t *off;
  const uint32_t *nbr;
  const uint32_t *frontier;
  uint64_t **bufs;
} bfs_par_ctx;

static void bfs_par_body(void *c, int tid, uint64_t f) {
  const bfs_par_ctx *b = (const bfs_par_ctx *)c;
  uint64_t *vb = b->bufs[tid];
  const uint32_t v = b->frontier[f];
  for (uint64_t k = b->off[v]; k < b->off[v + 1]; k++) vb[b->nbr[k] >> 6] |= 1ull << (b->nbr[k] & 63);
}

uint32_t orc_bfs_path_mt(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t start, uint32_t max_depth,
                         uint32_t *rounds, uint64_t *level_sizes, uint32_t cap, uint64_t *reached, int threads) {
  if (threads < 1) threads = 1;
  const size_t words = (size_t)n / 64 + 2;
  uint64_t *visited = (uint64_t *)calloc(words, 8), *old = (uint64_t *)calloc(words, 8);
  uint64_t **bufs = (uint64_t **)malloc(sizeof(uint64_t *) * (size_t)threads);
  for (int t = 0; t < threads; t++) bufs[t] = (uint64_t *)calloc(words, 8);
  uint32_t *frontier = (uint32_t *)malloc(4 * ((size_t)n + 1));
  size_t nf = 1;
  uint32_t path = 0, r = 0;
  frontier[0] = start;
  visited[start >> 6] |= 1ull << (start & 63);
  bfs_par_ctx c = {off, nbr, frontier, bufs};
  for (;;) {
    memcpy(old, visited, 8 * words);
    const int used = par_for_range(threads, 0, nf, 100, bfs_par_body, &c);
    for (int t = 0; t < used; t++) {
      for (size_t w = 0; w < words; w++) visited[w] |= bufs[t][w];
      memset(bufs[t], 0, 8 * words);
    }
    nf = 0;
    for (size_t w = 0; w < words; w++) {
      uint64_t d = visited[w] & ~old[w];
      while (d) {
        frontier[nf++] = (uint32_t)(64 * w + (size_t)__builtin_ctzll(d));
        d &= d - 1;
      }
    }
    if (r < cap && level_sizes) level_sizes[r] = nf;
    r++;
    if (nf == 0 || path >= max_depth) break;
    path++;
  }
  if (rounds) *rounds = r;
  if (reached) {
    uint64_t cnt = 0;
    for (size_t w = 0; w < words; w++) cnt += (uint64_t)__builtin_popcountll(visited[w]);
    *reached = cnt;
  }
  for (int t = 0; t < threads; t++) free(bufs[t]);
  free(bufs);
  free(visited);
  free(old);
  free(frontier);
  return path;
}

/* set_difference (src/set/ops/difference.hpp:5-84) on bitset images: A's words AND NOT
 * B's over their common span, A's words outside it as they are */
size_t orc_bitset_difference(const uint8_t *a, size_t a_bytes, const uint8_t *b, size_t b_bytes, uint32_t *out) {
  if (a_bytes < 8) return 0;
  const uint64_t a0 = ld64(a), na = (a_bytes - 8) / 8;
  const uint64_t b0 = b_bytes >= 8 ? ld64(b) : 0, nb = b_bytes >= 8 ? (b_bytes - 8) / 8 : 0;
  size_t c = 0;
  for (uint64_t i = 0; i < na; i++) {
    const uint64_t w = a0 + i;
    uint64_t r = ld64(a + 8 + 8 * i);
    if (w >= b0 && w < b0 + nb) r &= ~ld64(b + 8 + 8 * (w - b0));
    while (r) {
      out[c++] = (uint32_t)(64 * w + (uint64_t)__builtin_ctzll(r));
      r &= r - 1;
    }
  }
  return c;
}

/* ------------------------------------------------------------------------- */
/* independent brute force on the symmetric graph (no orientation, no        */
/* layouts): marker arrays over vertex ids.                                   */
/* ------------------------------------------------------------------------- */

uint64_t orc_brute_triangles(uint32_t n, const uint64_t *off, const uint32_t *nbr) {
  uint8_t *mark = (uint8_t *)calloc((size_t)n + 1, 1);
  uint64_t t = 0;
  for (uint32_t u = 0; u < n; u++) {
    for (uint64_t a = off[u]; a < off[u + 1]; a++) mark[nbr[a]] = 1;
    for (uint64_t a = off[u]; a < off[u + 1]; a++) {
      uint32_t v = nbr[a];
      if (v <= u) continue;
      for (uint64_t b = off[v]; b < off[v + 1]; b++) {
        uint32_t w = nbr[b];
        if (w > v && mark[w]) t++;
      }
    }
    for (uint64_t a = off[u]; a < off[u + 1]; a++) mark[nbr[a]] = 0;
  }
  free(mark);
  return t;
}

static uint64_t brute_extend(uint32_t n, const uint64_t *off, const uint32_t *nbr,
                             const uint32_t *cand, size_t nc, int need, uint32_t *scratch,
                             uint8_t *mark) {
  if (need == 0) return 1;
  if (need == 1) return nc;
  uint64_t t = 0;
  for (size_t a = 0; a < nc; a++) {
    uint32_t v = cand[a];
    /* next candidates: members of cand after v that are adjacent to v */
    for (uint64_t b = off[v]; b < off[v + 1]; b++) mark[nbr[b]] = 1;
    size_t k = 0;
    uint32_t *next = scratch;
    for (size_t c = a + 1; c < nc; c++)
      if (mark[cand[c]]) next[k++] = cand[c];
    for (uint64_t b = off[v]; b < off[v + 1]; b++) mark[nbr[b]] = 0;
    t += brute_extend(n, off, nbr, next, k, need - 1, scratch + k, mark);
  }
  return t;
}

uint64_t orc_brute_cliques(uint32_t n, const uint64_t *off, const uint32_t *nbr, int k) {
  if (k < 1) return 0;
  uint64_t maxd = 0;
  for (uint32_t u = 0; u < n; u++)
    if (off[u + 1] - off[u] > maxd) maxd = off[u + 1] - off[u];
  uint32_t *scratch = (uint32_t *)malloc(4 * ((size_t)maxd + 1) * (size_t)(k + 1));
  uint32_t *cand = (uint32_t *)malloc(4 * ((size_t)maxd + 1));
  uint8_t *mark = (uint8_t *)calloc((size_t)n + 1, 1);
  uint64_t t = 0;
  for (uint32_t u = 0; u < n; u++) {
    size_t nc = 0;
    for (uint64_t a = off[u]; a < off[u + 1]; a++)
      if (nbr[a] > u) cand[nc++] = nbr[a];
    t += brute_extend(n, off, nbr, cand, nc, k - 1, scratch, mark);
  }
  free(scratch);
  free(cand);
  free(mark);
  return t;
}

/* ------------------------------------------------------------------------- */
/* seeded generators: this project's specification, restated for checking    */
/* the device generators.                                                    */
/*   mix64       SplitMix64 output finaliser                                  */
/*   stream key  mix64(seed + mix64(stream))                                  */
/*   draw(k,c)   mix64(k + (c+1) * 0x9E3779B97F4A7C15)                         */
/* ------------------------------------------------------------------------- */

#define GOLDEN 0x9E3779B97F4A7C15ull

uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_stream_key(uint64_t seed, uint64_t stream) { return orc_mix64(seed + orc_mix64(stream)); }

uint64_t orc_draw(uint64_t key, uint64_t ctr) { return orc_mix64(key + (ctr + 1) * GOLDEN); }

static inline uint32_t scale32(uint32_t r, uint32_t n) { return (uint32_t)(((uint64_t)r * n) >> 32); }

/* Erdos-Renyi G(n, m draws): draw e gives endpoints from its two 32-bit halves. */
void orc_gen_er(uint32_t n, uint64_t m, uint64_t key, uint32_t *src, uint32_t *dst) {
  for (uint64_t e = 0; e < m; e++) {
    uint64_t r = orc_draw(key, e);
    src[e] = scale32((uint32_t)r, n);
    dst[e] = scale32((uint32_t)(r >> 32), n);
  }
}

void orc_rmat_thresholds(double a, double b, double c, uint32_t thr[3]) {
  thr[0] = (uint32_t)(a * 4294967296.0);
  thr[1] = (uint32_t)((a + b) * 4294967296.0);
  thr[2] = (uint32_t)((a + b + c) * 4294967296.0);
}

/* RMAT: one 32-bit draw per level, most significant bit first; quadrant
 * (0,0) below thr[0], (0,1) below thr[1], (1,0) below thr[2], else (1,1). */
void orc_gen_rmat(uint32_t scale, uint64_t m, const uint32_t thr[3], uint64_t key, uint32_t *src,
                  uint32_t *dst) {
  uint64_t per = (scale + 1) / 2;
  for (uint64_t e = 0; e < m; e++) {
    uint32_t u = 0, v = 0;
    uint64_t r = 0;
    for (uint32_t l = 0; l < scale; l++) {
      if ((l & 1) == 0) r = orc_draw(key, e * per + l / 2);
      uint32_t x = (l & 1) ? (uint32_t)(r >> 32) : (uint32_t)r;
      uint32_t bu = x >= thr[1], bv = (x >= thr[0] && x < thr[1]) || x >= thr[2];
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    src[e] = u;
    dst[e] = v;
  }
}

/* Chung-Lu weights w_i = (i+1)^(-1/(gamma-1)); cdf[i] = floor(2^32 * prefix/total). */
void orc_chung_lu_cdf(uint32_t n, double gamma, uint64_t *cdf) {
  double ex = -1.0 / (gamma - 1.0), total = 0.0;
  for (uint32_t i = 0; i < n; i++) total += pow((double)i + 1.0, ex);
  double acc = 0.0;
  for (uint32_t i = 0; i < n; i++) {
    acc += pow((double)i + 1.0, ex);
    double q = acc / total;
    cdf[i] = (uint64_t)(q * 4294967296.0);
  }
  cdf[n - 1] = 4294967296ull;
}

static uint32_t cdf_pick(const uint64_t *cdf, uint32_t n, uint32_t r) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)r < cdf[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}

void orc_gen_chung_lu(uint32_t n, uint64_t m, const uint64_t *cdf, uint64_t key, uint32_t *src,
                      uint32_t *dst) {
  for (uint64_t e = 0; e < m; e++) {
    uint64_t r = orc_draw(key, e);
    src[e] = cdf_pick(cdf, n, (uint32_t)r);
    dst[e] = cdf_pick(cdf, n, (uint32_t)(r >> 32));
  }
}

/* Planted communities: community c covers ids [c*(n/ncomm), +csize); pair
 * (s<t) is present when the top bit of draw(stream_key(seed, 1+c), s*csize+t)
 * is set.  Returns the number of edges written. */
uint64_t orc_gen_communities(uint32_t n, uint32_t ncomm, uint32_t csize, uint64_t seed,
                             uint32_t *src, uint32_t *dst) {
  uint64_t w = 0;
  uint32_t stride = n / ncomm;
  for (uint32_t c = 0; c < ncomm; c++) {
    uint64_t key = orc_stream_key(seed, 1 + (uint64_t)c);
    uint32_t base = c * stride;
    for (uint32_t s = 0; s < csize; s++)
      for (uint32_t t = s + 1; t < csize; t++) {
        if (orc_draw(key, (uint64_t)s * csize + t) >> 63) {
          if (src) { src[w] = base + s; dst[w] = base + t; }
          w++;
        }
      }
  }
  return w;
}

/* Batch member g: n uniform in [64, 8192], mean degree uniform in [4, 64),
 * m draws = n * degree / 2, all from stream 0 of the batch seed. */
void orc_batch_params(uint64_t seed, uint32_t g, uint32_t *n_out, uint64_t *m_out) {
  uint64_t key = orc_stream_key(seed, 0);
  uint64_t r = orc_draw(key, g);
  uint32_t n = 64 + (uint32_t)(((uint64_t)(uint32_t)r * 8129ull) >> 32);
  uint64_t deg_q32 = 4ull * 4294967296ull + 60ull * (uint64_t)(uint32_t)(r >> 32);
  *n_out = n;
  *m_out = ((uint64_t)n * deg_q32) >> 33;
}

/* ------------------------------------------------------------------------- */
/* FP64 reference of the layout-selector MLP 16 -> 64 -> 64 -> 4 (ReLU).     */
/* Parameter order: W1[64][16], b1[64], W2[64][64], b2[64], W3[4][64], b3[4]. */
/* Loss: masked mean squared error over the 4 outputs, averaged over the      */
/* labelled outputs of the batch.                                             */
/* ------------------------------------------------------------------------- */

#define MLP_IN 16
#define MLP_H 64
#define MLP_OUT 4

static void mlp_offsets(size_t *w1, size_t *b1, size_t *w2, size_t *b2, size_t *w3, size_t *b3) {
  *w1 = 0;
  *b1 = *w1 + MLP_H * MLP_IN;
  *w2 = *b1 + MLP_H;
  *b2 = *w2 + MLP_H * MLP_H;
  *w3 = *b2 + MLP_H;
  *b3 = *w3 + MLP_OUT * MLP_H;
}

void orc_mlp_forward(const float *P, const float *x, int batch, double *out) {
  size_t w1, b1, w2, b2, w3, b3;
  mlp_offsets(&w1, &b1, &w2, &b2, &w3, &b3);
  for (int s = 0; s < batch; s++) {
    double h1[MLP_H], h2[MLP_H];
    for (int o = 0; o < MLP_H; o++) {
      double a = P[b1 + o];
      for (int i = 0; i < MLP_IN; i++) a += (double)P[w1 + o * MLP_IN + i] * x[s * MLP_IN + i];
      h1[o] = a > 0 ? a : 0;
    }
    for (int o = 0; o < MLP_H; o++) {
      double a = P[b2 + o];
      for (int i = 0; i < MLP_H; i++) a += (double)P[w2 + o * MLP_H + i] * h1[i];
      h2[o] = a > 0 ? a : 0;
    }
    for (int o = 0; o < MLP_OUT; o++) {
      double a = P[b3 + o];
      for (int i = 0; i < MLP_H; i++) a += (double)P[w3 + o * MLP_H + i] * h2[i];
      out[s * MLP_OUT + o] = a;
    }
  }
}

double orc_mlp_loss_grad(const float *P, const float *x, const float *target, const float *mask,
                         int batch, double *grad) {
  size_t w1, b1, w2, b2, w3, b3;
  mlp_offsets(&w1, &b1, &w2, &b2, &w3, &b3);
  size_t np = b3 + MLP_OUT;
  for (size_t i = 0; i < np; i++) grad[i] = 0;
  double lab = 0;
  for (int s = 0; s < batch * MLP_OUT; s++) lab += mask[s];
  if (lab <= 0) return 0;
  double loss = 0;
  for (int s = 0; s < batch; s++) {
    double z1[MLP_H], h1[MLP_H], z2[MLP_H], h2[MLP_H], y[MLP_OUT];
    const float *xs = x + s * MLP_IN;
    for (int o = 0; o < MLP_H; o++) {
      double a = P[b1 + o];
      for (int i = 0; i < MLP_IN; i++) a += (double)P[w1 + o * MLP_IN + i] * xs[i];
      z1[o] = a;
      h1[o] = a > 0 ? a : 0;
    }
    for (int o = 0; o < MLP_H; o++) {
      double a = P[b2 + o];
      for (int i = 0; i < MLP_H; i++) a += (double)P[w2 + o * MLP_H + i] * h1[i];
      z2[o] = a;
      h2[o] = a > 0 ? a : 0;
    }
    double dy[MLP_OUT];
    for (int o = 0; o < MLP_OUT; o++) {
      double a = P[b3 + o];
      for (int i = 0; i < MLP_H; i++) a += (double)P[w3 + o * MLP_H + i] * h2[i];
      y[o] = a;
      double m = mask[s * MLP_OUT + o];
      double d = y[o] - target[s * MLP_OUT + o];
      loss += m * d * d;
      dy[o] = m * 2.0 * d / lab;
    }
    double dh2[MLP_H] = {0};
    for (int o = 0; o < MLP_OUT; o++) {
      grad[b3 + o] += dy[o];
      for (int i = 0; i < MLP_H; i++) {
        grad[w3 + o * MLP_H + i] += dy[o] * h2[i];
        dh2[i] += dy[o] * P[w3 + o * MLP_H + i];
      }
    }
    double dh1[MLP_H] = {0};
    for (int o = 0; o < MLP_H; o++) {
      double dz = z2[o] > 0 ? dh2[o] : 0;
      grad[b2 + o] += dz;
      for (int i = 0; i < MLP_H; i++) {
        grad[w2 + o * MLP_H + i] += dz * h1[i];
        dh1[i] += dz * P[w2 + o * MLP_H + i];
      }
    }
    for (int o = 0; o < MLP_H; o++) {
      double dz = z1[o] > 0 ? dh1[o] : 0;
      grad[b1 + o] += dz;
      for (int i = 0; i < MLP_IN; i++) grad[w1 + o * MLP_IN + i] += dz * xs[i];
    }
  }
  return loss / lab;
}

/* per-source-row counts of the triangle (k == 3) or k-clique pass */
uint64_t orc_count_rows(const orc_matrix *M, int k, int threads, uint64_t *per_row) {
  if (threads < 1) threads = 1;
  memset(per_row, 0, 8 * (size_t)M->n);
  uint64_t *cnt = (uint64_t *)calloc((size_t)threads * ORC_PAD, 8);
  uint64_t *edg = (uint64_t *)calloc((size_t)threads * ORC_PAD, 8);
  uint64_t total = 0;
  if (k == 3) {
    tc_ctx x = {M, cnt, edg, 1, 0, 0, NULL, 0, per_row};
    par_for_range(threads, 0, M->n, 100, tc_row, &x);
  } else {
    size_t bb = 8 * ((size_t)M->n / 64 + 4) + 24 * ((size_t)M->max_row + 16) + 64; /* new_type: 8 + 20 per element */
    uint8_t **bufs = (uint8_t **)malloc(sizeof(uint8_t *) * (size_t)threads);
    for (int t = 0; t < threads; t++) bufs[t] = (uint8_t *)malloc(bb * (size_t)k);
    tc_ctx x = {M, cnt, edg, 1, 0, k, bufs, bb, per_row};
    par_for_range(threads, 0, M->n, 100, clq_row, &x);
    for (int t = 0; t < threads; t++) free(bufs[t]);
    free(bufs);
  }
  for (int t = 0; t < threads; t++) total += cnt[(size_t)t * ORC_PAD];
  free(cnt);
  free(edg);
  return total;
}
