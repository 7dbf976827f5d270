/*
 * eh_oracle.h -- CPU restatement of EmptyHeaded's set-intersection pattern engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * simdgraphprocessing_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls into it.
 *
 * Every function names the reference file:line whose behaviour it restates
 * (paths are relative to the reference repository root).  The code is written
 * from the algorithm descriptions, not copied.
 */
#ifndef EH_ORACLE_H
#define EH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Set type tags: same values as common::type (src/common.hpp:171-181). */
enum { ORC_BITSET = 0, ORC_PSHORT = 1, ORC_UINTEGER = 2, ORC_HYBRID = 5, ORC_BITSET_NEW = 7,
       ORC_NEW_TYPE = 8 };

/* Whole-matrix layout policies.  UINT/PSHORT/BITSET are the single-layout
 * --layout=uint|pshort|bs paths (src/emptyheaded.hpp:45-56); HYBRID is the
 * live two-way rule (src/set/layouts/hybrid.hpp:73-100); HYBRID3 adds the
 * pshort band of compute_type (src/set/layouts/hybrid.hpp:55-63). */
enum { ORC_POLICY_UINT = 0, ORC_POLICY_PSHORT = 1, ORC_POLICY_BITSET = 2,
       ORC_POLICY_HYBRID = 3, ORC_POLICY_HYBRID3 = 4 };
/* Block-bitset policies (128-value blocks, src/common.hpp:28-30): every row
 * bitset_new (--layout=bitset_new), every row new_type (--layout=new_type)
 * (src/emptyheaded.hpp:60-65), and the hybrid rule built with NEW_BITSET, whose
 * dense rows are bitset_new (src/set/layouts/hybrid.hpp:83-87). */
enum { ORC_POLICY_BITSET_NEW = 7, ORC_POLICY_NEW_TYPE = 8, ORC_POLICY_HYBRID_BLOCK = 9 };

/* Intersection algorithm identifiers for orc_isect_uint (tests pin each). */
enum { ORC_ALG_SCALAR = 0, ORC_ALG_STANDARD = 1, ORC_ALG_GALLOPING = 2, ORC_ALG_SELECT = 3 };

/* ---- graph I/O and preprocessing ---- */
int orc_load_binary(const char *path, uint32_t *n_out, uint64_t **off_out, uint32_t **nbr_out,
                    uint64_t **ext_out);
int orc_write_binary(const char *path, uint32_t n, const uint64_t *off, const uint32_t *nbr,
                     const uint64_t *ext);
void orc_free(void *p);
/* canonical symmetric CSR from an edge list: drop loops, dedup, symmetrise */
int orc_edges_to_csr(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                     uint64_t **off_out, uint32_t **nbr_out);
/* degree relabel: new id order = (degree desc, old id asc) */
int orc_relabel_by_degree(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t *new2old,
                          uint64_t *off_out, uint32_t *nbr_out);
/* orientation nbr < node; off_out has n+1 entries, nbr_out holds the kept entries */
uint64_t orc_orient_lower(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint64_t *off_out,
                          uint32_t *nbr_out);

/* ---- layouts ---- */
int orc_row_type(const uint32_t *data, size_t len, int policy);
size_t orc_uint_build(uint8_t *out, const uint32_t *data, size_t len);
size_t orc_pshort_build(uint8_t *out, const uint32_t *data, size_t len);
size_t orc_bitset_build(uint8_t *out, const uint32_t *data, size_t len);
size_t orc_bitset_new_build(uint8_t *out, const uint32_t *data, size_t len);
size_t orc_new_type_build(uint8_t *out, const uint32_t *data, size_t len);
/* flattened row image as stored by SparseMatrix (hybrid: u64 tag + sub-layout) */
size_t orc_row_image(uint8_t *out, const uint32_t *data, size_t len, int policy);
/* decode a set payload of the given type into uint32 values in the reference's
 * foreach order (sorted, except new_type: its uint part, then its blocks); returns count */
size_t orc_decode(int type, const uint8_t *payload, size_t card, size_t nbytes, uint32_t *out);
/* upper bound on the bytes orc_row_image writes for a row */
size_t orc_row_image_bound(const uint32_t *data, size_t len);
size_t orc_row_image_bound_policy(const uint32_t *data, size_t len, int policy);

/* vertex orderings (0 degree, 1 rev_degree, 2 strong_run, 3 bfs, 4 shingles, 5 random,
 * 6 the_game) as new2old; seeded permutations from orc_perm */
void orc_perm(uint64_t seed, uint64_t stream, uint32_t n, uint32_t *new2old);
int orc_order(uint32_t n, const uint64_t *off, const uint32_t *nbr, int ordering, uint64_t seed, uint32_t *new2old);

/* ---- sets and intersections ---- */
typedef struct {
  uint8_t *data;
  uint64_t cardinality;
  uint64_t number_of_bytes;
  int32_t type;
  int32_t pad;
} orc_set;

size_t orc_isect_uint(int alg, const uint32_t *a, size_t la, const uint32_t *b, size_t lb,
                      uint32_t *out);
/* dispatcher over typed sets; out is NULL for count-only */
uint64_t orc_set_intersect(orc_set *C, const orc_set *A, const orc_set *B);

/* ---- matrix and apps ---- */
typedef struct orc_matrix orc_matrix;
orc_matrix *orc_matrix_build(uint32_t n, const uint64_t *off, const uint32_t *nbr, int lower_only,
                             int policy, int threads);
void orc_matrix_free(orc_matrix *M);
uint32_t orc_matrix_rows(const orc_matrix *M);
uint64_t orc_matrix_cardinality(const orc_matrix *M);
uint32_t orc_matrix_max_row(const orc_matrix *M);
int orc_matrix_get_row(const orc_matrix *M, uint32_t row, orc_set *out);
void orc_matrix_type_counts(const orc_matrix *M, uint64_t counts[3]);
uint64_t orc_matrix_bytes(const orc_matrix *M);

uint64_t orc_triangle_count(const orc_matrix *M, int threads, double *seconds);
uint64_t orc_triangle_count_sample(const orc_matrix *M, int threads, uint32_t stride,
                                   uint32_t phase, uint64_t *edges_done, double *seconds);
uint64_t orc_clique_count(const orc_matrix *M, int k, int threads, double *seconds);
uint64_t orc_clique_count_sample(const orc_matrix *M, int k, int threads, uint32_t stride,
                                 uint32_t phase, uint64_t *edges_done, double *seconds);
uint64_t orc_pair_mix(const orc_matrix *M, uint64_t mix[4]);
uint64_t orc_count_rows(const orc_matrix *M, int k, int threads, uint64_t *per_row);

/* ---- BFS path length (apps/n_path.cpp:27-87) ---- */
/* Level-synchronous BFS over rows [0, n) of a CSR from `start`: each round ORs the
 * frontier's rows into the visited bitset (set_union, src/set/ops/union.hpp:6-127) and
 * takes the new frontier as visited minus the previous visited (set_difference,
 * src/set/ops/difference.hpp:5-84), in id order.  Stops when the frontier is empty or
 * path_length reached max_depth.  Returns path_length; *rounds = frontier expansions
 * run, level_sizes[r] (r < cap) = size of the frontier produced by round r, *reached =
 * visited vertices. */
uint32_t orc_bfs_path(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t start, uint32_t max_depth,
                      uint32_t *rounds, uint64_t *level_sizes, uint32_t cap, uint64_t *reached);
/* the same search with each round's set_union over `threads` threads (per-thread visited
 * bitsets merged after the round, as apps/n_path.cpp:48-77 does); same results */
uint32_t orc_bfs_path_mt(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t start, uint32_t max_depth,
                         uint32_t *rounds, uint64_t *level_sizes, uint32_t cap, uint64_t *reached, int threads);
/* set_difference(C, A, B) of two bitset images [u64 first word][words] into a sorted
 * uint list: the elements of A not in B.  Returns the count. */
size_t orc_bitset_difference(const uint8_t *a, size_t a_bytes, const uint8_t *b, size_t b_bytes, uint32_t *out);

/* ---- independent brute force (symmetric CSR) ---- */
uint64_t orc_brute_triangles(uint32_t n, const uint64_t *off, const uint32_t *nbr);
uint64_t orc_brute_cliques(uint32_t n, const uint64_t *off, const uint32_t *nbr, int k);

/* ---- generators (this project's seeded specification, restated) ---- */
uint64_t orc_mix64(uint64_t x);
uint64_t orc_stream_key(uint64_t seed, uint64_t stream);
uint64_t orc_draw(uint64_t key, uint64_t ctr);
void orc_gen_er(uint32_t n, uint64_t m, uint64_t key, uint32_t *src, uint32_t *dst);
void orc_gen_rmat(uint32_t scale, uint64_t m, const uint32_t thr[3], uint64_t key, uint32_t *src,
                  uint32_t *dst);
void orc_rmat_thresholds(double a, double b, double c, uint32_t thr[3]);
void orc_chung_lu_cdf(uint32_t n, double gamma, uint64_t *cdf);
void orc_gen_chung_lu(uint32_t n, uint64_t m, const uint64_t *cdf, uint64_t key, uint32_t *src,
                      uint32_t *dst);
uint64_t orc_gen_communities(uint32_t n, uint32_t ncomm, uint32_t csize, uint64_t seed,
                             uint32_t *src, uint32_t *dst);
void orc_batch_params(uint64_t seed, uint32_t g, uint32_t *n_out, uint64_t *m_out);

/* ---- learned-selector reference math (fp64) ---- */
void orc_mlp_forward(const float *params, const float *x, int batch, double *out);
double orc_mlp_loss_grad(const float *params, const float *x, const float *target,
                         const float *mask, int batch, double *grad);

#ifdef __cplusplus
}
#endif
#endif
