/*
 * eh.h -- C ABI of the MI355X (gfx950) set-intersection pattern engine.
 *
 * This is the drop-in boundary for EmptyHeaded's hot path: SparseMatrix
 * construction with a per-neighbourhood layout choice, the ops::set_intersect
 * family, and the undirected_triangle_counting / n_clique apps.  Every entry
 * point takes plain pointers and sizes, returns an eh_status, and never throws.
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository root).
 *
 * Threading: a context owns one HIP stream on one device.  Calls on different
 * contexts may run concurrently; calls on one context must be serialised by
 * the caller.  Graphs, matrices and batches belong to the context that made
 * them.
 */
#ifndef EH_H
#define EH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EH_ABI_VERSION 5

typedef enum {
  EH_OK = 0,
  EH_ERR_INVALID = 1,     /* bad argument, malformed CSR, unknown layout or type pair */
  EH_ERR_OOM = 2,         /* device or host allocation failed */
  EH_ERR_HIP = 3,         /* HIP runtime error */
  EH_ERR_IO = 4,          /* file could not be read or written */
  EH_ERR_UNSUPPORTED = 5, /* request outside what this build implements */
  EH_ERR_CAPACITY = 6     /* caller-provided buffer too small */
} eh_status;

/* Set type tags, same values as common::type (src/common.hpp:171-181). */
typedef enum {
  EH_BITSET = 0,
  EH_PSHORT = 1,
  EH_UINTEGER = 2,
  EH_HYBRID = 5,
  EH_BITSET_NEW = 7, /* 128-value blocks (src/set/layouts/bitset_new.hpp:66-108) */
  EH_NEW_TYPE = 8    /* uint part + dense blocks (src/set/layouts/new_type.hpp:50-88) */
} eh_set_type;

/* Matrix layout policy (src/emptyheaded.hpp:45-56 --layout, plus the
 * selectors this engine adds). */
typedef enum {
  EH_LAYOUT_UINT = 0,     /* every row uinteger            (--layout=uint)   */
  EH_LAYOUT_PSHORT = 1,   /* every row pshort              (--layout=pshort) */
  EH_LAYOUT_BITSET = 2,   /* every row bitset              (--layout=bs)     */
  EH_LAYOUT_HYBRID = 3,   /* hybrid::get_type, uint/bitset (--layout=hybrid) */
  EH_LAYOUT_HYBRID3 = 4,  /* hybrid with compute_type's pshort band          */
  EH_LAYOUT_LEARNED = 5,  /* per-row choice by the learned selector          */
  EH_LAYOUT_COST = 6,     /* per-row choice by the byte-cost model, with a
                             bitset_new arm                                   */
  EH_LAYOUT_BITSET_NEW = 7, /* every row bitset_new        (--layout=bitset_new,
                               src/emptyheaded.hpp:63-65)                     */
  EH_LAYOUT_NEW_TYPE = 8,   /* every row new_type          (--layout=new_type,
                               src/emptyheaded.hpp:60-62)                     */
  EH_LAYOUT_HYBRID_BLOCK = 9 /* hybrid::get_type built with NEW_BITSET: dense rows
                                bitset_new (src/set/layouts/hybrid.hpp:83-87)  */
} eh_layout_policy;

/* Edge selection applied by the matrix build (SparseMatrix.hpp:249-266). */
typedef enum {
  EH_FILTER_ALL = 0,   /* keep every entry (edge_selection == true)        */
  EH_FILTER_LOWER = 1, /* keep nbr < node (apps' orientation predicate)    */
  EH_FILTER_MASK = 2   /* keep entries whose byte in edge_mask is nonzero  */
} eh_edge_filter;

/* Execution schedule of the counting kernels. */
typedef enum {
  EH_SCHED_PLAIN = 0, /* reference dispatch: each row staged in its own layout,
                         per-pair kernel by (typeA,typeB) and the size ratio  */
  EH_SCHED_TUNED = 1  /* source rows staged as LDS bitmaps when they fit,
                         autotuned task split and lane grouping            */
} eh_schedule;

typedef struct eh_ctx eh_ctx;
typedef struct eh_graph eh_graph;
typedef struct eh_matrix eh_matrix;
typedef struct eh_batch eh_batch;

/* Host view of one set, in the reference's in-memory formats (Set.hpp:17-21):
 * UINTEGER data = uint32[cardinality]; BITSET data = [u64 first word][u64 words];
 * PSHORT data = u16 stream of [prefix, count-1, low16...] partitions;
 * BITSET_NEW data = [u32 block ids][2 u64 words per block], 20 bytes per block;
 * NEW_TYPE data = [u64 uint bytes][uint32 uint part][BITSET_NEW part]. */
typedef struct {
  const void *data;
  uint64_t cardinality;
  uint64_t number_of_bytes;
  int32_t type; /* eh_set_type */
  int32_t reserved;
} eh_row_view;

typedef struct {
  uint32_t n;            /* vertices */
  uint32_t max_degree;   /* largest row */
  uint64_t entries;      /* stored entries (2x undirected edges for a graph) */
} eh_graph_info;

typedef struct {
  uint32_t n;
  uint32_t max_row;
  uint64_t cardinality;  /* kept entries (oriented edges under EH_FILTER_LOWER) */
  uint64_t rows_by_type[3]; /* non-empty rows typed BITSET, PSHORT, UINTEGER */
  uint64_t payload_bytes;   /* bytes of the B-side payload (uint + pshort + bitset) */
  uint64_t ref_row_bytes;   /* bytes the reference's row images would take */
  uint64_t wedges;          /* W = sum over kept (i,j) of card(j) */
  double build_ms;          /* device time of the build (filter+stats+layouts) */
} eh_matrix_info;

typedef struct {
  int32_t schedule;       /* eh_schedule */
  int32_t lanes_per_row;  /* 0 = autotuned/default; else 4,8,16,32,64 */
  uint32_t task_work;     /* 0 = default; target work units per task */
  uint32_t lds_words;     /* 0 = default; per-wave LDS budget in 32-bit words (512..16384) */
  int32_t waves_per_block;/* 0 = default (tuned 2, plain 4); at most 4, and at most
                             64 KB of LDS slices per block */
  int32_t repeat;         /* run the counting kernel this many times (>=1); stats average */
  int32_t small_degree;   /* TUNED on oriented matrices: rows with at most this many
                             entries are counted edge-centrically; 0 = default, <0 = off */
  int32_t stream_ratio;   /* TUNED: a bitset B row is streamed as its uint list instead of
                             probed when 16 |B| < stream_ratio * probes; 0 = default, <0 = never */
  int32_t bstat_split;    /* TUNED triangles on oriented (EH_FILTER_LOWER) matrices: pairs (i,j)
                             of rows above small_degree whose B row id j is below this bound are
                             counted B-stationary (row(j) staged once as an LDS bitmap, the part
                             of every row(i) below j streamed through it); 0 = default (65536),
                             <0 = off, else 64..65536 */
  int32_t local_max;      /* TUNED 4-cliques on oriented matrices: rows with more than small_degree
                             and at most this many entries count the triangles of their local
                             graph (bit-rows in LDS); 0 = default (512), <0 = off, at most 512 */
  int32_t time_parts;     /* nonzero: record HIP events around each kernel class of the pass and
                             report their device times in eh_run_stats.part_ms */
  int32_t bsh_max;        /* TUNED triangles with the split on: a row whose B rows at or above the
                             split all have at most this many entries counts those pairs
                             B-stationary too (k_bshl); 0 = default (512), <0 = off, else 1..512 */
  int32_t short_queues;   /* TUNED on batch arenas: the edge-centric short rows are cut into 8
                             ranges of equal wedge work, one queue per XCD, taken first by the
                             blocks of that XCD (each L2 then mostly holds its own graphs);
                             0 = default (on), <0 = one queue */
} eh_run_opts;

/* Kernel classes of one counting pass (indices of eh_run_stats.part_ms). */
enum {
  EH_PART_SHORT_ROWS = 0, /* edge-centric short rows (k_small_rows, k_small_cliques4)        */
  EH_PART_BSTAT = 1,      /* B-stationary pairs below the split (k_bstat)                      */
  EH_PART_BSTAT_ABOVE = 2,/* B-stationary pairs whose B row is at or above the split (k_bshl)   */
  EH_PART_LOCAL_MID = 3,  /* 4-cliques of rows up to 128 entries by their local graph (k_local4) */
  EH_PART_LOCAL_BIG = 4,  /* the same for rows of 129..512 entries                             */
  EH_PART_TASKS_BIG = 5,  /* A-side tasks of the large-LDS class (k_count)                    */
  EH_PART_TASKS = 6,      /* A-side tasks (k_count); also the whole plain schedule             */
  EH_PARTS = 8
};

typedef struct {
  double kernel_ms;       /* average device time of one counting pass (HIP events) */
  double sched_ms;        /* device time to build the task schedule (0 if cached) */
  uint64_t tasks;
  uint64_t alg_bytes;     /* algorithmic bytes of one pass on the canonical uint32 CSR: triangles
                             B_TC = 4 (m + W) + 8 (n + 1), m kept entries, W = sum over kept (i,j)
                             of card(j); 4-cliques B_4C = B_TC + 4 sum over triangles i>j>k of
                             card(k); 0 for deeper cliques */
  uint64_t pairs[4];      /* (i,j) pairs by B-side type: U, PS, BS, empty */
  uint64_t work_units;    /* elements + words touched on the B side */
  int32_t schedule;
  int32_t lanes_per_row;
  uint32_t task_work;
  uint32_t lds_words;
  double part_ms[EH_PARTS]; /* with eh_run_opts.time_parts: average device time of each kernel
                               class (EH_PART_*) of a pass, 0 for classes the pass did not run */
} eh_run_stats;

/* Generator specification (this engine's seeded synthetic graphs). */
typedef enum {
  EH_GEN_ER = 0,       /* m pair draws (loops and repeats dropped by the canonicaliser) */
  EH_GEN_RMAT = 1,
  EH_GEN_CHUNG_LU = 2,
  EH_GEN_ER_EXACT = 3  /* G(n, m): exactly m distinct non-loop pairs, the first m distinct
                          pairs of EH_GEN_ER's draw stream */
} eh_gen_kind;
typedef struct {
  int32_t kind;          /* eh_gen_kind */
  uint32_t n;            /* vertices (ER, Chung-Lu) */
  uint32_t scale;        /* RMAT: n = 2^scale */
  uint32_t edge_factor;  /* RMAT: draws = edge_factor * n */
  uint64_t m;            /* ER: pair draws; ER_EXACT: distinct pairs; Chung-Lu: endpoint-pair draws */
  uint64_t seed;
  double a, b, c;        /* RMAT quadrant probabilities (d = 1-a-b-c) */
  double gamma;          /* Chung-Lu degree exponent */
  uint32_t communities;  /* Chung-Lu: planted communities */
  uint32_t community_size;
  int32_t relabel;       /* 1 = degree relabel (degree desc, id asc) */
  int32_t reserved;
} eh_gen_spec;

/* ---------------------------------------------------------------- context */
int eh_abi_version(void);
const char *eh_status_str(eh_status s);
const char *eh_last_error(void);            /* thread-local detail of the last failure */
/* No reference counterpart: the reference has no device (src/emptyheaded.hpp:33-82 runs on the host). */
eh_status eh_ctx_create(int device, eh_ctx **out);
void eh_ctx_destroy(eh_ctx *ctx);
eh_status eh_ctx_device_name(eh_ctx *ctx, char *buf, size_t cap);

/* ---------------------------------------------------------------- graphs */
/* MutableGraph::undirectedFromBinary (src/MutableGraph.hpp:316-360): host
 * symmetric CSR (rows sorted, no self-loops) uploaded to the device. */
eh_status eh_graph_from_csr(eh_ctx *ctx, uint32_t n, const uint64_t *row_off, const uint32_t *nbrs,
                            eh_graph **out);
/* Edge list -> canonical symmetric graph (loops dropped, duplicates merged). */
eh_status eh_graph_from_edges(eh_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *src,
                              const uint32_t *dst, int relabel, eh_graph **out);
/* Seeded device generator + canonicalisation (+ degree relabel). */
eh_status eh_graph_generate(eh_ctx *ctx, const eh_gen_spec *spec, eh_graph **out);
/* MutableGraph::reorder_by_degree (src/MutableGraph.hpp:20-32,100-123,268-272). */
eh_status eh_graph_relabel_by_degree(eh_graph *g);
/* The orderings of tools/undirectedEdgeListToBinary.cpp:26-32 (MutableGraph::reorder_*,
 * src/MutableGraph.hpp:126-285) as device passes.  The reference draws its shuffles
 * with std::random_shuffle; here `seed` drives counter-based permutations, so BFS,
 * SHINGLES, RANDOM and THE_GAME are reproducible.  Ties the reference leaves to an
 * unstable sort keep id order; STRONG_RUN places vertices that appear in no row last. */
typedef enum {
  EH_ORDER_DEGREE = 0,     /* reorder_by_degree: degree desc, id asc       */
  EH_ORDER_REV_DEGREE = 1, /* reorder_by_rev_degree: degree asc            */
  EH_ORDER_STRONG_RUN = 2, /* reorder_strong_run                           */
  EH_ORDER_BFS = 3,        /* reorder_bfs                                  */
  EH_ORDER_SHINGLES = 4,   /* reorder_by_shingles                          */
  EH_ORDER_RANDOM = 5,     /* reorder_random                               */
  EH_ORDER_THE_GAME = 6    /* reorder_by_the_game: BFS, then degree        */
} eh_ordering;
eh_status eh_graph_reorder(eh_graph *g, eh_ordering ordering, uint64_t seed);
eh_status eh_graph_info_get(const eh_graph *g, eh_graph_info *out);
/* copy the symmetric CSR to host (row_off: n+1 entries, nbrs: entries) */
eh_status eh_graph_download(const eh_graph *g, uint64_t *row_off, uint32_t *nbrs);
/* device new->old id map after a relabel (n entries; identity otherwise) */
eh_status eh_graph_download_perm(const eh_graph *g, uint32_t *new2old);
void eh_graph_free(eh_graph *g);

/* ---------------------------------------------------------------- matrices */
/* SparseMatrix<T,R>::build (src/SparseMatrix.hpp:90-93,268-290,406-494) from a
 * device graph; edge_mask is only read for EH_FILTER_MASK (one byte per stored
 * entry, host memory). */
eh_status eh_matrix_build(eh_graph *g, eh_edge_filter filter, const uint8_t *edge_mask,
                          eh_layout_policy policy, eh_matrix **out);
/* Same from a host symmetric CSR (upload + build). */
eh_status eh_matrix_build_host(eh_ctx *ctx, uint32_t n, const uint64_t *row_off, const uint32_t *nbrs,
                               eh_edge_filter filter, const uint8_t *edge_mask,
                               eh_layout_policy policy, eh_matrix **out);
eh_status eh_matrix_info_get(eh_matrix *m, eh_matrix_info *out);
/* SparseMatrix::get_row (src/SparseMatrix.hpp:156-160): copies row's set, in
 * its reference format, into buf (cap bytes); view->data points into buf. */
eh_status eh_matrix_row(eh_matrix *m, uint32_t row, void *buf, size_t cap, eh_row_view *view);
/* kept entries as a host CSR (row_off: n+1, nbrs: cardinality) */
eh_status eh_matrix_download(eh_matrix *m, uint64_t *row_off, uint32_t *nbrs);
/* per-row B-side layout types (n bytes, eh_set_type values; 3 is the uint layout
 * searched by galloping) */
eh_status eh_matrix_row_types(eh_matrix *m, uint8_t *types);
void eh_matrix_free(eh_matrix *m);

/* ---------------------------------------------------------------- set ops */
/* ops::set_intersect(Set<X>* C, const Set<Y>* A, const Set<Z>* B)
 * (src/set/ops/intersection.hpp:829,880,956,1019,1023,1059,1062,1321,1325,
 * 1364,1476,1507,1555,2067): intersect two host sets on the device.
 * materialize=0 only counts (out may be NULL); otherwise the result is written
 * to out_buf in the reference's result layout (U for pairs with a uint side,
 * BS for BSxBS, PS for PSxPS and PSxBS, BITSET_NEW for BITSET_NEW x BITSET_NEW,
 * NEW_TYPE for NEW_TYPE x NEW_TYPE) and described by *out.  Block sets pair only
 * as the reference's overloads do: BITSET_NEW with BITSET_NEW or UINTEGER,
 * NEW_TYPE with NEW_TYPE. */
eh_status eh_intersect(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, int materialize,
                       void *out_buf, size_t out_cap, eh_row_view *out);
/* ops::set_union(Set<bitset>* A, Set<X>* B) (src/set/ops/union.hpp:6-127): *out = A's
 * bitset with B's elements added, in A's word span (the reference writes in place; pass
 * A's own buffer as out_buf to do the same).  B is BITSET (OR over the common words),
 * UINTEGER or PSHORT; B's elements outside A's span are dropped.  out->cardinality is
 * the result's element count. */
eh_status eh_union(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, void *out_buf, size_t out_cap,
                   eh_row_view *out);
/* ops::set_difference(Set<uinteger>* C, const Set<bitset>* A, const Set<bitset>* B)
 * (src/set/ops/difference.hpp:5-84): the elements of A not in B, as a sorted uint set. */
eh_status eh_difference(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, void *out_buf, size_t out_cap,
                        eh_row_view *out);
/* Many pairs of rows of one matrix at once: counts[p] = |row(a[p]) & row(b[p])|
 * (the per-pair loop of apps/undirected_triangle_counting.cpp:45-56 and
 * apps/similar_nodes.cpp:45-60 as one device call; host arrays in and out). */
eh_status eh_intersect_rows(eh_matrix *m, const uint32_t *a_rows, const uint32_t *b_rows,
                            uint64_t npairs, uint64_t *counts);
/* The nested ops::set_intersect calls of apps run under par_for_range -- row(i) & row(j)
 * (apps/undirected_triangle_counting.cpp:45-52) and (row(i) & row(j)) & row(k) ...
 * (apps/n_clique.cpp:30-50) -- answered many at once: request q intersects the rows
 * row_ids[row_off[q] .. row_off[q+1]) of one matrix (1 to EH_ROW_SET_MAX rows; row_off has
 * nreq + 1 entries, row_off[0] = 0) and counts[q] receives the result's size.  elem_off (may
 * be NULL) receives nreq + 1 running sums of the counts; with elems non-NULL the sorted
 * elements of request q are written to elems[elem_off[q] ..), and a total above elems_cap
 * fails with EH_ERR_CAPACITY after counts and elem_off are filled. */
#define EH_ROW_SET_MAX 16
eh_status eh_intersect_row_sets(eh_matrix *m, const uint32_t *row_ids, const uint64_t *row_off, uint64_t nreq,
                                uint64_t *counts, uint32_t *elems, uint64_t elems_cap, uint64_t *elem_off);
/* No counterpart in the reference.  eh_intersect_rows for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): same arguments and results, but the pairs and counts are staged in a static
 * device array of EH_STATIC_MAX_PAIRS pairs (16 bytes a pair: 16 MiB on every GPU the
 * library runs on, held until the process ends) instead of buffers allocated per call.
 * npairs above EH_STATIC_MAX_PAIRS is rejected with EH_ERR_CAPACITY.  All calls on one GPU
 * share that array: calls that overlap on one GPU (different contexts or threads) race,
 * and the caller must serialise them. */
#define EH_STATIC_MAX_PAIRS 1048576ull
eh_status eh_intersect_rows_static(eh_matrix *m, const uint32_t *a_rows, const uint32_t *b_rows,
                                   uint64_t npairs, uint64_t *counts);

/* ---------------------------------------------------------------- apps */
/* undirected_triangle_counting<T,R>::run (apps/undirected_triangle_counting.cpp:27-65):
 * sum over kept (i,j) of |row(i) & row(j)|. */
eh_status eh_triangle_count(eh_matrix *m, const eh_run_opts *opts, uint64_t *count,
                            eh_run_stats *stats);
/* n_clique<T,R>::run (apps/n_clique.cpp:52-101), query_depth = k, 3 <= k <= 16:
 * k = 3, 4 run the tuned/plain schedules of eh_run_opts; k >= 5 runs the general
 * level-set recursion (opts are not used). */
eh_status eh_clique_count(eh_matrix *m, uint32_t k, const eh_run_opts *opts, uint64_t *count,
                          eh_run_stats *stats);

/* n_path<T,R>::run (apps/n_path.cpp:27-87): breadth-first search from `start` over
 * the rows [0, n_vertices) of a matrix built with EH_FILTER_ALL (0: all rows; a
 * directed matrix's out-rows are rows [0, n)).  Each round ORs the frontier's rows
 * into the visited set (set_union, src/set/ops/union.hpp:6-127) and takes the new
 * frontier as the vertices not visited before (set_difference,
 * src/set/ops/difference.hpp:5-84); the search stops when the frontier is empty or
 * path_length reached max_depth (query_depth; n_path's default is 0xEFFFFFF).
 * level_sizes[r] (r < level_cap, may be NULL with level_cap 0) receives the size of the
 * frontier round r produced. */
typedef enum {
  EH_BFS_TOP_DOWN = 0,  /* every round expands the frontier's rows */
  EH_BFS_SYMMETRIC = 1, /* rows are symmetric: wide rounds may run bottom-up over the rows */
  EH_BFS_IN_ROWS = 2    /* rows [n_vertices, 2 n_vertices) hold in-neighbourhoods (a directed
                           matrix as the C++ surface builds it): wide rounds run bottom-up over them */
} eh_bfs_mode;
typedef struct {
  uint32_t path_length;  /* n_path::path_length */
  uint32_t rounds;       /* frontier expansions run */
  uint64_t reached;      /* visited vertices, the start included */
  double kernel_ms;      /* device time of the search */
  uint32_t bottom_up;    /* rounds run bottom-up */
  uint32_t reserved;
} eh_bfs_info;
/* mode: how wide rounds may be run; every mode gives the same frontiers, path length and
 * level sizes (a round's new frontier is the set of unvisited vertices adjacent to the
 * current one, whichever side enumerates the adjacency).  EH_BFS_SYMMETRIC and
 * EH_BFS_IN_ROWS need a matrix built with EH_FILTER_ALL (EH_ERR_INVALID otherwise). */
eh_status eh_bfs_path(eh_matrix *m, uint32_t n_vertices, uint32_t start, uint32_t max_depth, eh_bfs_mode mode,
                      uint64_t *level_sizes, uint32_t level_cap, eh_bfs_info *info);
/* No reference counterpart (the reference times whole phases only,
 * src/common.hpp:87-99): one counting pass with the shader-clock cycles of
 * every task recorded; task_rows/task_kinds/task_cycles get ntasks entries
 * (cap is their capacity).  The engine-generated timings the selector trains on. */
eh_status eh_count_profile(eh_matrix *m, uint32_t k, const eh_run_opts *opts, uint32_t *task_rows,
                           uint32_t *task_kinds, uint64_t *task_cycles, uint64_t cap, uint64_t *ntasks);

/* ---------------------------------------------------------------- batches */
/* No reference counterpart (the reference processes one graph per process,
 * src/emptyheaded.hpp:33-82): many independent graphs in one arena. */
eh_status eh_batch_generate(eh_ctx *ctx, uint64_t seed, uint32_t ngraphs, eh_layout_policy policy,
                            eh_batch **out);
eh_status eh_batch_info(eh_batch *b, uint32_t *ngraphs, uint64_t *total_vertices,
                        uint64_t *total_edges);
/* per-graph vertex counts and edge-draw counts used by the generator */
eh_status eh_batch_graph_params(eh_batch *b, uint32_t *n, uint64_t *m_draws);
/* graph g's canonical symmetric CSR, local ids (row_off: n_g+1) */
eh_status eh_batch_graph_download(eh_batch *b, uint32_t g, uint64_t *row_off, uint32_t *nbrs,
                                  uint64_t cap);
/* mode 0: one persistent launch over every graph, replayed from a captured
 * hipGraph; mode 1: one launch sequence per graph.  tri/k4 get ngraphs counts
 * each (either may be NULL: that pass is not run).  tri_opts / k4_opts set the
 * schedule and launch shape of the triangle and 4-clique passes (NULL: defaults);
 * the graph is recorded again whenever either pass changes.  When both passes are tuned
 * and every row of both is a short row (one small_degree for both), the 4-clique kernel
 * counts the triangles too and the triangle pass is not launched; the counts are those of
 * the two passes run apart. */
eh_status eh_batch_count(eh_batch *b, int mode, const eh_run_opts *tri_opts,
                         const eh_run_opts *k4_opts, uint64_t *tri, uint64_t *k4, double *ms);
/* No reference counterpart: build the arena's layouts again from the same graphs, under
 * `policy`, or with the per-row arms row_types (total_vertices entries, as
 * eh_matrix_build_typed takes them; policy is then ignored).  The schedules, the
 * captured graph and the per-graph plans are built again on the next eh_batch_count. */
eh_status eh_batch_rebuild(eh_batch *b, eh_layout_policy policy, const uint8_t *row_types);
/* No reference counterpart: the arena's oriented matrix (rows are the union graph's
 * vertices, every graph's a contiguous range), for eh_matrix_info_get, eh_matrix_row_types,
 * eh_matrix_features and eh_matrix_download.  Borrowed: valid until eh_batch_rebuild or
 * eh_batch_free, and never passed to eh_matrix_free. */
eh_status eh_batch_matrix(eh_batch *b, eh_matrix **out);
/* No reference counterpart: the algorithmic bytes of the batch's passes, summed over its
 * graphs (B_TC of the triangle pass, B_4C of the 4-clique pass, as in eh_run_stats). */
eh_status eh_batch_alg_bytes(eh_batch *b, uint64_t *tri_bytes, uint64_t *k4_bytes);
void eh_batch_free(eh_batch *b);

/* ---------------------------------------------------------------- learned selector */
/* No reference counterpart: the reference's selector is the fixed density rule
 * (src/set/layouts/hybrid.hpp:73-100).  FP32 MLP 16->64->64->4, ReLU. */
#define EH_MLP_IN 16
#define EH_MLP_HIDDEN 64
#define EH_MLP_OUT 4
#define EH_MLP_PARAMS 5508
eh_status eh_selector_set_params(eh_ctx *ctx, const float *params);
eh_status eh_selector_get_params(eh_ctx *ctx, float *params);
/* Per-row selector arms (eh_matrix_row_types values) and what the tuned triangle pass runs
 * for each when the row is the B side: 0 bitset (queries probe a bit; pairs probe B's words
 * with A's ids or AND them with A's bitmap), 1 pshort (queries walk the partitions; pairs
 * stream B's list), 2 uint (queries binary-search; pairs stream B's list), 3 uint searched
 * (queries binary-search; pairs search each of A's ids within B's first and last id in B's
 * list: the per-pair algorithm for long B rows few of A's ids fall in). */
/* EH_LAYOUT_LEARNED's rule: a bitset only where its words are at most guard_words per id
 * (default 8: density 1/512, 16x the uint row's bytes); fixed_arm -1 takes the network's arm (default), 0..3 gives
 * that arm to every row whose arm can matter (the ablations the network is measured against).
 * Rows of fewer than two ids, or held by no row, are uint under every rule.  No counterpart
 * in the reference. */
eh_status eh_selector_set_rule(eh_ctx *ctx, uint32_t guard_words, int32_t fixed_arm);
/* The tuned triangle schedule EH_LAYOUT_LEARNED selects for (eh_run_opts.bstat_split and
 * bsh_max of the pass that will read the matrix; 0 = their defaults, 65536 and 512).  A row's
 * arm only changes the pairs and short-row queries whose B row it is and that the pass runs
 * A-side or edge-centric: with the split on, the pairs at or above the split of the rows that
 * hold a B row there of more than bsh_max ids (the others are B-stationary and read canonical
 * lists).  Only rows some such pair reads are candidates, and a row's "uses" feature counts
 * those rows; the others are uint.  bstat_split < 0: every pair is A-side (every row held by
 * some row is a candidate), as with bstat_split < 0 in the pass; bsh_max < 0: no row's pairs
 * above the split are B-stationary.  Unoriented matrices and batch arenas always count every
 * use.  No counterpart in the reference. */
eh_status eh_selector_set_schedule(eh_ctx *ctx, int32_t bstat_split, int32_t bsh_max);
/* per-row features (n x 16 floats) of an oriented matrix's rows; in a batch arena each row's
 * ids, index and n are its own graph's, and feature 15 is 0 (1 for a single graph) */
eh_status eh_matrix_features(eh_matrix *m, float *features);
/* A matrix whose rows take the caller's arms (row_types: n values 0..3; a bitset only within
 * the guard of eh_selector_set_rule, rows of fewer than two ids uint): the selector's labels
 * time the tuned pass over such matrices.  No counterpart in the reference. */
eh_status eh_matrix_build_typed(eh_graph *g, eh_edge_filter filter, const uint8_t *row_types, eh_matrix **out);
/* device forward over a host batch (x: batch x 16, y: batch x 4) */
eh_status eh_mlp_forward(eh_ctx *ctx, const float *params, const float *x, int batch, float *y);
/* device loss + gradient (grad: 5508 floats) */
eh_status eh_mlp_loss_grad(eh_ctx *ctx, const float *params, const float *x, const float *target,
                           const float *mask, int batch, float *loss, float *grad);
/* device Adam training of the context's selector on host samples */
eh_status eh_selector_train(eh_ctx *ctx, const float *x, const float *target, const float *mask,
                            int nsamples, int epochs, int batch, float lr, uint64_t seed,
                            float *final_loss);

#ifdef __cplusplus
}
#endif
#endif
