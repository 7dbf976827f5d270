// eh_internal.hpp -- shared declarations of the gfx950 set-intersection engine.
//
// Device layout of a built matrix (one contiguous set of HBM buffers per matrix):
//   card[n]      u32   kept entries per row (d+ under the nbr<node filter)
//   aoff[n]      u32   row start in `unbr`, in units of 4 elements (16-byte aligned rows)
//   unbr[]       u32   canonical sorted uint32 rows, each padded to a multiple of 4
//                      with 0xFFFFFFFF (never a vertex id)
//   desc[n]      RowDesc (16 B): the B-side payload of the row in its chosen layout
//   bsw[]        u64   bitset words (reference bitset: word index = value >> 6)
//   psd[]        u16   pshort streams [prefix, count-1, low16 x count] per partition,
//                      each row 8-byte aligned
//   bid[], bwd[] u32 block id (value >> 7) and the block's two u64 words (uint4),
//                      one slot per 128-value block of a bitset_new / new_type row;
//                      a new_type row's slots start with two header slots
//                      [uint part count, uint part offset in ntu / 4]
//   ntu[]        u32   new_type uint parts, padded to 4 with 0xFFFFFFFF
// The uint rows are always present (they drive the j-enumeration of a source
// row); the pshort/bitset/block payloads exist for rows whose selected layout is
// that type.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/eh.h"

#define EH_U 2u
#define EH_PS 1u
#define EH_BS 0u
#define EH_UG 3u  // uint layout, membership by galloping (exponential) search from the front
#define EH_BN 4u  // bitset_new: 128-value blocks (src/set/layouts/bitset_new.hpp:66-108)
#define EH_NT 5u  // new_type: uint part + dense blocks (src/set/layouts/new_type.hpp:50-88)
#define EH_NT_HEADER 2u  // header slots before a new_type row's blocks
#define EH_PAD_VALUE 0xFFFFFFFFu
// common::type values of the block layouts (src/common.hpp:179-180)
#define EH_SET_BITSET_NEW 7
#define EH_SET_NEW_TYPE 8

struct RowDesc {
  uint32_t card;   // element count
  uint32_t tinfo;  // bits 0-2: type (0 BS, 1 PS, 2 U, 3 U searched by galloping,
                   // 4 bitset_new, 5 new_type);
                   // bits 3-31: BS words / PS stream shorts / blocks
  uint32_t w0;     // BS: first u64 word index; PS: number of partitions;
                   // BN, NT: first slot in bid/bwd
  uint32_t off;    // U, UG, BN, NT: canonical list offset/4; BS: word offset; PS: short offset/4
};
static_assert(sizeof(RowDesc) == 16, "RowDesc must be 16 bytes");

__host__ __device__ inline uint32_t desc_type(const RowDesc &d) { return d.tinfo & 7u; }
// rows whose canonical uint list sits at `off`: the two uint arms (that list is
// their payload) and the block layouts (kept beside their blocks)
__host__ __device__ inline bool is_uint_type(uint32_t t) { return t >= EH_U; }
__host__ __device__ inline bool is_block_type(uint32_t t) { return t >= EH_BN; }
__host__ __device__ inline uint32_t desc_aux(const RowDesc &d) { return d.tinfo >> 3; }
__host__ __device__ inline uint32_t make_tinfo(uint32_t t, uint32_t aux) { return t | (aux << 3); }

// ---------------------------------------------------------------- errors

void eh_set_error(const char *fmt, ...);

#define EH_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      eh_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? EH_ERR_OOM : EH_ERR_HIP;                     \
    }                                                                                 \
  } while (0)

#define EH_CHECK(expr)                  \
  do {                                  \
    eh_status s_ = (expr);              \
    if (s_ != EH_OK) return s_;         \
  } while (0)

// ---------------------------------------------------------------- device buffers

// A context's cache of device blocks for the temporaries of its builds and schedules: a
// released block goes back here instead of to hipFree (which waits for the whole device), and
// a later request it fits takes it again.  Every use of a context's memory is ordered on the
// context's stream (a blocking stream: copies on the null stream wait for it), so a block
// handed out again is only touched by work queued after the work that used it before.  Blocks
// are sized in 256-byte steps below 1 MiB and 1 MiB steps above; a request takes the smallest
// cached block of at least its size if that is at most twice it.  Beyond CAP bytes the largest
// cached blocks are freed.
struct BlockCache {
  static constexpr size_t CAP = 8ull << 30;
  std::mutex mu;
  std::multimap<size_t, void *> blocks;  // by size
  size_t cached = 0;
  static size_t round(size_t b) {
    const size_t q = b < (1u << 20) ? 256u : (1u << 20);
    return (b + q - 1) / q * q;
  }
  void *get(size_t b, size_t *got) {
    std::lock_guard<std::mutex> g(mu);
    auto it = blocks.lower_bound(b);
    if (it == blocks.end() || it->first > 2 * b) return nullptr;
    void *p = it->second;
    *got = it->first;
    cached -= it->first;
    blocks.erase(it);
    return p;
  }
  void put(void *p, size_t b) {
    std::lock_guard<std::mutex> g(mu);
    blocks.emplace(b, p);
    cached += b;
    while (cached > CAP && !blocks.empty()) {
      auto last = std::prev(blocks.end());
      (void)hipFree(last->second);
      cached -= last->first;
      blocks.erase(last);
    }
  }
  void clear() {
    std::lock_guard<std::mutex> g(mu);
    for (auto &kv : blocks) (void)hipFree(kv.second);
    blocks.clear();
    cached = 0;
  }
};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  BlockCache *cache = nullptr;  // the block returns here on release (a context's temporaries)
  size_t block = 0;             // its cached size
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), cache(o.cache), block(o.block) {
    o.p = nullptr;
    o.bytes = 0;
    o.cache = nullptr;
    o.block = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      cache = o.cache;
      block = o.block;
      o.p = nullptr;
      o.bytes = 0;
      o.cache = nullptr;
      o.block = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      if (cache) cache->put(p, block);
      else (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    cache = nullptr;
    block = 0;
  }
  // c: take the block from (and later return it to) a context's cache
  eh_status alloc(size_t b, BlockCache *c = nullptr) {
    release();
    if (b == 0) b = 16;
    if (c) {
      size_t got = 0;
      if ((p = c->get(b, &got)) != nullptr) {
        bytes = b;
        cache = c;
        block = got;
        return EH_OK;
      }
      const size_t r = BlockCache::round(b);
      if (hipMalloc(&p, r) == hipSuccess) {
        bytes = b;
        cache = c;
        block = r;
        return EH_OK;
      }
      (void)hipGetLastError();
      // out of memory with blocks cached: give them back and try once more
      c->clear();
    }
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      eh_set_error("hipMalloc(%zu): %s", b, hipGetErrorString(e));
      return EH_ERR_OOM;
    }
    bytes = b;
    return EH_OK;
  }
  template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// ---------------------------------------------------------------- handles

struct TuneParams {
  int lanes = 64;          // lanes cooperating on one B row
  uint32_t task_work = 0;  // work units per task (0: derived)
  uint32_t lds_small = 2048;
  uint32_t lds_big = 16384;
  int waves_per_block = 4;
  int big_waves = 1;       // waves per block of the large-LDS task class
  uint32_t small_deg = 0;  // rows up to this many entries use the edge-centric kernel
  uint32_t stream_ratio = 0;  // tuned: stream bitset B rows when 16 |B| < ratio * probes (0: never)
  uint32_t bstat_j = 0;       // tuned triangles: pairs with B row id below this run B-stationary (0: off)
  uint32_t bsh_max = 0;       // and rows whose B rows above it have at most this many entries run
                              // those pairs B-stationary, k_bshl (0: off)
  uint32_t grab = 2;          // tasks a wave takes from the queue per atomic
  uint32_t local_max = 0;     // 4-cliques: rows above small_deg and up to this run k_local4 (0: off)
  bool short_queues = true;   // batch arenas: short rows from the per-XCD queues (Schedule::xq)
};

// device-time marks around one kernel class of a counting pass
struct PartEv {
  hipEvent_t b = nullptr, e = nullptr;
  bool used = false;
};

struct eh_ctx {
  int device = 0;
  PartEv parts[EH_PARTS];  // created on the first pass that asks for per-class times
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int num_cus = 0;
  std::string name;
  std::vector<float> mlp;  // selector parameters (EH_MLP_PARAMS) or empty
  DevBuf mlp_dev;          // their device copy for the selection kernel
  uint32_t sel_guard = 8;  // LEARNED: a bitset only where its words are at most this many per id
  int sel_fixed = -1;      // LEARNED: -1 the network's arm, 0..3 that arm for every row (ablations)
  // LEARNED: the tuned triangle schedule the selection serves (the B-stationary split and the
  // k_bshl row bound; split 0: every pair A-side): only rows some A-side or edge-centric pair of
  // that schedule reads are candidates
  uint32_t sel_split = 65536, sel_bsh = 512;
  BlockCache cache;        // device blocks of the builds' and schedules' temporaries
  DevBuf scratch;          // reusable temp storage for primitives
  DevBuf counter;          // small device counters
  DevBuf pass_words;       // a counting pass's queue heads, error word and count (kept between
                           // passes: a per-pass hipMalloc / hipFree pair synchronised the device)
};

struct eh_graph {
  eh_ctx *ctx = nullptr;
  uint32_t n = 0;
  uint64_t entries = 0;  // stored entries (2x undirected edges)
  uint32_t max_degree = 0;
  DevBuf off;      // u64[n+1]
  DevBuf nbr;      // u32[entries]
  DevBuf new2old;  // u32[n] (empty => identity)
  bool symmetric = false;  // built by canonicalisation: every edge stored both ways
};

struct Schedule {
  bool valid = false;
  int sched = -1;
  uint32_t task_work = 0;
  uint32_t lds_small = 0, lds_big = 0;
  DevBuf tasks;  // uint4 {i, jb, je, graph}
  uint64_t n_small = 0, n_big = 0;  // tasks [0,n_big) need the big LDS budget
  double ms = 0;
  uint64_t stats[6] = {0, 0, 0, 0, 0, 0};  // pairs by B type (U, PS, BS, empty), B units, wedges
  DevBuf small_rows;        // rows counted by the edge-centric kernel
  uint64_t n_small_rows = 0;
  uint32_t small_deg = 0;
  // B-stationary part (tuned triangles): in_pairs holds, per B row j < bstat_j, the
  // (row offset / 4, position of j) of every task row containing j; btasks are
  // {j, first pair, end pair, 0}
  uint32_t bstat_j = 0, bstat_req = 0;  // in effect / as asked (0 where it does not apply)
  uint32_t bsh_req = 0;                 // the k_bshl row bound the schedule was built for
  DevBuf in_pairs, btasks, jstart;  // jstart: per row, entries counted B-stationary
  DevBuf lo16;                      // those entries as 16-bit ids, 8 per 16-byte vector
  uint32_t local_req = 0;           // 4-cliques: the local_max the schedule was built for
  // B-stationary part above the split (k_bshl): B rows at or above the split with in-pairs (row order), the
  // offset of each one's in-pairs, the first row of each bin, the pairs and the tasks
  DevBuf hj, hjlo, hoff, bstart, hpairs, htasks;  // hjlo: each such B row's ids below the split
  uint64_t n_htasks = 0;
  uint32_t bsh_longest = 0;         // the longest of those B rows (ids)
  DevBuf local_rows[2];             // rows of the two k_local4 classes (<= 128, <= 512 entries)
  // batch arenas: small_rows cut into 8 ranges of equal wedge work, xq[0..8] their bounds and
  // the queue heads of a pass, each on a line of its own (one queue per XCD; empty for a single graph)
  DevBuf xq;
  uint64_t n_local[2] = {0, 0};
  uint64_t n_btasks = 0, n_in_pairs = 0;
};

struct eh_matrix {
  eh_ctx *ctx = nullptr;
  uint32_t n = 0;
  uint64_t card = 0;
  uint32_t max_row = 0;
  int policy = 0;
  int filter = 0;
  uint64_t unbr_words = 0;  // u32 entries in unbr (padded)
  uint64_t bs_words = 0;
  uint64_t ps_shorts = 0;
  uint64_t blk_slots = 0;   // bid/bwd slots
  uint64_t ntu_words = 0;   // u32 entries in ntu
  uint64_t rows_by_type[3] = {0, 0, 0};
  uint64_t ref_row_bytes = 0;
  uint64_t wedges = 0;
  double build_ms = 0;
  bool tw_valid = false;    // tri_weight computed
  uint64_t tri_weight = 0;  // sum over triangles i > j > x of card(x) (4-clique byte model)
  DevBuf dcard, daoff, dunbr, ddesc, dbsw, dpsd, dbid, dbwd, dntu;
  DevBuf dgraph_of;  // u32 per row: graph id (batch arenas); empty => all 0
  Schedule sched[2][2];  // [triangles, 4-cliques][plain, tuned]
};

// ---------------------------------------------------------------- primitives (eh_prim.hip)

// renumber a graph by a device new2old permutation (eh_graph.hip)
eh_status graph_apply_order(eh_graph *g, const uint32_t *d_new2old);

// k-cliques for k >= 3 by the general level-set recursion (eh_clique.hip)
eh_status clique_count_deep(eh_matrix *m, uint32_t k, uint64_t *count, double *ms);

// keys sorted on bits [0, end_bit) (always from bit 0: rocPRIM's merge-sort path, taken for
// inputs up to its merge-sort limit, builds its digit mask as (1 << (begin_bit + bits)) - 1, a
// shift by 64 for a range ending at bit 64 with begin_bit > 0, and then orders by the wrong bits)
eh_status prim_sort_u64(eh_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int end_bit);
eh_status prim_sort_u64_range(eh_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int begin_bit, int end_bit);
eh_status prim_sort_pairs_u32(eh_ctx *ctx, uint32_t *keys, uint32_t *keys_tmp, uint32_t *vals,
                              uint32_t *vals_tmp, uint64_t n);
eh_status prim_sort_pairs_u64_u32(eh_ctx *ctx, uint64_t *keys, uint64_t *keys_tmp, uint32_t *vals,
                                  uint32_t *vals_tmp, uint64_t n);
eh_status prim_sort_u32(eh_ctx *ctx, uint32_t *keys, uint32_t *tmp, uint64_t n);
eh_status prim_unique_u64(eh_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n,
                          uint64_t *n_out);
eh_status prim_exscan_u64(eh_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n);
eh_status prim_exscan_u32_to_u64(eh_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n);
eh_status prim_reduce_max_u32(eh_ctx *ctx, const uint32_t *in, uint64_t n, uint32_t *out);
eh_status prim_reduce_sum_u64(eh_ctx *ctx, const uint64_t *in, uint64_t n, uint64_t *out);

// ---------------------------------------------------------------- internal entry points

eh_status graph_canonicalize(eh_ctx *ctx, uint32_t n, DevBuf &keys, uint64_t nkeys, int relabel,
                             eh_graph **out, const uint32_t *vgraph);
eh_status graph_relabel(eh_graph *g);
eh_status matrix_build_dev(eh_ctx *ctx, uint32_t n, const uint64_t *d_off, const uint32_t *d_nbr,
                           uint64_t entries, eh_edge_filter filter, const uint8_t *d_mask,
                           eh_layout_policy policy, const uint32_t *d_graph_of, bool symmetric,
                           eh_matrix **out, const uint8_t *d_types_in = nullptr);
// the learned (or fixed-arm) layout of every row; sym_off: the row offsets of the symmetric
// graph whose lower triangle m holds (its degrees give the uses), or null
eh_status matrix_select_learned(eh_matrix *m, uint8_t *d_types, const uint64_t *sym_off);
eh_status count_run(eh_matrix *m, int k, const eh_run_opts *opts, uint64_t *d_counts,
                    uint32_t ncounts, uint64_t *count, eh_run_stats *stats, uint64_t *d_task_cycles);
eh_status schedule_build(eh_matrix *m, int sched, int k, const TuneParams &tp, Schedule &s);
// the schedule of a k-pass under opts: built, or reused while every parameter it was built for
// holds; *sched_ms = its build time when built here, else 0
eh_status schedule_ensure(eh_matrix *m, int k, const eh_run_opts *opts, double *sched_ms);
// d_counts3 (4-cliques only, may be null): the triangle counters of the same rows, counted in
// the 4-clique pass (every row of the schedule must be a short row: the triangle pass is then
// not run); see k_small_cliques4
eh_status schedule_launch(eh_matrix *m, int k, int sched, const TuneParams &tp, Schedule &s,
                          uint64_t *d_counts, uint32_t *d_next, uint64_t *d_task_cycles, PartEv *pev = nullptr,
                          uint64_t *d_counts3 = nullptr);
TuneParams tune_defaults(eh_matrix *m, int sched, int k, const eh_run_opts *opts);
eh_status launch_parts(eh_matrix *m, int k, int sched, const TuneParams &tp, const uint32_t *small_rows,
                       uint64_t n_small_rows, const uint4 *tasks_big, uint64_t n_big, const uint4 *tasks_small,
                       uint64_t n_small, uint32_t lds_small, uint32_t lds_big, uint64_t *d_counts, uint32_t *d_next,
                       uint64_t *d_task_cycles, const uint4 *btasks = nullptr, uint64_t n_btasks = 0,
                       const uint2 *in_pairs = nullptr, uint32_t bstat_j = 0, const uint32_t *jstart = nullptr,
                       const uint4 *lo16 = nullptr, const uint32_t *local_rows0 = nullptr, uint64_t n_local0 = 0,
                       const uint32_t *local_rows1 = nullptr, uint64_t n_local1 = 0,
                       const Schedule *hs = nullptr, PartEv *pev = nullptr, uint64_t *d_counts3 = nullptr);
eh_status matrix_triangle_weight(eh_matrix *m, uint32_t ncounts, uint64_t *out);
eh_status ctx_part_events(eh_ctx *ctx);

// records this device's address of eh_intersect_rows_static's staging array (device current)
eh_status static_scratch_init(int device);

// timing helper: device ms between two events recorded on ctx->stream
double ctx_elapsed_ms(eh_ctx *ctx);

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
