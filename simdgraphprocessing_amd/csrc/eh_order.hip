// eh_order.hip -- the vertex orderings of the reference's conversion tool
// (tools/undirectedEdgeListToBinary.cpp:21-32, MutableGraph::reorder_*,
// src/MutableGraph.hpp:100-285) as device passes.  Each ordering produces a
// new2old permutation that graph_apply_order applies to the CSR.
//
//   DEGREE      degree descending, id ascending (reorder_by_degree)
//   REV_DEGREE  degree ascending, id ascending (reorder_by_rev_degree; the reference's
//               unstable sort leaves ties unspecified, here they keep id order)
//   STRONG_RUN  rows taken in DEGREE order; each vertex gets the next id where it first
//               appears in their concatenation (reorder_strong_run); vertices that appear
//               in no row follow in id order (the reference assumes there are none)
//   BFS         breadth-first from the first vertex of a seeded permutation; a vertex's
//               id follows the order it is first reached in (earlier parent first, then
//               its position in the parent's row); when a component is exhausted the next
//               unvisited vertex of the permutation starts the next one (reorder_bfs)
//   SHINGLES    two seeded rankings; each vertex keyed by the lowest-ranked neighbour
//               under each, sorted by the pair of keys, ties in id order
//               (reorder_by_shingles)
//   RANDOM      a seeded permutation (reorder_random)
//   THE_GAME    BFS, then DEGREE (reorder_by_the_game)
//
// The reference draws its permutations with std::random_shuffle; here a permutation is
// the ids sorted by a counter-based 32-bit draw (ties by id), so every one is a pure
// function of (seed, stream) and the oracle reproduces it.
//
// BFS runs all components at once: each vertex is first labelled with the smallest
// permutation position in its component (min-label propagation with pointer jumping),
// which is exactly the reference's restart order, and the component's root is the
// vertex at that position.  Levels then grow from all roots together; a vertex reached
// at level L + 1 keeps the smallest (rank of parent in level L, index in the parent's
// row) key, and level L + 1 is ranked by sorting on (component, key).  The final order
// is (component, level, rank).
#include "eh_internal.hpp"
#include "eh_rng.hpp"

namespace {

constexpr uint32_t UNSET = 0xFFFFFFFFu;

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = (work + block - 1) / block;
  if (g == 0) g = 1;
  return (unsigned)(g > 65536 ? 65536 : g);
}

__global__ void k_perm_keys(uint64_t key, uint32_t n, uint64_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = ((eh_draw(key, i) >> 32) << 32) | i;
}

__global__ void k_low32(const uint64_t *in, uint32_t n, uint32_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)in[i];
}

__global__ void k_iota32(uint32_t *p, uint32_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

__global__ void k_degrees(const uint64_t *off, uint32_t n, uint32_t *deg) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    deg[i] = (uint32_t)(off[i + 1] - off[i]);
}

__global__ void k_gather32(const uint32_t *src, const uint32_t *idx, uint32_t n, uint32_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = src[idx[i]];
}

__global__ void k_not32(const uint32_t *in, uint32_t n, uint32_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = ~in[i];
}

// pos[perm[p]] = p
__global__ void k_positions(const uint32_t *perm, uint32_t n, uint32_t *pos) {
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x)
    pos[perm[p]] = (uint32_t)p;
}

__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill64(unsigned long long *p, uint64_t n, unsigned long long v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------- strong run

// entry k of the row ranked r sits at position pre[r] + k of the concatenation
__global__ void k_strong_first(const uint64_t *off, const uint32_t *nbr, const uint32_t *order, const uint64_t *pre,
                               uint32_t n, unsigned long long *first) {
  for (uint64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const uint32_t v = order[r];
    const uint64_t b = off[v], e = off[v + 1];
    for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x)
      atomicMin(&first[nbr[k]], (unsigned long long)(pre[r] + (k - b)));
  }
}

// flag[p] = 1 where position p is the first appearance of its vertex; val[p] = vertex
__global__ void k_strong_flags(const uint64_t *off, const uint32_t *nbr, const uint32_t *order, const uint64_t *pre,
                               uint32_t n, const unsigned long long *first, uint32_t *flag, uint32_t *val) {
  for (uint64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const uint32_t v = order[r];
    const uint64_t b = off[v], e = off[v + 1];
    for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
      const uint64_t p = pre[r] + (k - b);
      const uint32_t u = nbr[k];
      flag[p] = first[u] == p ? 1u : 0u;
      val[p] = u;
    }
  }
}

__global__ void k_scatter_flagged(const uint32_t *flag, const uint32_t *val, const uint64_t *pos, uint64_t m,
                                  uint64_t base, uint32_t *out) {
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < m; p += (uint64_t)gridDim.x * blockDim.x)
    if (flag[p]) out[base + pos[p]] = val ? val[p] : (uint32_t)p;
}

__global__ void k_unseen(const unsigned long long *first, uint32_t n, uint32_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = (i < n && first[i] == ~0ull) ? 1u : 0u;
}

// ---------------------------------------------------------------- shingles

// one wave per row: the neighbour with the smallest rank; an empty row's shingle is vertex 0
// (src/MutableGraph.hpp:230-240)
__global__ void __launch_bounds__(256) k_shingle(const uint64_t *off, const uint32_t *nbr, const uint32_t *rank,
                                                 uint32_t n, uint32_t *out) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    const uint64_t b = off[i], e = off[i + 1];
    uint64_t best = ~0ull;
    for (uint64_t k = b + lane; k < e; k += 64) {
      const uint32_t x = nbr[k];
      const uint64_t key = ((uint64_t)rank[x] << 32) | x;
      best = key < best ? key : best;
    }
    for (int o = 32; o; o >>= 1) {
      const uint64_t other = __shfl_xor(best, o, 64);
      best = other < best ? other : best;
    }
    if (lane == 0) out[i] = best == ~0ull ? 0u : (uint32_t)best;
  }
}

// ---------------------------------------------------------------- BFS

// label propagation over every entry: label[v] = min(label[v], label[u])
__global__ void k_cc_step(const uint64_t *off, const uint32_t *nbr, uint32_t n, uint32_t *label, uint32_t *changed) {
  for (uint64_t v = blockIdx.x; v < n; v += gridDim.x) {
    const uint64_t b = off[v], e = off[v + 1];
    uint32_t m = label[v];
    for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
      const uint32_t lu = label[nbr[k]];
      m = lu < m ? lu : m;
    }
    if (m < label[v]) {
      atomicMin(&label[v], m);
      *changed = 1u;
    }
  }
}

// pointer jumping: a label is a permutation position; the vertex there has a label <= it
__global__ void k_cc_jump(const uint32_t *perm, uint32_t n, uint32_t *label, uint32_t *changed) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t l = label[v];
    const uint32_t ll = label[perm[l]];
    if (ll < l) {
      atomicMin(&label[v], ll);
      *changed = 1u;
    }
  }
}

// roots: the vertex at its component's label position; level 0, rank 0
__global__ void k_bfs_roots(const uint32_t *label, const uint32_t *pos, uint32_t n, uint32_t *level, uint32_t *flag) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= n; v += (uint64_t)gridDim.x * blockDim.x) {
    const bool root = v < n && label[v] == pos[v];
    if (v < n) level[v] = root ? 0u : UNSET;
    flag[v] = root ? 1u : 0u;
  }
}

// frontier vertex f (rank = its index) offers (rank << 32 | k) to each unvisited neighbour
__global__ void k_bfs_expand(const uint64_t *off, const uint32_t *nbr, const uint32_t *frontier, uint32_t nf,
                             const uint32_t *level, unsigned long long *disc) {
  for (uint64_t f = blockIdx.x; f < nf; f += gridDim.x) {
    const uint32_t v = frontier[f];
    const uint64_t b = off[v], e = off[v + 1];
    for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
      const uint32_t u = nbr[k];
      if (level[u] == UNSET) atomicMin(&disc[u], ((unsigned long long)f << 32) | (k - b));
    }
  }
}

__global__ void k_bfs_mark(uint32_t n, uint32_t lv, const unsigned long long *disc, uint32_t *level, uint32_t *flag) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= n; v += (uint64_t)gridDim.x * blockDim.x) {
    const bool hit = v < n && level[v] == UNSET && disc[v] != ~0ull;
    if (hit) level[v] = lv;
    flag[v] = hit ? 1u : 0u;
  }
}

__global__ void k_bfs_keys(const uint32_t *list, uint32_t m, const unsigned long long *disc, const uint32_t *label,
                           uint32_t *klo, uint32_t *khi, uint32_t *kcomp) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = list[i];
    const unsigned long long d = disc[v];
    klo[i] = (uint32_t)d;
    khi[i] = (uint32_t)(d >> 32);
    kcomp[i] = label[v];
  }
}

// rank of every vertex of the sorted level list = its index
__global__ void k_bfs_rank(const uint32_t *list, uint32_t m, uint32_t *rank) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
    rank[list[i]] = (uint32_t)i;
}

// sort ids by key columns, last column first (each pass a stable radix sort)
eh_status sort_ids_by(eh_ctx *ctx, uint32_t *ids, uint32_t n, std::initializer_list<const uint32_t *> cols_last_first) {
  hipStream_t st = ctx->stream;
  DevBuf k, kt, vt;
  EH_CHECK(k.alloc(4ull * n + 4));
  EH_CHECK(kt.alloc(4ull * n + 4));
  EH_CHECK(vt.alloc(4ull * n + 4));
  for (const uint32_t *col : cols_last_first) {
    k_gather32<<<grid_for(n), 256, 0, st>>>(col, ids, n, k.as<uint32_t>());
    EH_CHECK(prim_sort_pairs_u32(ctx, k.as<uint32_t>(), kt.as<uint32_t>(), ids, vt.as<uint32_t>(), n));
  }
  return EH_OK;
}

// a seeded permutation of 0..n-1 (new2old form)
eh_status seeded_perm(eh_ctx *ctx, uint64_t seed, uint64_t stream, uint32_t n, DevBuf &perm) {
  hipStream_t st = ctx->stream;
  DevBuf keys, tmp;
  EH_CHECK(keys.alloc(8ull * n + 8));
  EH_CHECK(tmp.alloc(8ull * n + 8));
  k_perm_keys<<<grid_for(n), 256, 0, st>>>(eh_stream_key(seed, stream), n, keys.as<uint64_t>());
  EH_CHECK(prim_sort_u64(ctx, keys.as<uint64_t>(), tmp.as<uint64_t>(), n, 64));
  EH_CHECK(perm.alloc(4ull * n + 4));
  k_low32<<<grid_for(n), 256, 0, st>>>(keys.as<uint64_t>(), n, perm.as<uint32_t>());
  return EH_OK;
}

// compact the flagged ids 0..n-1 (flag has n + 1 entries) into out[base..]; returns the count
eh_status compact_ids(eh_ctx *ctx, const uint32_t *flag, uint32_t n, uint32_t *out, uint64_t base, uint64_t *count) {
  DevBuf pos;
  EH_CHECK(pos.alloc(8ull * n + 16));
  EH_CHECK(prim_exscan_u32_to_u64(ctx, flag, pos.as<uint64_t>(), (uint64_t)n + 1));
  EH_HIP(hipMemcpy(count, pos.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost));
  k_scatter_flagged<<<grid_for(n), 256, 0, ctx->stream>>>(flag, nullptr, pos.as<uint64_t>(), n, base, out);
  EH_HIP(hipGetLastError());
  return EH_OK;
}

eh_status order_rev_degree(eh_graph *g, DevBuf &n2o) {
  eh_ctx *ctx = g->ctx;
  const uint32_t n = g->n;
  DevBuf deg;
  EH_CHECK(deg.alloc(4ull * n + 4));
  EH_CHECK(n2o.alloc(4ull * n + 4));
  k_degrees<<<grid_for(n), 256, 0, ctx->stream>>>(g->off.as<uint64_t>(), n, deg.as<uint32_t>());
  k_iota32<<<grid_for(n), 256, 0, ctx->stream>>>(n2o.as<uint32_t>(), n);
  return sort_ids_by(ctx, n2o.as<uint32_t>(), n, {deg.as<uint32_t>()});
}

eh_status order_strong_run(eh_graph *g, DevBuf &n2o) {
  eh_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t n = g->n;
  const uint64_t m = g->entries;
  // rows in degree order (degree descending, id ascending)
  DevBuf deg, negdeg, order, rdeg, pre, first, flag, val, pos;
  EH_CHECK(deg.alloc(4ull * n + 4));
  EH_CHECK(order.alloc(4ull * n + 4));
  k_degrees<<<grid_for(n), 256, 0, st>>>(g->off.as<uint64_t>(), n, deg.as<uint32_t>());
  EH_CHECK(negdeg.alloc(4ull * n + 4));
  k_not32<<<grid_for(n), 256, 0, st>>>(deg.as<uint32_t>(), n, negdeg.as<uint32_t>());  // descending = ascending ~deg
  k_iota32<<<grid_for(n), 256, 0, st>>>(order.as<uint32_t>(), n);
  EH_CHECK(sort_ids_by(ctx, order.as<uint32_t>(), n, {negdeg.as<uint32_t>()}));
  // concatenation offsets of the rows in that order
  EH_CHECK(rdeg.alloc(4ull * n + 8));
  k_gather32<<<grid_for(n), 256, 0, st>>>(deg.as<uint32_t>(), order.as<uint32_t>(), n, rdeg.as<uint32_t>());
  k_fill32<<<1, 64, 0, st>>>(rdeg.as<uint32_t>() + n, 1, 0u);
  EH_CHECK(pre.alloc(8ull * n + 16));
  EH_CHECK(prim_exscan_u32_to_u64(ctx, rdeg.as<uint32_t>(), pre.as<uint64_t>(), (uint64_t)n + 1));
  EH_CHECK(first.alloc(8ull * n + 8));
  k_fill64<<<grid_for(n), 256, 0, st>>>(first.as<unsigned long long>(), n, ~0ull);
  k_strong_first<<<grid_for((uint64_t)n * 64, 64), 64, 0, st>>>(g->off.as<uint64_t>(), g->nbr.as<uint32_t>(),
                                                                order.as<uint32_t>(), pre.as<uint64_t>(), n,
                                                                first.as<unsigned long long>());
  EH_CHECK(flag.alloc(4ull * m + 8));
  EH_CHECK(val.alloc(4ull * m + 8));
  k_fill32<<<1, 64, 0, st>>>(flag.as<uint32_t>() + m, 1, 0u);
  k_strong_flags<<<grid_for((uint64_t)n * 64, 64), 64, 0, st>>>(g->off.as<uint64_t>(), g->nbr.as<uint32_t>(),
                                                                order.as<uint32_t>(), pre.as<uint64_t>(), n,
                                                                first.as<unsigned long long>(), flag.as<uint32_t>(),
                                                                val.as<uint32_t>());
  EH_CHECK(pos.alloc(8ull * m + 16));
  EH_CHECK(prim_exscan_u32_to_u64(ctx, flag.as<uint32_t>(), pos.as<uint64_t>(), m + 1));
  uint64_t seen = 0;
  EH_HIP(hipMemcpy(&seen, pos.as<uint64_t>() + m, 8, hipMemcpyDeviceToHost));
  EH_CHECK(n2o.alloc(4ull * n + 4));
  k_scatter_flagged<<<grid_for(m), 256, 0, st>>>(flag.as<uint32_t>(), val.as<uint32_t>(), pos.as<uint64_t>(), m, 0,
                                                 n2o.as<uint32_t>());
  // vertices in no row follow in id order
  DevBuf uflag;
  EH_CHECK(uflag.alloc(4ull * n + 8));
  k_unseen<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(first.as<unsigned long long>(), n, uflag.as<uint32_t>());
  uint64_t unseen = 0;
  EH_CHECK(compact_ids(ctx, uflag.as<uint32_t>(), n, n2o.as<uint32_t>(), seen, &unseen));
  if (seen + unseen != n) {
    eh_set_error("strong-run ordering covered %llu of %u vertices", (unsigned long long)(seen + unseen), n);
    return EH_ERR_INVALID;
  }
  EH_HIP(hipGetLastError());
  return EH_OK;
}

eh_status order_shingles(eh_graph *g, uint64_t seed, DevBuf &n2o) {
  eh_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t n = g->n;
  DevBuf r0, r1, s0, s1;
  EH_CHECK(seeded_perm(ctx, seed, 103, n, r0));
  EH_CHECK(seeded_perm(ctx, seed, 104, n, r1));
  EH_CHECK(s0.alloc(4ull * n + 4));
  EH_CHECK(s1.alloc(4ull * n + 4));
  k_shingle<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(g->off.as<uint64_t>(), g->nbr.as<uint32_t>(),
                                                        r0.as<uint32_t>(), n, s0.as<uint32_t>());
  k_shingle<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(g->off.as<uint64_t>(), g->nbr.as<uint32_t>(),
                                                        r1.as<uint32_t>(), n, s1.as<uint32_t>());
  EH_CHECK(n2o.alloc(4ull * n + 4));
  k_iota32<<<grid_for(n), 256, 0, st>>>(n2o.as<uint32_t>(), n);
  return sort_ids_by(ctx, n2o.as<uint32_t>(), n, {s1.as<uint32_t>(), s0.as<uint32_t>()});
}

eh_status order_bfs(eh_graph *g, uint64_t seed, DevBuf &n2o) {
  eh_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t n = g->n;
  DevBuf perm, pos, label, changed, level, rank, disc, flag, frontier, klo, khi, kcomp;
  EH_CHECK(seeded_perm(ctx, seed, 102, n, perm));
  // pos[v] = position of v in the permutation, the starting label of v
  EH_CHECK(pos.alloc(4ull * n + 4));
  EH_CHECK(label.alloc(4ull * n + 4));
  k_positions<<<grid_for(n), 256, 0, st>>>(perm.as<uint32_t>(), n, pos.as<uint32_t>());
  EH_HIP(hipMemcpyAsync(label.p, pos.p, 4ull * n, hipMemcpyDeviceToDevice, st));
  // components: smallest permutation position in each
  EH_CHECK(changed.alloc(8));
  for (int it = 0;; it++) {
    k_fill32<<<1, 64, 0, st>>>(changed.as<uint32_t>(), 1, 0u);
    k_cc_step<<<grid_for((uint64_t)n * 64, 64), 64, 0, st>>>(g->off.as<uint64_t>(), g->nbr.as<uint32_t>(), n,
                                                             label.as<uint32_t>(), changed.as<uint32_t>());
    for (int j = 0; j < 4; j++)
      k_cc_jump<<<grid_for(n), 256, 0, st>>>(perm.as<uint32_t>(), n, label.as<uint32_t>(), changed.as<uint32_t>());
    uint32_t h = 0;
    EH_HIP(hipMemcpy(&h, changed.p, 4, hipMemcpyDeviceToHost));
    if (!h) break;
    if (it > (int)n + 2) { eh_set_error("BFS ordering: components did not converge"); return EH_ERR_HIP; }
  }
  // levels from all roots at once
  EH_CHECK(level.alloc(4ull * n + 4));
  EH_CHECK(rank.alloc(4ull * n + 4));
  EH_CHECK(disc.alloc(8ull * n + 8));
  EH_CHECK(flag.alloc(4ull * n + 8));
  EH_CHECK(frontier.alloc(4ull * n + 4));
  EH_CHECK(klo.alloc(4ull * n + 4));
  EH_CHECK(khi.alloc(4ull * n + 4));
  EH_CHECK(kcomp.alloc(4ull * n + 4));
  k_fill64<<<grid_for(n), 256, 0, st>>>(disc.as<unsigned long long>(), n, ~0ull);
  k_bfs_roots<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(label.as<uint32_t>(), pos.as<uint32_t>(), n,
                                                         level.as<uint32_t>(), flag.as<uint32_t>());
  uint64_t nf = 0;
  EH_CHECK(compact_ids(ctx, flag.as<uint32_t>(), n, frontier.as<uint32_t>(), 0, &nf));
  // roots sorted by component (their own label), rank = index
  EH_CHECK(sort_ids_by(ctx, frontier.as<uint32_t>(), (uint32_t)nf, {label.as<uint32_t>()}));
  k_bfs_rank<<<grid_for(nf), 256, 0, st>>>(frontier.as<uint32_t>(), (uint32_t)nf, rank.as<uint32_t>());
  uint64_t visited = nf;
  for (uint32_t lv = 1; nf > 0 && visited < n; lv++) {
    k_bfs_expand<<<grid_for(nf * 64, 64), 64, 0, st>>>(g->off.as<uint64_t>(), g->nbr.as<uint32_t>(),
                                                       frontier.as<uint32_t>(), (uint32_t)nf, level.as<uint32_t>(),
                                                       disc.as<unsigned long long>());
    k_bfs_mark<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(n, lv, disc.as<unsigned long long>(), level.as<uint32_t>(),
                                                          flag.as<uint32_t>());
    EH_CHECK(compact_ids(ctx, flag.as<uint32_t>(), n, frontier.as<uint32_t>(), 0, &nf));
    if (nf == 0) break;
    // the new level ranked by (component, parent rank, index in the parent's row)
    k_bfs_keys<<<grid_for(nf), 256, 0, st>>>(frontier.as<uint32_t>(), (uint32_t)nf, disc.as<unsigned long long>(),
                                             label.as<uint32_t>(), klo.as<uint32_t>(), khi.as<uint32_t>(),
                                             kcomp.as<uint32_t>());
    {
      // the key columns are per list position: sort positions, then map to vertices
      DevBuf idx, k, kt, vt, ord;
      EH_CHECK(idx.alloc(4ull * nf + 4));
      EH_CHECK(k.alloc(4ull * nf + 4));
      EH_CHECK(kt.alloc(4ull * nf + 4));
      EH_CHECK(vt.alloc(4ull * nf + 4));
      k_iota32<<<grid_for(nf), 256, 0, st>>>(idx.as<uint32_t>(), (uint32_t)nf);
      for (const uint32_t *col : {klo.as<uint32_t>(), khi.as<uint32_t>(), kcomp.as<uint32_t>()}) {
        k_gather32<<<grid_for(nf), 256, 0, st>>>(col, idx.as<uint32_t>(), (uint32_t)nf, k.as<uint32_t>());
        EH_CHECK(prim_sort_pairs_u32(ctx, k.as<uint32_t>(), kt.as<uint32_t>(), idx.as<uint32_t>(), vt.as<uint32_t>(), nf));
      }
      EH_CHECK(ord.alloc(4ull * nf + 4));
      k_gather32<<<grid_for(nf), 256, 0, st>>>(frontier.as<uint32_t>(), idx.as<uint32_t>(), (uint32_t)nf,
                                               ord.as<uint32_t>());
      EH_HIP(hipMemcpyAsync(frontier.p, ord.p, 4ull * nf, hipMemcpyDeviceToDevice, st));
      EH_HIP(hipStreamSynchronize(st));
    }
    k_bfs_rank<<<grid_for(nf), 256, 0, st>>>(frontier.as<uint32_t>(), (uint32_t)nf, rank.as<uint32_t>());
    visited += nf;
  }
  if (visited != n) {
    eh_set_error("BFS ordering reached %llu of %u vertices", (unsigned long long)visited, n);
    return EH_ERR_HIP;
  }
  // final order: (component, level, rank)
  EH_CHECK(n2o.alloc(4ull * n + 4));
  k_iota32<<<grid_for(n), 256, 0, st>>>(n2o.as<uint32_t>(), n);
  EH_CHECK(sort_ids_by(ctx, n2o.as<uint32_t>(), n, {rank.as<uint32_t>(), level.as<uint32_t>(), label.as<uint32_t>()}));
  EH_HIP(hipGetLastError());
  return EH_OK;
}

}  // namespace

eh_status graph_relabel(eh_graph *g);

extern "C" eh_status eh_graph_reorder(eh_graph *g, eh_ordering ordering, uint64_t seed) {
  if (!g) { eh_set_error("null graph"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(g->ctx->device));
  DevBuf n2o;
  switch (ordering) {
    case EH_ORDER_DEGREE: return graph_relabel(g);
    case EH_ORDER_REV_DEGREE: EH_CHECK(order_rev_degree(g, n2o)); break;
    case EH_ORDER_STRONG_RUN: EH_CHECK(order_strong_run(g, n2o)); break;
    case EH_ORDER_BFS: EH_CHECK(order_bfs(g, seed, n2o)); break;
    case EH_ORDER_SHINGLES: EH_CHECK(order_shingles(g, seed, n2o)); break;
    case EH_ORDER_RANDOM: EH_CHECK(seeded_perm(g->ctx, seed, 101, g->n, n2o)); break;
    case EH_ORDER_THE_GAME:
      EH_CHECK(order_bfs(g, seed, n2o));
      EH_CHECK(graph_apply_order(g, n2o.as<uint32_t>()));
      return graph_relabel(g);
    default: eh_set_error("unknown ordering %d", (int)ordering); return EH_ERR_INVALID;
  }
  return graph_apply_order(g, n2o.as<uint32_t>());
}
