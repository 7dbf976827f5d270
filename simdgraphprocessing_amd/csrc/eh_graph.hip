// eh_graph.hip -- seeded generators and the device canonicaliser.
//
// Pipeline (all on the device):
//   draws -> undirected keys (min<<32 | max; self-loops become a sentinel)
//   -> radix sort -> unique -> [degree count -> (degree desc, id asc) relabel]
//   -> both directions -> radix sort -> symmetric CSR (sorted rows).
// The relabel restates MutableGraph::reorder_by_degree (src/MutableGraph.hpp:20-32,
// 100-123, 268-272); the reference does it offline on the host.
#include <cmath>
#include <cstring>

#include "eh_internal.hpp"
#include "eh_rng.hpp"

#define EH_SENTINEL 0xFFFFFFFFFFFFFFFFull

namespace {

__device__ __forceinline__ uint64_t undirected_key(uint32_t u, uint32_t v) {
  if (u == v) return EH_SENTINEL;
  uint32_t a = u < v ? u : v, b = u < v ? v : u;
  return ((uint64_t)a << 32) | b;
}

__global__ void k_gen_er(uint64_t *keys, uint64_t m, uint32_t n, uint64_t key) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = eh_draw(key, e);
    keys[e] = undirected_key(eh_scale32((uint32_t)r, n), eh_scale32((uint32_t)(r >> 32), n));
  }
}

// G(n, m) with exactly m distinct pairs: the ER draws with their draw index
__global__ void k_gen_er_idx(uint64_t *keys, uint32_t *idx, uint64_t m, uint32_t n, uint64_t key) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = eh_draw(key, e);
    keys[e] = undirected_key(eh_scale32((uint32_t)r, n), eh_scale32((uint32_t)(r >> 32), n));
    idx[e] = (uint32_t)e;
  }
}

// after a stable sort by key: the first entry of each pair's run holds its first draw
__global__ void k_first_draw_flags(const uint64_t *keys, uint64_t m, uint32_t *flag) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x)
    flag[e] = (keys[e] != EH_SENTINEL && (e == 0 || keys[e] != keys[e - 1])) ? 1u : 0u;
}

__global__ void k_first_draw_scatter(const uint32_t *flag, const uint64_t *pos, const uint32_t *idx, uint64_t m,
                                     uint32_t *first) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x)
    if (flag[e]) first[pos[e]] = idx[e];
}

__global__ void k_er_keys_of_draws(const uint32_t *draws, uint64_t m, uint32_t n, uint64_t key, uint64_t *keys) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < m;
       q += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = eh_draw(key, draws[q]);
    keys[q] = undirected_key(eh_scale32((uint32_t)r, n), eh_scale32((uint32_t)(r >> 32), n));
  }
}

__global__ void k_gen_rmat(uint64_t *keys, uint64_t m, uint32_t scale, uint32_t t0, uint32_t t1,
                           uint32_t t2, uint64_t key) {
  const uint64_t per = (scale + 1) / 2;
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = 0, v = 0;
    uint64_t r = 0;
    for (uint32_t l = 0; l < scale; l++) {
      if ((l & 1) == 0) r = eh_draw(key, e * per + l / 2);
      uint32_t x = (l & 1) ? (uint32_t)(r >> 32) : (uint32_t)r;
      uint32_t bu = x >= t1;
      uint32_t bv = (x >= t0 && x < t1) || x >= t2;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    keys[e] = undirected_key(u, v);
  }
}

__device__ __forceinline__ uint32_t cdf_pick(const uint64_t *cdf, uint32_t n, uint32_t r) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)r < cdf[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}

__global__ void k_gen_chung_lu(uint64_t *keys, uint64_t m, const uint64_t *cdf, uint32_t n,
                               uint64_t key) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = eh_draw(key, e);
    keys[e] = undirected_key(cdf_pick(cdf, n, (uint32_t)r), cdf_pick(cdf, n, (uint32_t)(r >> 32)));
  }
}

// one slot per ordered pair (s,t) of every community; absent pairs write the sentinel
__global__ void k_gen_communities(uint64_t *keys, uint32_t n, uint32_t ncomm, uint32_t csize,
                                  uint64_t seed) {
  const uint64_t per = (uint64_t)csize * csize;
  const uint64_t total = per * ncomm;
  const uint32_t stride = n / ncomm;
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < total;
       p += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = (uint32_t)(p / per);
    uint32_t rem = (uint32_t)(p % per);
    uint32_t s = rem / csize, t = rem % csize;
    uint64_t k = EH_SENTINEL;
    if (s < t) {
      uint64_t key = eh_stream_key(seed, 1 + (uint64_t)c);
      if (eh_draw(key, (uint64_t)s * csize + t) >> 63) {
        uint32_t base = c * stride;
        k = ((uint64_t)(base + s) << 32) | (base + t);
      }
    }
    keys[p] = k;
  }
}

__global__ void k_edges_to_keys(uint64_t *keys, const uint32_t *src, const uint32_t *dst, uint64_t m,
                                uint32_t n) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = src[e], v = dst[e];
    keys[e] = (u >= n || v >= n) ? EH_SENTINEL : undirected_key(u, v);
  }
}

// Keys sorted by their high word: equal highs form runs, so a wave adds each run
// it holds with one atomic (hub rows would otherwise serialise on one counter).
// Called by every lane of a wave together; `valid` marks lanes holding a key.
__device__ __forceinline__ void add_sorted_runs(uint32_t hi, bool valid, uint32_t *cnt) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t prev = __shfl_up(hi, 1, 64);
  const bool start = valid && (lane == 0 || prev != hi);
  const uint64_t starts = __ballot(start);
  const uint64_t valids = __ballot(valid);
  if (start) {
    const uint64_t later = lane == 63 ? 0ull : (starts >> (lane + 1));
    const uint32_t next = later ? lane + 1 + (uint32_t)__builtin_ctzll(later) : 64u;
    const uint64_t upto = next == 64 ? ~0ull : ((1ull << next) - 1ull);
    const uint64_t span = upto & ~((1ull << lane) - 1ull);
    atomicAdd(&cnt[hi], (uint32_t)__popcll(valids & span));
  }
}

// degrees from sorted unique (min << 32 | max) keys
__global__ void k_degree_from_keys(const uint64_t *keys, uint64_t m, uint32_t *deg) {
  // wave-uniform grid stride: every lane of a wave runs the same iterations
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < m; base += stride) {
    const uint64_t e = base + (threadIdx.x & 63u);
    const bool valid = e < m;
    const uint64_t k = valid ? keys[e] : 0ull;
    add_sorted_runs((uint32_t)(k >> 32), valid, deg);
    if (valid) atomicAdd(&deg[(uint32_t)k], 1u);
  }
}

// (degree desc, id asc) as one ascending 64-bit key
__global__ void k_relabel_keys(const uint32_t *deg, uint32_t n, uint64_t *vk) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    vk[i] = ((uint64_t)(0xFFFFFFFFu - deg[i]) << 32) | (uint32_t)i;
}

// batch arenas: (graph, degree desc, id asc) so every graph keeps its id range
__global__ void k_relabel_keys_grouped(const uint32_t *deg, const uint32_t *vgraph, uint32_t n, uint64_t *vk) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    vk[i] = ((uint64_t)vgraph[i] << 52) | ((uint64_t)(0xFFFFFFu - min(deg[i], 0xFFFFFFu)) << 26) | (uint32_t)i;
}

__global__ void k_perm_from_sorted_masked(const uint64_t *vk, uint32_t n, uint32_t *new2old, uint32_t *old2new) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t o = (uint32_t)(vk[r] & 0x3FFFFFFull);
    new2old[r] = o;
    old2new[o] = (uint32_t)r;
  }
}

__global__ void k_perm_from_sorted(const uint64_t *vk, uint32_t n, uint32_t *new2old,
                                   uint32_t *old2new) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t o = (uint32_t)vk[r];
    new2old[r] = o;
    old2new[o] = (uint32_t)r;
  }
}

__global__ void k_both_directions(const uint64_t *ukeys, uint64_t m, const uint32_t *old2new,
                                  uint64_t *out) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = ukeys[e];
    uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
    if (old2new) { a = old2new[a]; b = old2new[b]; }
    out[2 * e] = ((uint64_t)a << 32) | b;
    out[2 * e + 1] = ((uint64_t)b << 32) | a;
  }
}

// row lengths from keys sorted by row (row << 32 | nbr)
__global__ void k_row_counts(const uint64_t *keys, uint64_t m, uint32_t *cnt) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < m; base += stride) {
    const uint64_t e = base + (threadIdx.x & 63u);
    const bool valid = e < m;
    const uint64_t k = valid ? keys[e] : 0ull;
    add_sorted_runs((uint32_t)(k >> 32), valid, cnt);
  }
}

__global__ void k_low_words(const uint64_t *keys, uint64_t m, uint32_t *nbr) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m;
       e += (uint64_t)gridDim.x * blockDim.x)
    nbr[e] = (uint32_t)keys[e];
}

__global__ void k_off_diff(const uint64_t *off, uint32_t n, uint32_t *deg) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    deg[i] = (uint32_t)(off[i + 1] - off[i]);
}

// remap a CSR through old2new into keys (new_row << 32 | new_nbr)
__global__ void k_remap_csr(const uint64_t *off, const uint32_t *nbr, uint32_t n,
                            const uint32_t *old2new, uint64_t *keys) {
  for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {
    uint64_t b = off[i], e = off[i + 1];
    uint64_t r = old2new[i];
    for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x)
      keys[k] = (r << 32) | old2new[nbr[k]];
  }
}

// validate a host CSR: strictly increasing rows, ids < n, no self-loops
__global__ void k_validate_csr(const uint64_t *off, const uint32_t *nbr, uint32_t n, uint32_t *bad) {
  for (uint64_t i = blockIdx.x; i < n; i += gridDim.x) {
    uint64_t b = off[i], e = off[i + 1];
    for (uint64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
      uint32_t v = nbr[k];
      bool ok = v < n && v != (uint32_t)i && (k == b || nbr[k - 1] < v);
      if (!ok) atomicOr(bad, 1u);
    }
  }
}

__global__ void k_iota(uint32_t *p, uint32_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

__global__ void k_inverse_perm(const uint32_t *new2old, uint32_t n, uint32_t *old2new) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x)
    old2new[new2old[r]] = (uint32_t)r;
}

__global__ void k_compose_perm(const uint32_t *outer, const uint32_t *inner, uint32_t n, uint32_t *out) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x)
    out[r] = outer[inner[r]];
}

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = ceil_div(work ? work : 1, block);
  return (unsigned)(g > 65536 ? 65536 : g);
}

// sorted unique undirected keys -> symmetric CSR (optionally relabelled)
eh_status finish_graph(eh_ctx *ctx, uint32_t n, const uint64_t *ukeys, uint64_t mu, int relabel,
                       eh_graph *g, const uint32_t *vgraph) {
  hipStream_t st = ctx->stream;
  DevBuf old2new;
  g->n = n;
  g->entries = 2 * mu;
  EH_CHECK(g->new2old.alloc(4 * (size_t)n + 4));
  if (relabel) {
    DevBuf deg, vk, vk2;
    EH_CHECK(deg.alloc(4 * (size_t)n + 4));
    EH_HIP(hipMemsetAsync(deg.p, 0, 4 * (size_t)n + 4, st));
    k_degree_from_keys<<<grid_for(mu), 256, 0, st>>>(ukeys, mu, deg.as<uint32_t>());
    EH_CHECK(vk.alloc(8 * (size_t)n + 8));
    EH_CHECK(vk2.alloc(8 * (size_t)n + 8));
    EH_CHECK(old2new.alloc(4 * (size_t)n + 4));
    if (vgraph) {
      k_relabel_keys_grouped<<<grid_for(n), 256, 0, st>>>(deg.as<uint32_t>(), vgraph, n, vk.as<uint64_t>());
      EH_CHECK(prim_sort_u64(ctx, vk.as<uint64_t>(), vk2.as<uint64_t>(), n, 64));
      k_perm_from_sorted_masked<<<grid_for(n), 256, 0, st>>>(vk.as<uint64_t>(), n, g->new2old.as<uint32_t>(),
                                                            old2new.as<uint32_t>());
    } else {
      k_relabel_keys<<<grid_for(n), 256, 0, st>>>(deg.as<uint32_t>(), n, vk.as<uint64_t>());
      EH_CHECK(prim_sort_u64(ctx, vk.as<uint64_t>(), vk2.as<uint64_t>(), n, 64));
      k_perm_from_sorted<<<grid_for(n), 256, 0, st>>>(vk.as<uint64_t>(), n, g->new2old.as<uint32_t>(),
                                                     old2new.as<uint32_t>());
    }
  } else {
    k_iota<<<grid_for(n), 256, 0, st>>>(g->new2old.as<uint32_t>(), n);
  }
  DevBuf sym, tmp, cnt;
  EH_CHECK(sym.alloc(16 * mu + 16));
  EH_CHECK(tmp.alloc(16 * mu + 16));
  k_both_directions<<<grid_for(mu), 256, 0, st>>>(ukeys, mu, relabel ? old2new.as<uint32_t>() : nullptr,
                                                  sym.as<uint64_t>());
  EH_CHECK(prim_sort_u64(ctx, sym.as<uint64_t>(), tmp.as<uint64_t>(), 2 * mu, 64));
  EH_CHECK(cnt.alloc(4 * (size_t)n + 4));
  EH_HIP(hipMemsetAsync(cnt.p, 0, 4 * (size_t)n + 4, st));
  k_row_counts<<<grid_for(2 * mu), 256, 0, st>>>(sym.as<uint64_t>(), 2 * mu, cnt.as<uint32_t>());
  EH_CHECK(g->off.alloc(8 * ((size_t)n + 1)));
  EH_CHECK(prim_exscan_u32_to_u64(ctx, cnt.as<uint32_t>(), g->off.as<uint64_t>(), (uint64_t)n + 1));
  EH_CHECK(g->nbr.alloc(4 * (2 * mu) + 16));
  k_low_words<<<grid_for(2 * mu), 256, 0, st>>>(sym.as<uint64_t>(), 2 * mu, g->nbr.as<uint32_t>());
  EH_CHECK(prim_reduce_max_u32(ctx, cnt.as<uint32_t>(), n, &g->max_degree));
  EH_HIP(hipStreamSynchronize(st));
  return EH_OK;
}

}  // namespace

eh_status graph_canonicalize(eh_ctx *ctx, uint32_t n, DevBuf &keys, uint64_t nkeys, int relabel,
                             eh_graph **out, const uint32_t *vgraph) {
  DevBuf tmp, uk;
  EH_CHECK(tmp.alloc(8 * nkeys + 8));
  EH_CHECK(prim_sort_u64(ctx, keys.as<uint64_t>(), tmp.as<uint64_t>(), nkeys, 64));
  uint64_t mu = 0;
  EH_CHECK(prim_unique_u64(ctx, keys.as<uint64_t>(), tmp.as<uint64_t>(), nkeys, &mu));
  if (mu) {
    uint64_t last = 0;
    EH_HIP(hipMemcpy(&last, tmp.as<uint64_t>() + (mu - 1), 8, hipMemcpyDeviceToHost));
    if (last == EH_SENTINEL) mu--;
  }
  keys.release();
  eh_graph *g = new eh_graph();
  g->ctx = ctx;
  eh_status s = finish_graph(ctx, n, tmp.as<uint64_t>(), mu, relabel, g, vgraph);
  if (s != EH_OK) { delete g; return s; }
  g->symmetric = true;  // every undirected key is stored as both of its entries
  *out = g;
  return EH_OK;
}

// Renumber the graph by new2old (device, n entries): rows and entries are remapped,
// every row re-sorted, and the permutation composed with any earlier one.
eh_status graph_apply_order(eh_graph *g, const uint32_t *d_new2old) {
  eh_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t n = g->n;
  DevBuf o2n, keys, tmp, cnt;
  EH_CHECK(o2n.alloc(4 * (size_t)n + 4));
  k_inverse_perm<<<grid_for(n), 256, 0, st>>>(d_new2old, n, o2n.as<uint32_t>());
  EH_CHECK(keys.alloc(8 * g->entries + 8));
  EH_CHECK(tmp.alloc(8 * g->entries + 8));
  k_remap_csr<<<grid_for((uint64_t)n * 64, 64), 64, 0, st>>>(g->off.as<uint64_t>(), g->nbr.as<uint32_t>(),
                                                             n, o2n.as<uint32_t>(), keys.as<uint64_t>());
  EH_CHECK(prim_sort_u64(ctx, keys.as<uint64_t>(), tmp.as<uint64_t>(), g->entries, 64));
  EH_CHECK(cnt.alloc(4 * (size_t)n + 4));
  EH_HIP(hipMemsetAsync(cnt.p, 0, 4 * (size_t)n + 4, st));
  k_row_counts<<<grid_for(g->entries), 256, 0, st>>>(keys.as<uint64_t>(), g->entries, cnt.as<uint32_t>());
  EH_CHECK(prim_exscan_u32_to_u64(ctx, cnt.as<uint32_t>(), g->off.as<uint64_t>(), (uint64_t)n + 1));
  k_low_words<<<grid_for(g->entries), 256, 0, st>>>(keys.as<uint64_t>(), g->entries, g->nbr.as<uint32_t>());
  // compose with any earlier renumbering: new2old_total[r] = old_map[new2old[r]]
  DevBuf composed;
  EH_CHECK(composed.alloc(4 * (size_t)n + 4));
  k_compose_perm<<<grid_for(n), 256, 0, st>>>(g->new2old.as<uint32_t>(), d_new2old, n, composed.as<uint32_t>());
  EH_HIP(hipGetLastError());
  EH_HIP(hipStreamSynchronize(st));
  g->new2old = std::move(composed);
  return EH_OK;
}

// MutableGraph::reorder_by_degree (src/MutableGraph.hpp:20-32,268-272): degree
// descending, old id ascending
eh_status graph_relabel(eh_graph *g) {
  eh_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  uint32_t n = g->n;
  DevBuf deg, vk, vk2, o2n, n2o;
  EH_CHECK(deg.alloc(4 * (size_t)n + 4));
  k_off_diff<<<grid_for(n), 256, 0, st>>>(g->off.as<uint64_t>(), n, deg.as<uint32_t>());
  EH_CHECK(vk.alloc(8 * (size_t)n + 8));
  EH_CHECK(vk2.alloc(8 * (size_t)n + 8));
  k_relabel_keys<<<grid_for(n), 256, 0, st>>>(deg.as<uint32_t>(), n, vk.as<uint64_t>());
  EH_CHECK(prim_sort_u64(ctx, vk.as<uint64_t>(), vk2.as<uint64_t>(), n, 64));
  EH_CHECK(o2n.alloc(4 * (size_t)n + 4));
  EH_CHECK(n2o.alloc(4 * (size_t)n + 4));
  k_perm_from_sorted<<<grid_for(n), 256, 0, st>>>(vk.as<uint64_t>(), n, n2o.as<uint32_t>(),
                                                 o2n.as<uint32_t>());
  return graph_apply_order(g, n2o.as<uint32_t>());
}

// ---------------------------------------------------------------- C ABI

// The first m distinct non-loop pairs of the ER draw stream of (n, seed) into keys (m of
// them): K draws are sorted stably by pair with their draw index, the first draw of every
// pair is kept, and the m smallest first draws name the pairs.  K grows until the draws
// hold m distinct pairs; any K that does gives the same m pairs.
static eh_status gen_er_exact(eh_ctx *ctx, uint32_t n, uint64_t m, uint64_t seed, DevBuf &keys) {
  hipStream_t st = ctx->stream;
  const uint64_t key = eh_stream_key(seed, 0);
  for (uint64_t K = m + m / 4 + 4096;; K *= 2) {
    if (K > 0xFFFFFFFFull) { eh_set_error("ER: %llu distinct pairs need more than 2^32 draws", (unsigned long long)m); return EH_ERR_UNSUPPORTED; }
    DevBuf dk, dk2, di, di2, flag, pos;
    EH_CHECK(dk.alloc(8 * K));
    EH_CHECK(dk2.alloc(8 * K));
    EH_CHECK(di.alloc(4 * K));
    EH_CHECK(di2.alloc(4 * K));
    EH_CHECK(flag.alloc(4 * K));
    EH_CHECK(pos.alloc(8 * K));
    k_gen_er_idx<<<grid_for(K), 256, 0, st>>>(dk.as<uint64_t>(), di.as<uint32_t>(), K, n, key);
    EH_CHECK(prim_sort_pairs_u64_u32(ctx, dk.as<uint64_t>(), dk2.as<uint64_t>(), di.as<uint32_t>(),
                                     di2.as<uint32_t>(), K));
    k_first_draw_flags<<<grid_for(K), 256, 0, st>>>(dk.as<uint64_t>(), K, flag.as<uint32_t>());
    EH_CHECK(prim_exscan_u32_to_u64(ctx, flag.as<uint32_t>(), pos.as<uint64_t>(), K));
    uint64_t last_pos = 0;
    uint32_t last_flag = 0;
    EH_HIP(hipMemcpyAsync(&last_pos, pos.as<uint64_t>() + K - 1, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipMemcpyAsync(&last_flag, flag.as<uint32_t>() + K - 1, 4, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    const uint64_t distinct = last_pos + last_flag;
    if (distinct < m) continue;
    DevBuf first, first2;
    EH_CHECK(first.alloc(4 * distinct));
    EH_CHECK(first2.alloc(4 * distinct));
    k_first_draw_scatter<<<grid_for(K), 256, 0, st>>>(flag.as<uint32_t>(), pos.as<uint64_t>(), di.as<uint32_t>(), K,
                                                       first.as<uint32_t>());
    EH_CHECK(prim_sort_u32(ctx, first.as<uint32_t>(), first2.as<uint32_t>(), distinct));
    EH_CHECK(keys.alloc(8 * m + 8));
    k_er_keys_of_draws<<<grid_for(m), 256, 0, st>>>(first.as<uint32_t>(), m, n, key, keys.as<uint64_t>());
    EH_HIP(hipGetLastError());
    EH_HIP(hipStreamSynchronize(st));
    return EH_OK;
  }
}

extern "C" eh_status eh_graph_generate(eh_ctx *ctx, const eh_gen_spec *spec, eh_graph **out) {
  if (!ctx || !spec || !out) { eh_set_error("eh_graph_generate: null argument"); return EH_ERR_INVALID; }
  hipStream_t st = ctx->stream;
  EH_HIP(hipSetDevice(ctx->device));
  DevBuf keys;
  uint64_t nkeys = 0;
  uint32_t n = 0;
  if (spec->kind == EH_GEN_ER) {
    n = spec->n;
    if (n < 2) { eh_set_error("ER needs n >= 2"); return EH_ERR_INVALID; }
    nkeys = spec->m;
    EH_CHECK(keys.alloc(8 * nkeys + 8));
    k_gen_er<<<grid_for(nkeys), 256, 0, st>>>(keys.as<uint64_t>(), nkeys, n,
                                              eh_stream_key(spec->seed, 0));
  } else if (spec->kind == EH_GEN_ER_EXACT) {
    n = spec->n;
    if (n < 2) { eh_set_error("ER needs n >= 2"); return EH_ERR_INVALID; }
    if (spec->m > (uint64_t)n * (n - 1) / 2) {
      eh_set_error("ER: %llu distinct pairs exceed the %llu a graph of %u vertices has", (unsigned long long)spec->m,
                   (unsigned long long)n * (n - 1) / 2, n);
      return EH_ERR_INVALID;
    }
    nkeys = spec->m;
    EH_CHECK(gen_er_exact(ctx, n, nkeys, spec->seed, keys));
  } else if (spec->kind == EH_GEN_RMAT) {
    if (spec->scale < 1 || spec->scale > 31) { eh_set_error("RMAT scale out of range"); return EH_ERR_INVALID; }
    n = 1u << spec->scale;
    nkeys = (uint64_t)spec->edge_factor << spec->scale;
    uint32_t t0 = (uint32_t)(spec->a * 4294967296.0);
    uint32_t t1 = (uint32_t)((spec->a + spec->b) * 4294967296.0);
    uint32_t t2 = (uint32_t)((spec->a + spec->b + spec->c) * 4294967296.0);
    EH_CHECK(keys.alloc(8 * nkeys + 8));
    k_gen_rmat<<<grid_for(nkeys), 256, 0, st>>>(keys.as<uint64_t>(), nkeys, spec->scale, t0, t1, t2,
                                                eh_stream_key(spec->seed, 0));
  } else if (spec->kind == EH_GEN_CHUNG_LU) {
    n = spec->n;
    if (n < 2 || spec->gamma <= 1.0) { eh_set_error("Chung-Lu needs n >= 2 and gamma > 1"); return EH_ERR_INVALID; }
    if (spec->communities && (uint64_t)spec->communities * spec->community_size > n) {
      eh_set_error("communities do not fit in n");
      return EH_ERR_INVALID;
    }
    // weight table w_i = (i+1)^(-1/(gamma-1)) as a 32-bit fixed-point CDF
    std::vector<uint64_t> cdf(n);
    double ex = -1.0 / (spec->gamma - 1.0), total = 0.0, acc = 0.0;
    for (uint32_t i = 0; i < n; i++) total += std::pow((double)i + 1.0, ex);
    for (uint32_t i = 0; i < n; i++) {
      acc += std::pow((double)i + 1.0, ex);
      double q = acc / total;
      cdf[i] = (uint64_t)(q * 4294967296.0);
    }
    cdf[n - 1] = 4294967296ull;
    DevBuf dcdf;
    EH_CHECK(dcdf.alloc(8 * (size_t)n));
    EH_HIP(hipMemcpy(dcdf.p, cdf.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
    uint64_t mc = spec->m;
    uint64_t ncomm_slots = (uint64_t)spec->communities * spec->community_size * spec->community_size;
    nkeys = mc + ncomm_slots;
    EH_CHECK(keys.alloc(8 * nkeys + 8));
    k_gen_chung_lu<<<grid_for(mc), 256, 0, st>>>(keys.as<uint64_t>(), mc, dcdf.as<uint64_t>(), n,
                                                 eh_stream_key(spec->seed, 0));
    if (ncomm_slots)
      k_gen_communities<<<grid_for(ncomm_slots), 256, 0, st>>>(keys.as<uint64_t>() + mc, n,
                                                               spec->communities,
                                                               spec->community_size, spec->seed);
    EH_HIP(hipStreamSynchronize(st));
  } else {
    eh_set_error("unknown generator kind %d", spec->kind);
    return EH_ERR_INVALID;
  }
  EH_HIP(hipGetLastError());
  return graph_canonicalize(ctx, n, keys, nkeys, spec->relabel, out, nullptr);
}

extern "C" eh_status eh_graph_from_edges(eh_ctx *ctx, uint32_t n, uint64_t m, const uint32_t *src,
                                         const uint32_t *dst, int relabel, eh_graph **out) {
  if (!ctx || !out || (m && (!src || !dst))) { eh_set_error("eh_graph_from_edges: null argument"); return EH_ERR_INVALID; }
  if (n == 0) { eh_set_error("eh_graph_from_edges: a graph needs at least one vertex"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(ctx->device));
  DevBuf ds, dd, keys;
  EH_CHECK(ds.alloc(4 * m + 4));
  EH_CHECK(dd.alloc(4 * m + 4));
  EH_CHECK(keys.alloc(8 * m + 8));
  if (m) {
    EH_HIP(hipMemcpy(ds.p, src, 4 * m, hipMemcpyHostToDevice));
    EH_HIP(hipMemcpy(dd.p, dst, 4 * m, hipMemcpyHostToDevice));
  }
  k_edges_to_keys<<<grid_for(m), 256, 0, ctx->stream>>>(keys.as<uint64_t>(), ds.as<uint32_t>(),
                                                        dd.as<uint32_t>(), m, n);
  EH_HIP(hipGetLastError());
  return graph_canonicalize(ctx, n, keys, m, relabel, out, nullptr);
}

extern "C" eh_status eh_graph_from_csr(eh_ctx *ctx, uint32_t n, const uint64_t *row_off,
                                       const uint32_t *nbrs, eh_graph **out) {
  if (!ctx || !out || !row_off) { eh_set_error("eh_graph_from_csr: null argument"); return EH_ERR_INVALID; }
  if (n == 0) { eh_set_error("eh_graph_from_csr: a graph needs at least one vertex"); return EH_ERR_INVALID; }
  if (row_off[0] != 0) { eh_set_error("row_off[0] must be 0"); return EH_ERR_INVALID; }
  for (uint32_t i = 0; i < n; i++)
    if (row_off[i + 1] < row_off[i]) { eh_set_error("row_off not monotone at %u", i); return EH_ERR_INVALID; }
  uint64_t entries = row_off[n];
  if (entries && !nbrs) { eh_set_error("eh_graph_from_csr: null nbrs"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(ctx->device));
  eh_graph *g = new eh_graph();
  g->ctx = ctx;
  g->n = n;
  g->entries = entries;
  eh_status s = EH_OK;
  uint32_t maxd = 0;
  for (uint32_t i = 0; i < n; i++) maxd = std::max<uint32_t>(maxd, (uint32_t)(row_off[i + 1] - row_off[i]));
  g->max_degree = maxd;
  if ((s = g->off.alloc(8 * ((size_t)n + 1))) != EH_OK || (s = g->nbr.alloc(4 * entries + 16)) != EH_OK ||
      (s = g->new2old.alloc(4 * (size_t)n + 4)) != EH_OK) {
    delete g;
    return s;
  }
  hipError_t e = hipMemcpy(g->off.p, row_off, 8 * ((size_t)n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && entries) e = hipMemcpy(g->nbr.p, nbrs, 4 * entries, hipMemcpyHostToDevice);
  if (e != hipSuccess) { eh_set_error("upload: %s", hipGetErrorString(e)); delete g; return EH_ERR_HIP; }
  k_iota<<<grid_for(n), 256, 0, ctx->stream>>>(g->new2old.as<uint32_t>(), n);
  DevBuf bad;
  if ((s = bad.alloc(4)) != EH_OK) { delete g; return s; }
  (void)hipMemsetAsync(bad.p, 0, 4, ctx->stream);
  k_validate_csr<<<grid_for((uint64_t)n * 64, 64), 64, 0, ctx->stream>>>(
      g->off.as<uint64_t>(), g->nbr.as<uint32_t>(), n, bad.as<uint32_t>());
  uint32_t hbad = 0;
  e = hipMemcpy(&hbad, bad.p, 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) { eh_set_error("validate: %s", hipGetErrorString(e)); delete g; return EH_ERR_HIP; }
  if (hbad) {
    eh_set_error("CSR rows must be strictly increasing, ids < n, without self-loops");
    delete g;
    return EH_ERR_INVALID;
  }
  *out = g;
  return EH_OK;
}

extern "C" eh_status eh_graph_relabel_by_degree(eh_graph *g) {
  if (!g) { eh_set_error("null graph"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(g->ctx->device));
  return graph_relabel(g);
}

extern "C" eh_status eh_graph_info_get(const eh_graph *g, eh_graph_info *out) {
  if (!g || !out) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  out->n = g->n;
  out->max_degree = g->max_degree;
  out->entries = g->entries;
  return EH_OK;
}

extern "C" eh_status eh_graph_download(const eh_graph *g, uint64_t *row_off, uint32_t *nbrs) {
  if (!g || !row_off || (g->entries && !nbrs)) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(g->ctx->device));
  EH_HIP(hipMemcpy(row_off, g->off.p, 8 * ((size_t)g->n + 1), hipMemcpyDeviceToHost));
  if (g->entries) EH_HIP(hipMemcpy(nbrs, g->nbr.p, 4 * g->entries, hipMemcpyDeviceToHost));
  return EH_OK;
}

extern "C" eh_status eh_graph_download_perm(const eh_graph *g, uint32_t *new2old) {
  if (!g || !new2old) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(g->ctx->device));
  if (g->n) EH_HIP(hipMemcpy(new2old, g->new2old.p, 4 * (size_t)g->n, hipMemcpyDeviceToHost));
  return EH_OK;
}

extern "C" void eh_graph_free(eh_graph *g) { delete g; }
