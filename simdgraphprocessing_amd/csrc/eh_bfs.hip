// eh_bfs.hip -- n_path's breadth-first search (apps/n_path.cpp:27-87) on the device.
//
// The reference grows the frontier top-down: every frontier row is OR-ed into the
// visited set (ops::set_union, src/set/ops/union.hpp:6-127) and the new frontier is the
// visited set minus the previous one (ops::set_difference, src/set/ops/difference.hpp:5-84).
// Here a round is one of two expansions followed by one collect pass, all over
// HBM-resident bitmaps (visited, next, and the current frontier as a bitmap and an id
// list):
//   top-down   a bitset row ORs its words into `next`, any other row sets one bit per
//              unvisited element; a narrow frontier (whose rows may be hub rows) is cut
//              into 256-unit chunks spread over all waves, a wide one takes a wave per row
//   bottom-up  one wave per 64 vertices: every unvisited vertex scans its in-row until it
//              meets a frontier vertex, and the wave stores its word of `next` (no atomics)
//   collect    per word new = next & ~visited, visited |= new, next = 0, the frontier bitmap
//              becomes new, and the new frontier is written in id order from a per-block
//              popcount scan, with the sum of its rows' lengths (m_f)
// Both expansions find exactly the unvisited vertices adjacent to the frontier, so the
// frontiers, level sizes and path length are those of the reference.  A round runs
// bottom-up when the caller allows it (symmetric rows, or in-rows beside the out-rows) and
// the frontier's rows hold more than 1/14 of the entries not yet explored.  The host reads
// the new frontier's size and m_f after each round (16 bytes) and stops as n_path does:
// an empty frontier, or path_length == max_depth.
#include <algorithm>

#include "eh_internal.hpp"

namespace {

constexpr int BFS_THREADS = 256;
constexpr uint32_t BFS_CHUNK = 256;  // units of one top-down work item of a narrow frontier
constexpr uint64_t BFS_WIDE = 4096;  // frontiers this wide expand one wave per vertex

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = ceil_div(work ? work : 1, block);
  return (unsigned)(g > 65536 ? 65536 : g);
}

__global__ void k_bfs_init(unsigned long long *visited, unsigned long long *fbm, uint32_t *frontier, uint32_t start) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    visited[start >> 6] = 1ull << (start & 63u);
    fbm[start >> 6] = 1ull << (start & 63u);
    frontier[0] = start;
  }
}

// units of a frontier row: 64-bit words for a bitset row, entries otherwise
__device__ __forceinline__ uint32_t row_units(const RowDesc &d, uint32_t card) {
  return (d.card && desc_type(d) == EH_BS) ? desc_aux(d) : card;
}

__global__ void k_bfs_chunks(const uint32_t *__restrict__ card, const RowDesc *__restrict__ desc,
                             const uint32_t *__restrict__ frontier, uint64_t nf, uint32_t *chunks) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = frontier[f];
    chunks[f] = (row_units(desc[v], card[v]) + BFS_CHUNK - 1) / BFS_CHUNK;
  }
}

// top-down over a wide frontier: one wave per frontier vertex
__global__ void __launch_bounds__(256)
k_bfs_union(const uint32_t *__restrict__ card, const uint32_t *__restrict__ aoff, const uint32_t *__restrict__ unbr,
            const RowDesc *__restrict__ desc, const uint64_t *__restrict__ bsw, const uint32_t *__restrict__ frontier,
            uint64_t nf, const unsigned long long *__restrict__ visited, unsigned long long *next) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t f = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; f < nf; f += nwaves) {
    const uint32_t v = frontier[f];
    const RowDesc d = desc[v];
    if (d.card && desc_type(d) == EH_BS) {
      const uint64_t *W = bsw + d.off;
      for (uint32_t k = lane; k < desc_aux(d); k += 64) {
        const uint64_t w = W[k];
        if (w & ~visited[d.w0 + k]) atomicOr(&next[d.w0 + k], (unsigned long long)w);
      }
    } else {
      const uint32_t *R = unbr + (uint64_t)aoff[v] * 4;
      const uint32_t c = card[v];
      for (uint32_t k = lane; k < c; k += 64) {
        const uint32_t x = R[k];
        const unsigned long long bit = 1ull << (x & 63u);
        if (!(visited[x >> 6] & bit)) atomicOr(&next[x >> 6], bit);
      }
    }
  }
}

// top-down over a narrow frontier (whose rows may be long): chunk c belongs to the
// frontier vertex f with coff[f] <= c < coff[f + 1]
__global__ void __launch_bounds__(256)
k_bfs_topdown(const uint32_t *__restrict__ card, const uint32_t *__restrict__ aoff, const uint32_t *__restrict__ unbr,
              const RowDesc *__restrict__ desc, const uint64_t *__restrict__ bsw, const uint32_t *__restrict__ frontier,
              uint64_t nf, const uint64_t *__restrict__ coff, const unsigned long long *__restrict__ visited,
              unsigned long long *next) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t total = coff[nf];
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < total; c += nwaves) {
    uint64_t lo = 0, len = nf;  // last f with coff[f] <= c
    while (len > 1) {
      const uint64_t half = len >> 1;
      lo = (coff[lo + half] <= c) ? lo + half : lo;
      len -= half;
    }
    const uint32_t v = frontier[lo];
    const RowDesc d = desc[v];
    const uint32_t k0 = (uint32_t)(c - coff[lo]) * BFS_CHUNK;
    if (d.card && desc_type(d) == EH_BS) {
      const uint64_t *W = bsw + d.off;
      const uint32_t e = min(k0 + BFS_CHUNK, desc_aux(d));
      for (uint32_t k = k0 + lane; k < e; k += 64) {
        const uint64_t w = W[k];
        if (w & ~visited[d.w0 + k]) atomicOr(&next[d.w0 + k], (unsigned long long)w);
      }
    } else {
      const uint32_t *R = unbr + (uint64_t)aoff[v] * 4;
      const uint32_t e = min(k0 + BFS_CHUNK, card[v]);
      for (uint32_t k = k0 + lane; k < e; k += 64) {
        const uint32_t x = R[k];
        const unsigned long long bit = 1ull << (x & 63u);
        if (!(visited[x >> 6] & bit)) atomicOr(&next[x >> 6], bit);
      }
    }
  }
}

// bottom-up: lane l of the wave for word w tests vertex 64 w + l (its row at in_base + v)
__global__ void __launch_bounds__(256)
k_bfs_bottomup(const uint32_t *__restrict__ card, const uint32_t *__restrict__ aoff, const uint32_t *__restrict__ unbr,
               uint32_t n_vertices, uint32_t in_base, const unsigned long long *__restrict__ visited,
               const unsigned long long *__restrict__ fbm, unsigned long long *next) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t nw = ((uint64_t)n_vertices + 63) / 64;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nw; w += nwaves) {
    const uint64_t v = 64 * w + lane;
    uint32_t found = 0;
    if (v < n_vertices && !((visited[w] >> lane) & 1ull)) {
      const uint32_t r = in_base + (uint32_t)v;
      const uint32_t *R = unbr + (uint64_t)aoff[r] * 4;
      const uint32_t c = card[r];
      for (uint32_t k = 0; k < c; k++) {
        const uint32_t u = R[k];
        if ((fbm[u >> 6] >> (u & 63u)) & 1ull) { found = 1; break; }
      }
    }
    const uint64_t m = __ballot(found);
    if (lane == 0 && m) next[w] = (unsigned long long)m;
  }
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < nw; w++) {
    before += (w < wave) ? sh[w] : 0u;
    all += sh[w];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// per block of BFS_THREADS words: how many vertices this round discovers there
__global__ void __launch_bounds__(BFS_THREADS)
k_bfs_count(const unsigned long long *next, const unsigned long long *visited, uint64_t nwords, uint32_t *bcount,
            unsigned long long *mf) {
  __shared__ uint32_t sh[BFS_THREADS / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) *mf = 0ull;  // filled by k_bfs_write
  const uint64_t w = (uint64_t)blockIdx.x * BFS_THREADS + threadIdx.x;
  const uint32_t c = w < nwords ? (uint32_t)__popcll(next[w] & ~visited[w]) : 0u;
  uint32_t tot;
  (void)block_excl_scan(c, sh, &tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}

// write the new vertices in id order, fold them into visited, make them the frontier
// bitmap, clear next, and add their row lengths to *mf
__global__ void __launch_bounds__(BFS_THREADS)
k_bfs_write(unsigned long long *next, unsigned long long *visited, unsigned long long *fbm, uint64_t nwords,
            const uint64_t *boff, const uint32_t *card, uint32_t *frontier, unsigned long long *mf) {
  __shared__ uint32_t sh[BFS_THREADS / 64];
  const uint64_t w = (uint64_t)blockIdx.x * BFS_THREADS + threadIdx.x;
  unsigned long long nw = 0;
  if (w < nwords) {
    nw = next[w] & ~visited[w];
    visited[w] |= nw;
    fbm[w] = nw;
    next[w] = 0ull;
  }
  uint32_t tot;
  uint64_t pos = boff[blockIdx.x] + block_excl_scan((uint32_t)__popcll(nw), sh, &tot);
  unsigned long long deg = 0;
  while (nw) {
    const uint32_t v = (uint32_t)(64ull * w + (uint64_t)__builtin_ctzll(nw));
    frontier[pos++] = v;
    deg += card[v];
    nw &= nw - 1;
  }
  for (int o = 32; o > 0; o >>= 1) deg += __shfl_xor(deg, o, 64);
  if ((threadIdx.x & 63u) == 0 && deg) atomicAdd(mf, deg);
}

__global__ void k_bfs_popcount(const unsigned long long *visited, uint64_t nwords, unsigned long long *out) {
  unsigned long long c = 0;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += (uint64_t)gridDim.x * blockDim.x)
    c += (unsigned long long)__popcll(visited[w]);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63u) == 0 && c) atomicAdd(out, c);
}

}  // namespace

extern "C" eh_status eh_bfs_path(eh_matrix *m, uint32_t n_vertices, uint32_t start, uint32_t max_depth,
                                 eh_bfs_mode mode, uint64_t *level_sizes, uint32_t level_cap, eh_bfs_info *info) {
  if (!m || !info || (level_cap && !level_sizes)) { eh_set_error("eh_bfs_path: null argument"); return EH_ERR_INVALID; }
  if (mode < EH_BFS_TOP_DOWN || mode > EH_BFS_IN_ROWS) { eh_set_error("eh_bfs_path: unknown mode %d", (int)mode); return EH_ERR_INVALID; }
  if (n_vertices == 0) n_vertices = mode == EH_BFS_IN_ROWS ? m->n / 2 : m->n;
  if (n_vertices > m->n || (mode == EH_BFS_IN_ROWS && 2ull * n_vertices > m->n)) {
    eh_set_error("eh_bfs_path: %u vertices (mode %d) but the matrix has %u rows", n_vertices, (int)mode, m->n);
    return EH_ERR_INVALID;
  }
  if (start >= n_vertices) { eh_set_error("eh_bfs_path: start %u outside [0, %u)", start, n_vertices); return EH_ERR_INVALID; }
  // bottom-up rounds find a vertex's frontier neighbours in its own row (or in-row): only a
  // matrix that kept every entry holds all of them
  if (mode != EH_BFS_TOP_DOWN && m->filter != EH_FILTER_ALL) {
    eh_set_error("eh_bfs_path: mode %d needs a matrix built with EH_FILTER_ALL (this one keeps a subset of "
                 "each row); use EH_BFS_TOP_DOWN", (int)mode);
    return EH_ERR_INVALID;
  }
  eh_ctx *ctx = m->ctx;
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // bitmaps over every row id (entries are < m->n by the build), so any row may be reached
  const uint64_t nwords = (uint64_t)m->n / 64 + 1;
  const uint64_t nblocks = ceil_div(nwords, BFS_THREADS);
  DevBuf visited, next, fbm, frontier, chunks, coff, bcount, boff, counters;
  EH_CHECK(visited.alloc(8 * nwords + 64));
  EH_CHECK(next.alloc(8 * nwords + 64));
  EH_CHECK(fbm.alloc(8 * nwords + 64));
  EH_CHECK(frontier.alloc(4ull * m->n + 64));
  EH_CHECK(chunks.alloc(4ull * m->n + 64));
  EH_CHECK(coff.alloc(8ull * m->n + 64));
  EH_CHECK(bcount.alloc(4 * nblocks + 16));
  EH_CHECK(boff.alloc(8 * nblocks + 16));
  EH_CHECK(counters.alloc(64));  // [0] m_f of the new frontier, [1] reached
  EH_HIP(hipMemsetAsync(visited.p, 0, 8 * nwords + 64, st));
  EH_HIP(hipMemsetAsync(next.p, 0, 8 * nwords + 64, st));
  EH_HIP(hipMemsetAsync(fbm.p, 0, 8 * nwords + 64, st));
  EH_HIP(hipMemsetAsync(counters.p, 0, 64, st));
  unsigned long long *mf = counters.as<unsigned long long>();
  // explored-edge bookkeeping for the direction choice (rows [0, n_vertices))
  uint32_t start_deg = 0;
  EH_HIP(hipMemcpyAsync(&start_deg, m->dcard.as<uint32_t>() + start, 4, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  const uint64_t total_edges = mode == EH_BFS_IN_ROWS ? m->card / 2 : m->card;
  const uint32_t in_base = mode == EH_BFS_IN_ROWS ? n_vertices : 0u;
  EH_HIP(hipEventRecord(ctx->ev0, st));
  k_bfs_init<<<1, 64, 0, st>>>(visited.as<unsigned long long>(), fbm.as<unsigned long long>(), frontier.as<uint32_t>(),
                               start);
  uint64_t nf = 1, frontier_edges = start_deg, explored = 0;
  uint32_t path = 0, rounds = 0, bottom_up = 0;
  for (;;) {
    explored += frontier_edges;
    const uint64_t unexplored = total_edges > explored ? total_edges - explored : 0;
    const bool bu = mode != EH_BFS_TOP_DOWN && 14 * frontier_edges > unexplored;
    if (bu) {
      bottom_up++;
      k_bfs_bottomup<<<grid_for(ceil_div(n_vertices, 64) * 64), 256, 0, st>>>(
          m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(), n_vertices, in_base,
          visited.as<unsigned long long>(), fbm.as<unsigned long long>(), next.as<unsigned long long>());
    } else if (nf >= BFS_WIDE) {
      k_bfs_union<<<grid_for(nf * 64), 256, 0, st>>>(
          m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(), m->ddesc.as<RowDesc>(),
          m->dbsw.as<uint64_t>(), frontier.as<uint32_t>(), nf, visited.as<unsigned long long>(),
          next.as<unsigned long long>());
    } else {
      k_bfs_chunks<<<grid_for(nf), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->ddesc.as<RowDesc>(),
                                                 frontier.as<uint32_t>(), nf, chunks.as<uint32_t>());
      EH_CHECK(prim_exscan_u32_to_u64(ctx, chunks.as<uint32_t>(), coff.as<uint64_t>(), nf + 1));
      // the chunk count is on the device: a grid of at least 8 blocks per CU, waves loop over chunks
      const unsigned td_grid = std::max<unsigned>(grid_for(nf * 64), 8u * (unsigned)(ctx->num_cus ? ctx->num_cus : 256));
      k_bfs_topdown<<<std::min<unsigned>(td_grid, 65536u), 256, 0, st>>>(
          m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(), m->ddesc.as<RowDesc>(),
          m->dbsw.as<uint64_t>(), frontier.as<uint32_t>(), nf, coff.as<uint64_t>(), visited.as<unsigned long long>(),
          next.as<unsigned long long>());
    }
    k_bfs_count<<<(unsigned)nblocks, BFS_THREADS, 0, st>>>(next.as<unsigned long long>(),
                                                            visited.as<unsigned long long>(), nwords,
                                                            bcount.as<uint32_t>(), mf);
    EH_CHECK(prim_exscan_u32_to_u64(ctx, bcount.as<uint32_t>(), boff.as<uint64_t>(), nblocks + 1));
    k_bfs_write<<<(unsigned)nblocks, BFS_THREADS, 0, st>>>(
        next.as<unsigned long long>(), visited.as<unsigned long long>(), fbm.as<unsigned long long>(), nwords,
        boff.as<uint64_t>(), m->dcard.as<uint32_t>(), frontier.as<uint32_t>(), mf);
    EH_HIP(hipGetLastError());
    uint64_t h[2] = {0, 0};
    EH_HIP(hipMemcpyAsync(&h[0], boff.as<uint64_t>() + nblocks, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipMemcpyAsync(&h[1], mf, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    nf = h[0];
    frontier_edges = h[1];
    if (rounds < level_cap) level_sizes[rounds] = nf;
    rounds++;
    if (nf == 0 || path >= max_depth) break;
    path++;
  }
  k_bfs_popcount<<<grid_for(nwords), 256, 0, st>>>(visited.as<unsigned long long>(), nwords, mf + 1);
  EH_HIP(hipEventRecord(ctx->ev1, st));
  unsigned long long r = 0;
  EH_HIP(hipMemcpyAsync(&r, mf + 1, 8, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  info->path_length = path;
  info->rounds = rounds;
  info->reached = r;
  info->kernel_ms = ctx_elapsed_ms(ctx);
  info->bottom_up = bottom_up;
  info->reserved = 0;
  return EH_OK;
}
