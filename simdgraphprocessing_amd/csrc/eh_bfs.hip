// eh_bfs.hip -- n_path's breadth-first search (apps/n_path.cpp:27-87) on the device.
//
// One round of n_path's loop is two passes over HBM-resident state:
//   union    the frontier's rows OR-ed into a `next` bitmap (ops::set_union,
//            src/set/ops/union.hpp:6-127): a bitset row contributes its 64-bit words
//            (one atomicOr per non-zero word), any other row its canonical uint list
//            (one bit per element, skipped when the vertex is already visited); one
//            wave per frontier vertex
//   collect  per 64-bit word, new = next & ~visited (ops::set_difference,
//            src/set/ops/difference.hpp:5-84), visited |= new, next = 0; the new
//            frontier is written in id order from a per-block popcount scan
// The host reads the new frontier's size after each round (one 8-byte copy) and stops
// as n_path does: an empty frontier, or path_length == max_depth.  Every word of the
// bitmaps belongs to exactly one thread in the collect pass, so only the union pass
// needs atomics.
#include "eh_internal.hpp"

namespace {

constexpr int BFS_THREADS = 256;

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = ceil_div(work ? work : 1, block);
  return (unsigned)(g > 65536 ? 65536 : g);
}

__global__ void k_bfs_init(unsigned long long *visited, uint32_t *frontier, uint32_t start) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    visited[start >> 6] = 1ull << (start & 63u);
    frontier[0] = start;
  }
}

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
k_bfs_count(const unsigned long long *next, const unsigned long long *visited, uint64_t nwords, uint32_t *bcount) {
  __shared__ uint32_t sh[BFS_THREADS / 64];
  const uint64_t w = (uint64_t)blockIdx.x * BFS_THREADS + threadIdx.x;
  const uint32_t c = w < nwords ? (uint32_t)__popcll(next[w] & ~visited[w]) : 0u;
  uint32_t tot;
  (void)block_excl_scan(c, sh, &tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}

// write the new vertices in id order, fold them into visited, clear next
__global__ void __launch_bounds__(BFS_THREADS)
k_bfs_write(unsigned long long *next, unsigned long long *visited, uint64_t nwords, const uint64_t *boff,
            uint32_t *frontier) {
  __shared__ uint32_t sh[BFS_THREADS / 64];
  const uint64_t w = (uint64_t)blockIdx.x * BFS_THREADS + threadIdx.x;
  unsigned long long nw = 0;
  if (w < nwords) {
    nw = next[w] & ~visited[w];
    visited[w] |= nw;
    next[w] = 0ull;
  }
  uint32_t tot;
  uint64_t pos = boff[blockIdx.x] + block_excl_scan((uint32_t)__popcll(nw), sh, &tot);
  while (nw) {
    frontier[pos++] = (uint32_t)(64ull * w + (uint64_t)__builtin_ctzll(nw));
    nw &= nw - 1;
  }
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
                                 uint64_t *level_sizes, uint32_t level_cap, eh_bfs_info *info) {
  if (!m || !info || (level_cap && !level_sizes)) { eh_set_error("eh_bfs_path: null argument"); return EH_ERR_INVALID; }
  if (n_vertices == 0) n_vertices = m->n;
  if (n_vertices > m->n) {
    eh_set_error("eh_bfs_path: %u vertices but the matrix has %u rows", n_vertices, m->n);
    return EH_ERR_INVALID;
  }
  if (start >= n_vertices) { eh_set_error("eh_bfs_path: start %u outside [0, %u)", start, n_vertices); return EH_ERR_INVALID; }
  eh_ctx *ctx = m->ctx;
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // bitmaps over every row id (entries are < m->n by the build), so any row may be reached
  const uint64_t nwords = (uint64_t)m->n / 64 + 1;
  const uint64_t nblocks = ceil_div(nwords, BFS_THREADS);
  DevBuf visited, next, frontier, bcount, boff, reached;
  EH_CHECK(visited.alloc(8 * nwords + 64));
  EH_CHECK(next.alloc(8 * nwords + 64));
  EH_CHECK(frontier.alloc(4ull * m->n + 64));
  EH_CHECK(bcount.alloc(4 * nblocks + 16));
  EH_CHECK(boff.alloc(8 * nblocks + 16));
  EH_CHECK(reached.alloc(16));
  EH_HIP(hipMemsetAsync(visited.p, 0, 8 * nwords + 64, st));
  EH_HIP(hipMemsetAsync(next.p, 0, 8 * nwords + 64, st));
  EH_HIP(hipMemsetAsync(reached.p, 0, 16, st));
  EH_HIP(hipEventRecord(ctx->ev0, st));
  k_bfs_init<<<1, 64, 0, st>>>(visited.as<unsigned long long>(), frontier.as<uint32_t>(), start);
  uint64_t nf = 1;
  uint32_t path = 0, rounds = 0;
  for (;;) {
    const uint64_t waves_per_block = 4;
    k_bfs_union<<<grid_for(nf * 64, 64 * waves_per_block), 64 * waves_per_block, 0, st>>>(
        m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(), m->ddesc.as<RowDesc>(),
        m->dbsw.as<uint64_t>(), frontier.as<uint32_t>(), nf, visited.as<unsigned long long>(),
        next.as<unsigned long long>());
    k_bfs_count<<<(unsigned)nblocks, BFS_THREADS, 0, st>>>(next.as<unsigned long long>(),
                                                            visited.as<unsigned long long>(), nwords,
                                                            bcount.as<uint32_t>());
    EH_CHECK(prim_exscan_u32_to_u64(ctx, bcount.as<uint32_t>(), boff.as<uint64_t>(), nblocks + 1));
    k_bfs_write<<<(unsigned)nblocks, BFS_THREADS, 0, st>>>(next.as<unsigned long long>(),
                                                            visited.as<unsigned long long>(), nwords,
                                                            boff.as<uint64_t>(), frontier.as<uint32_t>());
    EH_HIP(hipGetLastError());
    EH_HIP(hipMemcpyAsync(&nf, boff.as<uint64_t>() + nblocks, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (rounds < level_cap) level_sizes[rounds] = nf;
    rounds++;
    if (nf == 0 || path >= max_depth) break;
    path++;
  }
  k_bfs_popcount<<<grid_for(nwords), 256, 0, st>>>(visited.as<unsigned long long>(), nwords,
                                                   reached.as<unsigned long long>());
  EH_HIP(hipEventRecord(ctx->ev1, st));
  unsigned long long r = 0;
  EH_HIP(hipMemcpyAsync(&r, reached.p, 8, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  info->path_length = path;
  info->rounds = rounds;
  info->reached = r;
  info->kernel_ms = ctx_elapsed_ms(ctx);
  return EH_OK;
}
