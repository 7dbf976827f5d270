// eh_clique.hip -- k-clique counting for any query_depth k >= 3 (n_clique<T,R>::run
// and apply_function, apps/n_clique.cpp:30-101).  The triangle (k = 3) and
// 4-clique (k = 4) passes have their own tuned kernels (eh_count.hip); this is
// the general recursion.
//
// Work unit: one kept entry (i, j), taken in chunks from an atomic counter by a
// wavefront.  Level 1 is row(i) in HBM; level t+1 = level t & row(v) for each v of
// level t, down to level k-1, which is only counted:
//     count += |level(k-2) & row(v)|   for every v in level(k-2)
// Levels 2..k-2 are materialised in the wave's LDS slice, or in a per-wave HBM
// spill region when (k-3) * |row(i)| does not fit.  The walk over the levels is an
// explicit stack (position and size per level) kept in LDS.
//
// One intersection S & row(v) runs wave-wide over the shorter side: lanes take 64
// elements at a time and test membership in the other (sorted) side by binary
// search; hits are compacted with ballot/mbcnt, so every level stays sorted.
// On an oriented matrix (every entry of row(v) is below v) only the prefix of S
// below v can meet row(v), and S is cut there first.
#include "eh_internal.hpp"

namespace {

constexpr uint32_t DEEP_MAX_K = 16;
constexpr uint32_t DEEP_WAVES = 4;         // waves per block
constexpr uint32_t DEEP_META = 2 * DEEP_MAX_K;
constexpr uint32_t DEEP_SLICE = 4096 - DEEP_META;  // LDS words per wave for the levels (64 KB per block)
constexpr uint32_t DEEP_CHUNK = 16;        // entries per queue pull

struct DeepArgs {
  const uint32_t *__restrict__ card;
  const uint32_t *__restrict__ aoff;   // row start in unbr, 16-byte units
  const uint32_t *__restrict__ unbr;
  const uint64_t *__restrict__ rowstart;  // exclusive scan of card (n + 1)
  uint32_t n;
  uint64_t m;                  // kept entries = work units
  uint32_t k;
  uint32_t oriented;
  uint32_t *next;
  unsigned long long *count;
  uint32_t *spill;             // per-wave level storage in HBM (null: LDS only)
  uint64_t spill_words;        // words per wave in spill
};

__device__ __forceinline__ uint32_t rfl(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// first position in sorted L[0..n) whose value is >= x
__device__ __forceinline__ uint32_t lower_bound(const uint32_t *L, uint32_t n, uint32_t x) {
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (L[lo + half] < x) { lo += half + 1; len -= half + 1; }
    else len = half;
  }
  return lo;
}

__device__ __forceinline__ bool contains(const uint32_t *L, uint32_t n, uint32_t x) {
  const uint32_t p = lower_bound(L, n, x);
  return p < n && L[p] == x;
}

// |S & V| over sorted S[0..ns), V[0..nv); writes the common elements in order to
// out when out is not null.  Wave-uniform arguments; returns a wave-uniform count.
__device__ uint32_t wave_isect(const uint32_t *S, uint32_t ns, const uint32_t *V, uint32_t nv, uint32_t *out,
                               uint32_t lane) {
  const bool over_s = ns <= nv;
  const uint32_t *P = over_s ? S : V;  // the side the lanes walk
  const uint32_t *Q = over_s ? V : S;  // the side searched
  const uint32_t np = over_s ? ns : nv, nq = over_s ? nv : ns;
  uint32_t total = 0;
  for (uint32_t c = 0; c < np; c += 64) {
    const uint32_t t = c + lane;
    uint32_t x = 0;
    bool hit = false;
    if (t < np) {
      x = P[t];
      hit = contains(Q, nq, x);
    }
    const uint64_t b = __ballot(hit);
    if (out && hit) out[total + (uint32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] = x;
    total += (uint32_t)__popcll(b);
  }
  if (out) wave_sync();
  return total;
}

__global__ void __launch_bounds__(64 * DEEP_WAVES) k_cliques_deep(DeepArgs a) {
  extern __shared__ uint32_t lds[];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t *meta = lds + wv * (DEEP_META + DEEP_SLICE);
  uint32_t *pos = meta, *size = meta + DEEP_MAX_K;
  uint32_t *slice = meta + DEEP_META;
  const uint64_t wave_id = (uint64_t)blockIdx.x * DEEP_WAVES + wv;
  uint32_t *spill = a.spill ? a.spill + wave_id * a.spill_words : nullptr;
  const uint32_t k = a.k;
  uint64_t acc = 0;
  for (;;) {
    uint32_t e0 = 0;
    if (lane == 0) e0 = atomicAdd(a.next, DEEP_CHUNK);
    e0 = rfl(e0);
    if ((uint64_t)e0 >= a.m) break;
    const uint64_t e1 = min<uint64_t>(a.m, (uint64_t)e0 + DEEP_CHUNK);
    for (uint64_t e = e0; e < e1; e++) {
      // row of entry e: the last i with rowstart[i] <= e
      uint32_t lo = 0, hi = a.n;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.rowstart[mid] <= e) lo = mid; else hi = mid;
      }
      const uint32_t i = rfl(lo);
      const uint32_t *L1 = a.unbr + (uint64_t)a.aoff[i] * 4;
      const uint32_t n1 = a.card[i];
      const uint32_t j = rfl(L1[e - a.rowstart[i]]);
      const uint32_t *Nj = a.unbr + (uint64_t)a.aoff[j] * 4;
      const uint32_t cj = a.card[j];
      const uint32_t cut1 = a.oriented ? lower_bound(L1, n1, j) : n1;
      if (k == 3) {
        acc += wave_isect(L1, cut1, Nj, cj, nullptr, lane);
        continue;
      }
      // levels 2..k-2, each at most n1 elements
      const bool in_lds = (uint64_t)(k - 3) * n1 <= DEEP_SLICE;
      uint32_t *levels = in_lds ? slice : spill;
      auto level = [&](uint32_t t) { return levels + (uint64_t)(t - 2) * n1; };
      const uint32_t n2 = wave_isect(L1, cut1, Nj, cj, level(2), lane);
      if (lane == 0) { pos[2] = 0; size[2] = n2; }
      wave_sync();
      uint32_t t = 2;
      while (t >= 2) {
        const uint32_t p = pos[t], sz = size[t];
        if (p >= sz) {
          t--;
          if (t >= 2 && lane == 0) pos[t] = pos[t] + 1;
          wave_sync();
          continue;
        }
        const uint32_t *Lt = level(t);
        const uint32_t v = Lt[p];
        const uint32_t *Nv = a.unbr + (uint64_t)a.aoff[v] * 4;
        const uint32_t cv = a.card[v];
        const uint32_t cut = a.oriented ? lower_bound(Lt, sz, v) : sz;
        if (t == k - 2) {
          acc += wave_isect(Lt, cut, Nv, cv, nullptr, lane);
          if (lane == 0) pos[t] = p + 1;
          wave_sync();
        } else {
          const uint32_t nn = wave_isect(Lt, cut, Nv, cv, level(t + 1), lane);
          if (lane == 0) { pos[t + 1] = 0; size[t + 1] = nn; }
          wave_sync();
          t++;
        }
      }
    }
  }
  if (lane == 0 && acc) atomicAdd(a.count, (unsigned long long)acc);
}

__global__ void k_card_padded(const uint32_t *card, uint32_t n, uint32_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = i < n ? card[i] : 0u;
}

__global__ void k_zero_u64_one(unsigned long long *p) { *p = 0ull; }

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = (work + block - 1) / block;
  if (g == 0) g = 1;
  return (unsigned)(g > 65536 ? 65536 : g);
}

}  // namespace

eh_status clique_count_deep(eh_matrix *m, uint32_t k, uint64_t *count, double *ms) {
  if (k < 3 || k > DEEP_MAX_K) {
    eh_set_error("eh_clique_count: k=%u outside [3, %u]", k, DEEP_MAX_K);
    return EH_ERR_UNSUPPORTED;
  }
  eh_ctx *ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t n = m->n;
  DevBuf cp, rs, next, cnt, spill;
  EH_CHECK(cp.alloc(4ull * n + 8));
  EH_CHECK(rs.alloc(8ull * n + 16));
  EH_CHECK(next.alloc(16));
  EH_CHECK(cnt.alloc(16));
  k_card_padded<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(m->dcard.as<uint32_t>(), n, cp.as<uint32_t>());
  EH_CHECK(prim_exscan_u32_to_u64(ctx, cp.as<uint32_t>(), rs.as<uint64_t>(), (uint64_t)n + 1));
  const size_t lds_bytes = 4ull * DEEP_WAVES * (DEEP_META + DEEP_SLICE);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cliques_deep, 64 * DEEP_WAVES, lds_bytes) !=
          hipSuccess || per_cu < 1)
    per_cu = 1;
  const uint64_t blocks = (uint64_t)per_cu * (ctx->num_cus ? ctx->num_cus : 256);
  DeepArgs a;
  a.card = m->dcard.as<uint32_t>();
  a.aoff = m->daoff.as<uint32_t>();
  a.unbr = m->dunbr.as<uint32_t>();
  a.rowstart = rs.as<uint64_t>();
  a.n = n;
  a.m = m->card;
  a.k = k;
  a.oriented = m->filter == EH_FILTER_LOWER ? 1u : 0u;
  a.next = next.as<uint32_t>();
  a.count = cnt.as<unsigned long long>();
  a.spill = nullptr;
  a.spill_words = 0;
  if (a.m > 0xFFFFFFFFull - DEEP_CHUNK) {
    eh_set_error("eh_clique_count: more than 2^32 kept entries for k=%u", k);
    return EH_ERR_UNSUPPORTED;
  }
  if (k > 3 && (uint64_t)(k - 3) * m->max_row > DEEP_SLICE) {
    a.spill_words = (uint64_t)(k - 3) * m->max_row;
    EH_CHECK(spill.alloc(4ull * a.spill_words * blocks * DEEP_WAVES));
    a.spill = spill.as<uint32_t>();
  }
  k_zero_u64_one<<<1, 1, 0, st>>>(reinterpret_cast<unsigned long long *>(next.p));
  k_zero_u64_one<<<1, 1, 0, st>>>(a.count);
  EH_HIP(hipEventRecord(ctx->ev0, st));
  hipLaunchKernelGGL(k_cliques_deep, dim3((unsigned)blocks), dim3(64 * DEEP_WAVES), lds_bytes, st, a);
  EH_HIP(hipGetLastError());
  EH_HIP(hipEventRecord(ctx->ev1, st));
  EH_HIP(hipEventSynchronize(ctx->ev1));
  float t = 0;
  EH_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  if (ms) *ms = t;
  unsigned long long h = 0;
  EH_HIP(hipMemcpy(&h, a.count, 8, hipMemcpyDeviceToHost));
  *count = h;
  return EH_OK;
}
