// eh_count.hip -- triangle and 4-clique counting kernels for gfx950 (wave64).
//
// Work unit: a task = (source row i, j-range [jb, je) of row(i), graph id, staging kind).
// One wavefront runs one task at a time, pulling tasks from an atomic queue
// (persistent grid):
//   1. stage A = row(i) into the wave's LDS slice, in one of four kinds:
//        BITMAP  64-bit-word bitmap over [word(first), word(last)] plus the list
//                (the reference bitset's word span, bitset.hpp:65-96)
//        LIST    the sorted uint32 list (binary-search membership)
//        PSDIR   the list plus a directory of 16-bit prefix partitions
//                (pshort partitions, pshort.hpp:46-78)
//        GLOBAL  nothing staged; membership by binary search in HBM
//   2. for each j of the range, intersect A with B = row(j) read from HBM in
//      B's own layout (uint / pshort / bitset); 64 lanes split B's elements or
//      words and test membership against A, counts reduce per wave.
// Triangle counting follows apps/undirected_triangle_counting.cpp:39-61 (sum of
// |A & B| over kept (i,j)); the 4-clique count follows apps/n_clique.cpp:30-50
// with query_depth 4: C = A & row(j) is materialised in LDS and every k in C
// adds |C & row(k)|.
//
// Pair semantics follow the reference dispatcher (intersection.hpp:2066-2118):
//   U x U    lanes over B, binary search in A     (set_intersect_standard/galloping)
//   U x BS   lanes over the uint side probing bits (:1023-1058)
//   BS x BS  64-bit AND + popcount over the overlapping words (:880-955)
//   PS x PS  prefix match, then 16-bit runs        (:829-878)
//   U x PS   prefix-aligned run search             (:1062-1173)
//   PS x BS  pshort elements probing bits          (:956-1017)
//   BN x BN, NT x NT  blocks AND-ed with A's bitmap, new_type's uint part streamed
//                                                  (:1338-1413,1555-1649)
// A bitset_new / new_type B row facing a non-bitmap A is read through its
// canonical uint list, which every row keeps.
#include <mutex>
#include <type_traits>

#include "eh_internal.hpp"

namespace {

constexpr uint32_t KIND_BITMAP = 0, KIND_LIST = 1, KIND_PSDIR = 2, KIND_GLOBAL = 3, KIND_HASH = 4;
constexpr uint32_t KIND_SHIFT = 28;
constexpr uint32_t HASH_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t GID_MASK = (1u << KIND_SHIFT) - 1;

struct CountArgs {
  const uint32_t *__restrict__ card;
  const uint32_t *__restrict__ aoff;
  const uint32_t *__restrict__ unbr;
  const RowDesc *__restrict__ desc;
  const uint64_t *__restrict__ bsw;
  const uint16_t *__restrict__ psd;
  const uint4 *__restrict__ tasks;
  uint64_t ntasks;
  uint32_t *next;
  unsigned long long *counts;
  uint32_t lds_words;
  uint32_t tuned;     // flattened per-pair schedule (else the reference dispatch)
  const uint32_t *graph_of;         // per-row graph id of batch arenas (null: one graph)
  unsigned long long *task_cycles;  // optional: shader clock cycles spent per task
  uint64_t task_base;               // index of tasks[0] in the whole schedule
  uint32_t stream_ratio;            // tuned: stream a bitset B row's uint list when
                                    // 16 * |B| < stream_ratio * probes (0: always probe)
  const uint32_t *__restrict__ bid;  // block ids per slot (bitset_new / new_type rows)
  const uint4 *__restrict__ bwd;     // the two 64-bit words of each block slot
  const uint32_t *__restrict__ ntu;  // new_type uint parts
  uint32_t blocks;                   // the matrix has bitset_new / new_type rows (selects
                                     // the kernels with the block passes)
  uint32_t grab;                     // tasks a wave takes per queue atomic (one address is
                                     // shared by every wave: one task per atomic serialises)
  const uint32_t *jstart;            // B-stationary on: per row, the entries before this
                                     // position are counted by k_bstat (null: none)
  const uint4 *lo16;                 // B-stationary: those entries as 16-bit ids (they are all
                                     // below the split, at most 65536), padded to 8 per vector
  uint32_t *err;                     // device error word: a lane whose hash probe or insert
                                     // walked a whole table ORs ERR_HASH_FULL into it once, when
                                     // its task or kernel ends (the host reports it)
  uint32_t *xq;                      // short rows of a batch arena: bounds of 8 ranges of the
                                     // row list (xq[0..8]) and their queue heads (XQ_HEADS on);
                                     // null: one queue at next
  unsigned long long *counts3;       // k_small_cliques4 with the triangle pass fused in: per-graph
                                     // (or total) triangle counters for its local-edge hits (null:
                                     // 4-cliques only)
};

// bits of the device error word (CountArgs::err)
constexpr uint32_t ERR_HASH_FULL = 1u;

__device__ __forceinline__ uint32_t rfl(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t rdl(uint32_t x, uint32_t l) { return __builtin_amdgcn_readlane(x, l); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t lane, uint32_t *total) {
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// membership in a sorted array: largest position whose value <= x, then compare
template <class P>
__device__ __forceinline__ uint32_t sorted_contains(P L, uint32_t n, uint32_t x) {
  if (n == 0) return 0;
  uint32_t base = 0, len = n;
  while (len > 1) {
    uint32_t half = len >> 1;
    base = (L[base + half] <= x) ? base + half : base;
    len -= half;
  }
  return L[base] == x;
}

// A-side state of one staged set (row(i) for triangles; also C for 4-cliques)
struct AState {
  uint32_t *bm;           // LDS bitmap (u32 words); bit x at bm[(x >> 5) - base32]
  uint32_t base32;        // 2 * aw0
  uint32_t aw0, anw;      // 64-bit word span
  uint32_t lo, span;      // first element, last - first
  const uint32_t *list;   // LDS list (or HBM list for KIND_GLOBAL)
  uint32_t n;
  const uint32_t *dpre;   // PSDIR: partition prefixes (P), starts (P + 1)
  const uint32_t *dstart;
  uint32_t P;
  uint32_t *ht;           // HASH: LDS table of 2^hlg slots
  uint32_t hlg;
  mutable uint32_t bad;   // HASH: this lane walked a whole table (kept in a register and
                          // reported when the task ends, so no store sits in the probe loops)
};

template <uint32_t KIND>
__device__ __forceinline__ uint32_t memb(const AState &A, uint32_t x) {
  if (x - A.lo > A.span) return 0;
  if constexpr (KIND == KIND_BITMAP) {
    return (A.bm[(x >> 5) - A.base32] >> (x & 31u)) & 1u;
  } else if constexpr (KIND == KIND_PSDIR) {
    const uint32_t hi = x >> 16;
    uint32_t base = 0, len = A.P;
    while (len > 1) {
      uint32_t half = len >> 1;
      base = (A.dpre[base + half] <= hi) ? base + half : base;
      len -= half;
    }
    if (A.dpre[base] != hi) return 0;
    const uint32_t s = A.dstart[base], e = A.dstart[base + 1];
    return sorted_contains(A.list + s, e - s, x);
  } else if constexpr (KIND == KIND_HASH) {
    // a table of 2^hlg slots holds at most half that many ids, and it is only probed when
    // every insert found a slot (build_hash), so the walk meets a free slot
    const uint32_t mask = (1u << A.hlg) - 1u;
    uint32_t s = ((x * 2654435761u) >> (32u - A.hlg)) & mask;
    for (;;) {
      const uint32_t v = A.ht[s];
      if (v == x) return 1;
      if (v == HASH_EMPTY) return 0;
      s = (s + 1u) & mask;
    }
  } else {
    return sorted_contains(A.list, A.n, x);
  }
}

template <uint32_t KIND, bool W>
__device__ __forceinline__ uint64_t hit_value(const AState &A, const CountArgs &a, uint32_t x) {
  uint32_t h = memb<KIND>(A, x);
  if constexpr (W) return h ? (uint64_t)a.card[x] : 0ull;
  else return h;
}

// ---------------------------------------------------------------- count-only intersections

// |A & B| for a uint B row: 16-byte loads, four probes per lane per step
template <uint32_t KIND, bool W>
__device__ __forceinline__ uint64_t isect_b_uint(const AState &A, const CountArgs &a, const RowDesc &dj,
                                                 uint32_t lane) {
  const uint4 *B = reinterpret_cast<const uint4 *>(a.unbr) + dj.off;
  const uint32_t n4 = (dj.card + 3u) >> 2;
  uint64_t c = 0;
  for (uint32_t v = lane; v < n4; v += 64) {
    const uint4 x = B[v];
    c += hit_value<KIND, W>(A, a, x.x) + hit_value<KIND, W>(A, a, x.y) + hit_value<KIND, W>(A, a, x.z) +
         hit_value<KIND, W>(A, a, x.w);
  }
  return c;
}

// locate A's partition for prefix `pre` (uniform); returns false when absent
__device__ __forceinline__ bool psdir_find(const AState &A, uint32_t pre, uint32_t *s, uint32_t *e) {
  uint32_t base = 0, len = A.P;
  while (len > 1) {
    uint32_t half = len >> 1;
    base = (A.dpre[base + half] <= pre) ? base + half : base;
    len -= half;
  }
  if (A.P == 0 || A.dpre[base] != pre) return false;
  *s = A.dstart[base];
  *e = A.dstart[base + 1];
  return true;
}

template <uint32_t KIND, bool W>
__device__ __forceinline__ uint64_t isect_b_pshort(const AState &A, const CountArgs &a, const RowDesc &dj,
                                                   uint32_t lane) {
  const uint16_t *S = a.psd + (uint64_t)dj.off * 4;
  const uint32_t len = desc_aux(dj);
  const uint32_t a_hi = A.lo + A.span;
  uint64_t c = 0;
  uint32_t pos = 0;
  while (pos < len) {
    const uint32_t pre = rfl(S[pos]);
    const uint32_t cnt = rfl(S[pos + 1]) + 1u;
    const uint32_t base = pre << 16;
    if (base + 0xFFFFu >= A.lo && base <= a_hi) {
      if constexpr (KIND == KIND_PSDIR) {
        // PS x PS: only partitions with a matching prefix are read
        uint32_t s, e;
        if (psdir_find(A, pre, &s, &e)) {
          for (uint32_t k = lane; k < cnt; k += 64) {
            const uint32_t x = base | S[pos + 2 + k];
            uint32_t h = sorted_contains(A.list + s, e - s, x);
            if constexpr (W) c += h ? (uint64_t)a.card[x] : 0ull;
            else c += h;
          }
        }
      } else {
        for (uint32_t k = lane; k < cnt; k += 64) c += hit_value<KIND, W>(A, a, base | S[pos + 2 + k]);
      }
    }
    pos += cnt + 2u;
  }
  return c;
}

template <uint32_t KIND, bool W>
__device__ __forceinline__ uint64_t isect_b_bitset(const AState &A, const CountArgs &a, const RowDesc &dj,
                                                   uint32_t lane) {
  const uint32_t bw0 = dj.w0, bnw = desc_aux(dj);
  const uint64_t *BW = a.bsw + dj.off;
  uint64_t c = 0;
  if constexpr (KIND == KIND_BITMAP) {
    const uint32_t s = max(A.aw0, bw0);
    const uint32_t e = min(A.aw0 + A.anw, bw0 + bnw);
    for (uint32_t w = s + lane; w < e; w += 64) {
      const uint32_t i2 = 2u * (w - A.aw0);
      const uint64_t aw = (uint64_t)A.bm[i2] | ((uint64_t)A.bm[i2 + 1] << 32);
      uint64_t r = aw & BW[w - bw0];
      if constexpr (W) {
        while (r) {
          const uint32_t b = (uint32_t)__builtin_ctzll(r);
          c += a.card[64u * w + b];
          r &= r - 1;
        }
      } else {
        c += (uint64_t)__popcll(r);
      }
    }
  } else {
    // probe B's bits with A's elements
    for (uint32_t k = lane; k < A.n; k += 64) {
      const uint32_t x = A.list[k];
      const uint32_t w = (x >> 6) - bw0;
      if (w < bnw && ((BW[w] >> (x & 63u)) & 1ull)) {
        if constexpr (W) c += a.card[x];
        else c += 1;
      }
    }
  }
  return c;
}

// one 128-value block (slot s) AND-ed with A's LDS bitmap over their common words
template <bool W>
__device__ __forceinline__ uint64_t block_and(const AState &A, const CountArgs &a, uint32_t s) {
  const uint32_t id = a.bid[s];
  const uint4 q = a.bwd[s];
  uint64_t c = 0;
#pragma unroll
  for (uint32_t h = 0; h < 2; h++) {
    const uint32_t w = 2u * id + h;
    if (w - A.aw0 >= A.anw) continue;
    const uint32_t i2 = 2u * (w - A.aw0);
    uint64_t r = ((uint64_t)A.bm[i2] | ((uint64_t)A.bm[i2 + 1] << 32)) &
                 (h ? ((uint64_t)q.z | ((uint64_t)q.w << 32)) : ((uint64_t)q.x | ((uint64_t)q.y << 32)));
    if constexpr (W) {
      while (r) {
        c += a.card[64u * w + (uint32_t)__builtin_ctzll(r)];
        r &= r - 1;
      }
    } else {
      c += (uint64_t)__popcll(r);
    }
  }
  return c;
}

// |A & B| for a bitset_new / new_type B row against A's LDS bitmap: lanes over B's
// blocks, then (new_type) over its uint part
template <bool W>
__device__ __forceinline__ uint64_t isect_b_blocks(const AState &A, const CountArgs &a, const RowDesc &dj,
                                                   uint32_t lane) {
  const uint32_t t = desc_type(dj), nb = desc_aux(dj);
  const uint32_t first = dj.w0 + (t == EH_NT ? EH_NT_HEADER : 0u);
  uint64_t c = 0;
  for (uint32_t b = lane; b < nb; b += 64) c += block_and<W>(A, a, first + b);
  if (t == EH_NT) {
    const uint32_t nu = a.bid[dj.w0], uoff = a.bid[dj.w0 + 1];
    const uint4 *U = reinterpret_cast<const uint4 *>(a.ntu) + uoff;
    for (uint32_t v = lane; v < ((nu + 3u) >> 2); v += 64) {
      const uint4 x = U[v];
      c += hit_value<KIND_BITMAP, W>(A, a, x.x) + hit_value<KIND_BITMAP, W>(A, a, x.y) +
           hit_value<KIND_BITMAP, W>(A, a, x.z) + hit_value<KIND_BITMAP, W>(A, a, x.w);
    }
  }
  return c;
}

template <uint32_t KIND, bool W>
__device__ __forceinline__ uint64_t isect_count(const AState &A, const CountArgs &a, const RowDesc &dj,
                                                uint32_t lane) {
  if (dj.card == 0) return 0;
  const uint32_t t = desc_type(dj);
  if constexpr (KIND == KIND_BITMAP) {
    if (is_block_type(t)) return isect_b_blocks<W>(A, a, dj, lane);
  }
  if (is_uint_type(t)) return isect_b_uint<KIND, W>(A, a, dj, lane);
  if (t == EH_PS) return isect_b_pshort<KIND, W>(A, a, dj, lane);
  return isect_b_bitset<KIND, W>(A, a, dj, lane);
}

// ---------------------------------------------------------------- materialising intersection

// C = A & B into LDS list `out` (sorted); returns |C| (uniform)
template <uint32_t KIND>
__device__ __forceinline__ uint32_t isect_write(const AState &A, const CountArgs &a, const RowDesc &dj,
                                                uint32_t lane, uint32_t *out) {
  if (dj.card == 0) return 0;
  const uint32_t t = desc_type(dj);
  uint32_t nc = 0;
  if (is_uint_type(t)) {
    const uint32_t *B = a.unbr + (uint64_t)dj.off * 4;
    for (uint32_t k0 = 0; k0 < dj.card; k0 += 64) {
      const uint32_t k = k0 + lane;
      uint32_t x = 0, h = 0;
      if (k < dj.card) { x = B[k]; h = memb<KIND>(A, x); }
      const uint64_t m = __ballot(h);
      if (h) out[nc + mbcnt(m)] = x;
      nc += (uint32_t)__popcll(m);
    }
  } else if (t == EH_PS) {
    const uint16_t *S = a.psd + (uint64_t)dj.off * 4;
    const uint32_t len = desc_aux(dj);
    uint32_t pos = 0;
    while (pos < len) {
      const uint32_t pre = rfl(S[pos]);
      const uint32_t cnt = rfl(S[pos + 1]) + 1u;
      const uint32_t base = pre << 16;
      for (uint32_t k0 = 0; k0 < cnt; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint32_t x = 0, h = 0;
        if (k < cnt) { x = base | S[pos + 2 + k]; h = memb<KIND>(A, x); }
        const uint64_t m = __ballot(h);
        if (h) out[nc + mbcnt(m)] = x;
        nc += (uint32_t)__popcll(m);
      }
      pos += cnt + 2u;
    }
  } else {
    const uint32_t bw0 = dj.w0, bnw = desc_aux(dj);
    const uint64_t *BW = a.bsw + dj.off;
    if constexpr (KIND == KIND_BITMAP) {
      const uint32_t s = max(A.aw0, bw0);
      const uint32_t e = min(A.aw0 + A.anw, bw0 + bnw);
      for (uint32_t w0 = s; w0 < e; w0 += 64) {
        const uint32_t w = w0 + lane;
        uint64_t r = 0;
        if (w < e) {
          const uint32_t i2 = 2u * (w - A.aw0);
          r = ((uint64_t)A.bm[i2] | ((uint64_t)A.bm[i2 + 1] << 32)) & BW[w - bw0];
        }
        uint32_t tot;
        uint32_t p = nc + wave_excl_scan((uint32_t)__popcll(r), lane, &tot);
        while (r) {
          const uint32_t b = (uint32_t)__builtin_ctzll(r);
          out[p++] = 64u * w + b;
          r &= r - 1;
        }
        nc += tot;
      }
    } else {
      for (uint32_t k0 = 0; k0 < A.n; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint32_t x = 0, h = 0;
        if (k < A.n) {
          x = A.list[k];
          const uint32_t w = (x >> 6) - bw0;
          h = (w < bnw) ? (uint32_t)((BW[w] >> (x & 63u)) & 1ull) : 0u;
        }
        const uint64_t m = __ballot(h);
        if (h) out[nc + mbcnt(m)] = x;
        nc += (uint32_t)__popcll(m);
      }
    }
  }
  wave_sync();
  return nc;
}

// ---------------------------------------------------------------- staging

// Stage row(i) as `kind` into LDS slice L; returns the A state and the number
// of LDS words used.
__device__ __forceinline__ uint32_t stage_a(const CountArgs &a, uint32_t kind, uint32_t i, uint32_t lane,
                                            uint32_t *L, AState &A) {
  const uint32_t dA = a.card[i];
  const uint32_t *G = a.unbr + (uint64_t)a.aoff[i] * 4;
  const uint32_t lo = rfl(G[0]);
  const uint32_t hi = rfl(G[dA - 1]);
  A.lo = lo;
  A.span = hi - lo;
  A.n = dA;
  A.aw0 = lo >> 6;
  A.anw = (hi >> 6) - (lo >> 6) + 1u;
  A.base32 = 2u * A.aw0;
  A.P = 0;
  A.dpre = A.dstart = nullptr;
  A.ht = nullptr;
  A.bad = 0u;
  A.hlg = 0;
  A.bm = L;
  if (kind == KIND_GLOBAL) {
    A.list = G;
    return 0;
  }
  uint32_t used = 0;
  if (kind == KIND_BITMAP) {
    const uint32_t nw32 = 2u * A.anw;
    for (uint32_t w = lane; w < nw32; w += 64) L[w] = 0u;
    used = nw32;
  }
  uint32_t *list = L + used;
  for (uint32_t k = lane; k < dA; k += 64) list[k] = G[k];
  used += dA;
  wave_sync();
  if (kind == KIND_BITMAP) {
    for (uint32_t k = lane; k < dA; k += 64) {
      const uint32_t x = list[k];
      atomicOr(&L[(x >> 5) - A.base32], 1u << (x & 31u));
    }
  } else if (kind == KIND_PSDIR) {
    // directory of 16-bit prefixes: partition p starts where the prefix changes
    uint32_t *dpre = L + used;
    uint32_t np = 0;
    for (uint32_t k0 = 0; k0 < dA; k0 += 64) {
      const uint32_t k = k0 + lane;
      uint32_t f = 0;
      if (k < dA) f = (k == 0) || ((list[k] >> 16) != (list[k - 1] >> 16));
      np += (uint32_t)__popcll(__ballot(f));
    }
    uint32_t *dstart = dpre + np;
    uint32_t q = 0;
    for (uint32_t k0 = 0; k0 < dA; k0 += 64) {
      const uint32_t k = k0 + lane;
      uint32_t f = 0;
      if (k < dA) f = (k == 0) || ((list[k] >> 16) != (list[k - 1] >> 16));
      const uint64_t m = __ballot(f);
      if (f) {
        dpre[q + mbcnt(m)] = list[k] >> 16;
        dstart[q + mbcnt(m)] = k;
      }
      q += (uint32_t)__popcll(m);
    }
    if (lane == 0) dstart[np] = dA;
    A.P = np;
    A.dpre = dpre;
    A.dstart = dstart;
    used += 2u * np + 1u;
  }
  wave_sync();
  A.list = list;
  return used;
}

// C-state over a materialised LDS list (sorted) for the 4-clique inner level;
// for bitmap kinds the C bitmap shares A's word span.
template <uint32_t CKIND>
__device__ __forceinline__ void make_cstate(const AState &A, uint32_t *clist, uint32_t nc, uint32_t *cbm,
                                            uint32_t lane, AState &Cs) {
  Cs.list = clist;
  Cs.n = nc;
  Cs.lo = clist[0];
  Cs.span = clist[nc - 1] - clist[0];
  Cs.P = 0;
  Cs.dpre = Cs.dstart = nullptr;
  Cs.ht = nullptr;
  Cs.bad = 0u;
  Cs.hlg = 0;
  Cs.aw0 = A.aw0;
  Cs.anw = A.anw;
  Cs.base32 = A.base32;
  Cs.bm = cbm;
  if constexpr (CKIND == KIND_BITMAP) {
    const uint32_t nw32 = 2u * A.anw;
    for (uint32_t w = lane; w < nw32; w += 64) cbm[w] = 0u;
    wave_sync();
    for (uint32_t k = lane; k < nc; k += 64) {
      const uint32_t x = clist[k];
      atomicOr(&cbm[(x >> 5) - A.base32], 1u << (x & 31u));
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------- task bodies

template <uint32_t KIND, bool W>
__device__ uint64_t task_triangles(const CountArgs &a, const AState &A, uint32_t jb, uint32_t je,
                                   uint32_t lane) {
  uint64_t acc = 0;
  for (uint32_t jc = jb; jc < je; jc += 64) {
    const uint32_t nj = min(64u, je - jc);
    uint32_t jl = 0;
    RowDesc dl = {0u, 0u, 0u, 0u};
    if (lane < nj) {
      jl = A.list[jc + lane];
      dl = a.desc[jl];
    }
    for (uint32_t t = 0; t < nj; t++) {
      RowDesc dj;
      dj.card = rdl(dl.card, t);
      dj.tinfo = rdl(dl.tinfo, t);
      dj.w0 = rdl(dl.w0, t);
      dj.off = rdl(dl.off, t);
      acc += isect_count<KIND, W>(A, a, dj, lane);
    }
  }
  return acc;
}

template <uint32_t KIND>
__device__ uint64_t task_cliques4(const CountArgs &a, const AState &A, uint32_t jb, uint32_t je,
                                  uint32_t lane, uint32_t *free_lds) {
  constexpr uint32_t CKIND = (KIND == KIND_BITMAP) ? KIND_BITMAP : KIND_LIST;
  uint32_t *cbm = free_lds;
  uint32_t *clist = free_lds + ((CKIND == KIND_BITMAP) ? 2u * A.anw : 0u);
  uint64_t acc = 0;
  for (uint32_t jc = jb; jc < je; jc += 64) {
    const uint32_t nj = min(64u, je - jc);
    uint32_t jl = 0;
    RowDesc dl = {0u, 0u, 0u, 0u};
    if (lane < nj) {
      jl = A.list[jc + lane];
      dl = a.desc[jl];
    }
    for (uint32_t t = 0; t < nj; t++) {
      RowDesc dj;
      dj.card = rdl(dl.card, t);
      dj.tinfo = rdl(dl.tinfo, t);
      dj.w0 = rdl(dl.w0, t);
      dj.off = rdl(dl.off, t);
      const uint32_t nc = rfl(isect_write<KIND>(A, a, dj, lane, clist));
      if (nc == 0) continue;
      AState Cs;
      make_cstate<CKIND>(A, clist, nc, cbm, lane, Cs);
      for (uint32_t kc = 0; kc < nc; kc += 64) {
        const uint32_t nk = min(64u, nc - kc);
        RowDesc kl = {0u, 0u, 0u, 0u};
        if (lane < nk) kl = a.desc[clist[kc + lane]];
        for (uint32_t u = 0; u < nk; u++) {
          RowDesc dk;
          dk.card = rdl(kl.card, u);
          dk.tinfo = rdl(kl.tinfo, u);
          dk.w0 = rdl(kl.w0, u);
          dk.off = rdl(kl.off, u);
          acc += isect_count<CKIND, false>(Cs, a, dk, lane);
        }
      }
      wave_sync();
    }
  }
  return acc;
}

// ---------------------------------------------------------------- tuned path: flattened pair work
//
// For a chunk of up to 64 keys (the j's of a task, or the k's of a 4-clique
// level set) every pair (S, row(key)) is classified by the cheaper of
//   PROBE   S's elements inside B's word span test B's bitset bits (|S in span| items)
//   AND     S's LDS bitmap AND B's words over the overlapping span (overlap words)
//   STREAM  B's uint elements (16-byte vectors) test membership in S (|B|/4 items)
// and the items of all pairs of one kind are flattened over the 64 lanes, so a
// wave keeps 64 independent loads in flight instead of one pair at a time.
// pshort B rows are walked partition by partition, one pair after another.

constexpr uint32_t PASS_PROBE = 0, PASS_AND = 1, PASS_STREAM = 2, PASS_BLOCK = 3, PASS_NTSTREAM = 4,
                   PASS_STREAM16 = 5, PASS_SEARCH = 6;
// B-stationary pairs above the split (k_bshl): the longest B row, and the bins of consecutive
// B rows that form its tasks (a row weighs 64 plus the power of two at or above twice its ids
// at or above the split; a bin holds BSH_BIN_WEIGHT of weight, at most BSH_GROUP rows)
constexpr uint32_t BSH_MAX_D = 512, BSH_GROUP = 16, BSH_BIN_WEIGHT = 1024;
// flat_pass metadata: four rank-ordered arrays of 64 words and 64 slot-block start masks
constexpr uint32_t META_WORDS = 384;
constexpr uint32_t SMALL_META_WORDS = 260;

__device__ __forceinline__ uint32_t lds_lower_bound(const uint32_t *L, uint32_t n, uint32_t x) {
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (L[lo + half] < x) { lo += half + 1; len -= half + 1; }
    else len = half;
  }
  return lo;
}

// One element of a pass for a pair with metadata (v0, v1, v2):
//   PROBE   A's element S.list[v0 + r] tested in B's bitset words (bsw[v1 ..], first word v2)
//   AND     word v0 + r of A's LDS bitmap AND-ed with B's word (bsw[v1 ..], first word v2)
//   STREAM  B's elements unbr[4 (v1 + r) .. +4] tested against A's LDS structure
//   BLOCK   B's block slot v0 + r AND-ed with A's LDS bitmap
//   NTSTREAM  B's new_type uint part ntu[4 (v1 + r) .. +4] tested against A
//   SEARCH  A's element S.list[v0 + r] searched in B's sorted uint list (unbr[4 v1 ..], v2 ids)
// a lane's plain count in one pass adds at most 64 per element it visits, far below 2^32;
// weighted sums (card of every hit) need 64 bits
template <bool W>
using PassAcc = typename std::conditional<W, uint64_t, uint32_t>::type;

// the four ids of a streamed B vector tested against S; with LOCKSTEP a sorted list (in LDS,
// or in HBM for an unstaged row) is searched for all four together, so each step's four loads
// are in flight at once (B's padding 0xFFFFFFFF never equals an id)
template <uint32_t KIND, bool W, bool LOCKSTEP>
__device__ __forceinline__ PassAcc<W> stream_hits(const AState &S, const CountArgs &a, const uint4 x) {
  if constexpr ((KIND == KIND_LIST || KIND == KIND_GLOBAL) && !W && LOCKSTEP) {
    uint32_t len = S.n;
    if (len == 0) return 0;
    uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    while (len > 1) {
      const uint32_t half = len >> 1;
      const uint32_t v0 = S.list[b0 + half], v1 = S.list[b1 + half], v2 = S.list[b2 + half],
                     v3 = S.list[b3 + half];
      b0 = (v0 <= x.x) ? b0 + half : b0;
      b1 = (v1 <= x.y) ? b1 + half : b1;
      b2 = (v2 <= x.z) ? b2 + half : b2;
      b3 = (v3 <= x.w) ? b3 + half : b3;
      len -= half;
    }
    return (PassAcc<W>)((S.list[b0] == x.x) + (S.list[b1] == x.y) + (S.list[b2] == x.z) + (S.list[b3] == x.w));
  } else {
    return (PassAcc<W>)(hit_value<KIND, W>(S, a, x.x) + hit_value<KIND, W>(S, a, x.y) +
                        hit_value<KIND, W>(S, a, x.z) + hit_value<KIND, W>(S, a, x.w));
  }
}

// eight 16-bit ids against a bitmap that covers every id below the split from bit 0: no span
// test (ids at or past the pair's B row, and the padding, find zero bits)
__device__ __forceinline__ uint32_t bitmap16_hits(const uint32_t *bm, const uint4 x) {
  uint32_t h = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    // id w & 0xFFFF: word bits 5..15, bit w & 31; id w >> 16: word w >> 21, bit (w >> 16) & 31
    // (the bit-field extract reads only the low five bits of its offset)
    const uint32_t w = q == 0 ? x.x : q == 1 ? x.y : q == 2 ? x.z : x.w;
    h += __builtin_amdgcn_ubfe(bm[__builtin_amdgcn_ubfe(w, 5u, 11u)], w, 1u) +
         __builtin_amdgcn_ubfe(bm[w >> 21], w >> 16, 1u);
  }
  return h;
}

template <int PASS, uint32_t KIND, bool W, bool LOCKSTEP = false>
__device__ __forceinline__ PassAcc<W> pass_elem(const CountArgs &a, const AState &S, uint32_t v0, uint32_t v1,
                                                uint32_t v2, uint32_t r) {
  if constexpr (PASS == PASS_PROBE) {
    // the 32-bit half of B's 64-bit word that holds bit x
    const uint32_t x = S.list[v0 + r];
    const uint64_t w32 = 2ull * (uint64_t)(v1 + ((x >> 6) - v2)) + ((x >> 5) & 1u);
    const uint32_t h = (reinterpret_cast<const uint32_t *>(a.bsw)[w32] >> (x & 31u)) & 1u;
    if constexpr (W) return h ? (uint64_t)a.card[x] : 0ull;
    else return h;
  } else if constexpr (PASS == PASS_AND) {
    const uint32_t w = v0 + r;
    const uint32_t i2 = 2u * (w - S.aw0);
    uint64_t v = ((uint64_t)S.bm[i2] | ((uint64_t)S.bm[i2 + 1] << 32)) & a.bsw[v1 + (w - v2)];
    if constexpr (W) {
      uint64_t acc = 0;
      while (v) {
        acc += a.card[64u * w + (uint32_t)__builtin_ctzll(v)];
        v &= v - 1;
      }
      return acc;
    } else {
      return (uint32_t)__popcll(v);
    }
  } else if constexpr (PASS == PASS_SEARCH) {
    const uint32_t x = S.list[v0 + r];
    const uint32_t h = sorted_contains(a.unbr + 4ull * v1, v2, x);
    if constexpr (W) return h ? (uint64_t)a.card[x] : 0ull;
    else return h;
  } else if constexpr (PASS == PASS_BLOCK) {
    return (PassAcc<W>)block_and<W>(S, a, v0 + r);
  } else if constexpr (PASS == PASS_STREAM16) {
    return (PassAcc<W>)bitmap16_hits(S.bm, a.lo16[(uint64_t)v1 + r]);
  } else {
    const uint32_t *base = (PASS == PASS_NTSTREAM) ? a.ntu : a.unbr;
    return stream_hits<KIND, W, LOCKSTEP>(S, a, reinterpret_cast<const uint4 *>(base)[(uint64_t)v1 + r]);
  }
}

// Rank of the pair that owns slot t = 64 kb + lane of a flattened remainder: the pairs
// with remainder elements are numbered in lane order, `smask` holds per block of 64
// slots a 64-bit mask of the slots where one of them starts (two words), and `sb` is the
// lane's own pair's first slot (all ones if it has none).  The owner is the last pair
// starting at or before t: those starting before the block (a ballot, so every lane of
// the wave must be active) plus the block's starts at lanes 0..lane.  For slots past
// the last element it is the last pair.
__device__ __forceinline__ uint32_t slot_owner(const uint32_t *smask, uint32_t kb, uint32_t sb, uint32_t le_lo,
                                               uint32_t le_hi) {
  const uint32_t before = (uint32_t)__popcll(__ballot(sb < 64u * kb));
  return before + __popc(smask[2u * kb] & le_lo) + __popc(smask[2u * kb + 1u] & le_hi) - 1u;
}

// One pass over up to 64 pairs (one per lane, c elements each).  Whole rounds of 64
// elements run pair by pair with the pair's metadata wave-uniform (read from its
// lane into scalar registers); the remainders (< 64 per pair, so at most 63 blocks of
// 64 slots) are flattened: an exclusive scan gives every remainder element a slot t,
// lane t % 64 takes it and finds its pair with slot_owner.
template <int PASS, uint32_t KIND, bool W, int UNROLL>
__device__ __forceinline__ uint64_t flat_pass(const CountArgs &a, const AState &S, uint32_t c, uint32_t p0,
                                              uint32_t p1, uint32_t p2, uint32_t lane, uint32_t *meta) {
  if (__ballot(c != 0u) == 0ull) return 0;
  PassAcc<W> acc = 0;
  const uint32_t full = c & ~63u;
  uint64_t big = __ballot(full != 0u);
  while (big) {
    const uint32_t q = (uint32_t)__builtin_ctzll(big);
    big &= big - 1;
    const uint32_t fq = rdl(full, q), q0 = rdl(p0, q), q1 = rdl(p1, q), q2 = rdl(p2, q);
    if constexpr ((PASS == PASS_STREAM || PASS == PASS_STREAM16) && UNROLL > 1) {
      // fq is a multiple of 64: UNROLL vector loads per lane issued before any test
      const uint4 *vb = ((PASS == PASS_STREAM16) ? a.lo16 : reinterpret_cast<const uint4 *>(a.unbr)) + q1;
      for (uint32_t r0 = 0; r0 < fq; r0 += 64u * UNROLL) {
        // a block past fq is skipped (a wave-uniform branch)
        uint4 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++)
          if (r0 + 64u * u < fq) v[u] = vb[r0 + 64u * u + lane];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
          if (r0 + 64u * u < fq) {
            if constexpr (PASS == PASS_STREAM16) acc += (PassAcc<W>)bitmap16_hits(S.bm, v[u]);
            else acc += stream_hits<KIND, W, true>(S, a, v[u]);
          }
        }
      }
    } else {
      // fq is a multiple of 64: unrolled, each lane keeps several loads in flight
#pragma unroll UNROLL
      for (uint32_t r = lane; r < fq; r += 64) acc += pass_elem<PASS, KIND, W, (UNROLL > 1)>(a, S, q0, q1, q2, r);
    }
  }
  // remainders start after the whole rounds: PROBE/AND/BLOCK index by v0, the streams by v1
  c -= full;
  if constexpr (PASS == PASS_STREAM || PASS == PASS_NTSTREAM || PASS == PASS_STREAM16)
    p1 += full;
  else p0 += full;
  if (__ballot(c != 0u) == 0ull) return acc;
  uint32_t tot;
  const uint32_t base = wave_excl_scan(c, lane, &tot);
  tot = rfl(tot);
  uint32_t *mbase = meta, *m0 = meta + 64, *m1 = meta + 128, *m2 = meta + 192, *smask = meta + 256;
  const bool live = c != 0u;
  const uint32_t rank = mbcnt(__ballot(live));
  smask[lane] = 0u;
  smask[lane + 64u] = 0u;
  wave_sync();
  if (live) {
    mbase[rank] = base;
    m0[rank] = p0;
    m1[rank] = p1;
    m2[rank] = p2;
    atomicOr(&smask[base >> 5], 1u << (base & 31u));
  }
  wave_sync();
  const uint32_t sb = live ? base : 0xFFFFFFFFu;
  // slot-block lanes 0..lane (2u << 31 wraps to 0, so lane 31's low mask is all ones)
  const uint32_t le_lo = lane < 32u ? (2u << lane) - 1u : 0xFFFFFFFFu;
  const uint32_t le_hi = lane < 32u ? 0u : (2u << (lane - 32u)) - 1u;
  // the block loops are wave-uniform (slot_owner's ballot needs every lane)
  if constexpr ((PASS == PASS_STREAM || PASS == PASS_STREAM16) && UNROLL > 1) {
    // UNROLL blocks per step with their vector loads issued before any is tested; blocks
    // past the end are skipped (a wave-uniform branch), and in the last block a slot past
    // the end re-reads the last element (always a valid address) and counts zero
    const uint4 *vb = (PASS == PASS_STREAM16) ? a.lo16 : reinterpret_cast<const uint4 *>(a.unbr);
    for (uint32_t kb = 0; 64u * kb < tot; kb += UNROLL) {
      uint4 v[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; u++) {
        if (64u * (kb + u) < tot) {
          const uint32_t t = min(64u * (kb + u) + lane, tot - 1u);
          const uint32_t q = slot_owner(smask, kb + u, sb, le_lo, le_hi);
          v[u] = vb[(uint64_t)m1[q] + (t - mbase[q])];
        }
      }
      // keep the scheduler from moving the tests between the loads
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < UNROLL; u++) {
        if (64u * (kb + u) < tot) {
          const PassAcc<W> keep = (PassAcc<W>)0 - (PassAcc<W>)(64u * (kb + u) + lane < tot);
          PassAcc<W> h;
          if constexpr (PASS == PASS_STREAM16) h = (PassAcc<W>)bitmap16_hits(S.bm, v[u]);
          else h = stream_hits<KIND, W, true>(S, a, v[u]);
          acc += h & keep;
        }
      }
    }
  } else {
    for (uint32_t kb = 0; 64u * kb < tot; kb++) {
      const uint32_t t = 64u * kb + lane;
      const uint32_t q = slot_owner(smask, kb, sb, le_lo, le_hi);
      if (t < tot) acc += pass_elem<PASS, KIND, W>(a, S, m0[q], m1[q], m2[q], t - mbase[q]);
    }
  }
  wave_sync();
  return acc;
}

// sum over keys K[kb..ke) of |S & row(key)|
// UNROLL: depth of the whole-round loops (the triangle pass gains from 4; the
// 4-clique level sets, with more live state, run best at 1)
// BLK: the matrix has bitset_new / new_type rows (their block passes are compiled in)
template <uint32_t KIND, bool W, int UNROLL = 1, bool BLK = false>
__device__ __forceinline__ uint64_t sum_isect(const CountArgs &a, const AState &S, const uint32_t *K, uint32_t kb, uint32_t ke,
                                              uint32_t lane, uint32_t *meta) {
  uint64_t acc = 0;
  for (uint32_t kc = kb; kc < ke; kc += 64) {
    const uint32_t nk = min(64u, ke - kc);
    RowDesc d = {0u, 0u, 0u, 0u};
    uint32_t key = 0;
    if (lane < nk) {
      key = K[kc + lane];
      d = a.desc[key];
    }
    const uint32_t t = desc_type(d);
    // per pair, the cheapest of: probing A's elements in B's bitset words (PROBE),
    // AND-ing A's LDS bitmap with B's words (AND), streaming B's uint list against
    // A's LDS structure (STREAM).  Every row keeps its uint list, so pshort rows and
    // bitset rows whose probes would outnumber their elements are streamed.  A uint row
    // tagged for search (UG) is never streamed: A's elements between B's first and last
    // id are each searched in B's list (SEARCH), the per-pair algorithm for long B rows
    // that few of A's elements fall in.
    uint32_t pc = 0, ps = 0, ac = 0, as = 0, sc = 0, soff = d.off;
    uint32_t uc = 0, us = 0;                    // SEARCH items
    uint32_t bc = 0, bs = 0, nc = 0, noff = 0;  // BLOCK and NTSTREAM items
    if (d.card) {
      if (BLK && KIND == KIND_BITMAP && is_block_type(t)) {
        // blocks AND-ed with the bitmap (+ new_type's uint part) unless the canonical
        // list is shorter: a block and a 4-element vector cost about the same
        const uint32_t nb = desc_aux(d);
        uint32_t nu = 0, uoff = 0;
        if (t == EH_NT) { nu = a.bid[d.w0]; uoff = a.bid[d.w0 + 1]; }
        sc = (d.card + 3u) >> 2;
        if (nb + ((nu + 3u) >> 2) < sc) {
          sc = 0;
          bc = nb;
          bs = d.w0 + (t == EH_NT ? EH_NT_HEADER : 0u);
          nc = (nu + 3u) >> 2;
          noff = uoff;
        }
      } else if (t == EH_BS) {
        const uint32_t bw0 = d.w0, bnw = desc_aux(d);
        const uint64_t xlo = (uint64_t)bw0 * 64ull, xhi = (uint64_t)(bw0 + bnw) * 64ull;
        const uint32_t s = (xlo > S.lo + S.span) ? S.n : lds_lower_bound(S.list, S.n, (uint32_t)xlo);
        const uint32_t e = (xhi > 0xFFFFFFFFull) ? S.n : lds_lower_bound(S.list, S.n, (uint32_t)xhi);
        pc = e - s;
        ps = s;
        uint64_t best = 16ull * pc;  // costs in units of 1/16 probe
        if (a.stream_ratio && 16ull * d.card < (uint64_t)a.stream_ratio * pc) {
          best = (16ull * d.card) / a.stream_ratio;
          sc = (d.card + 3u) >> 2;
          soff = a.aoff[key];
          pc = 0;
        }
        if constexpr (KIND == KIND_BITMAP) {
          const uint32_t os = max(S.aw0, bw0), oe = min(S.aw0 + S.anw, bw0 + bnw);
          const uint32_t ov = oe > os ? oe - os : 0u;
          if (16ull * ov < best) { ac = ov; as = os; pc = 0; sc = 0; }
        }
      } else if (t == EH_UG) {
        const uint32_t *Bl = a.unbr + (uint64_t)d.off * 4;
        const uint32_t blo = Bl[0], bhi = Bl[d.card - 1];
        const uint32_t s = lds_lower_bound(S.list, S.n, blo);
        const uint32_t e = (bhi == 0xFFFFFFFFu) ? S.n : lds_lower_bound(S.list, S.n, bhi + 1u);
        uc = e > s ? e - s : 0u;
        us = s;
      } else {
        sc = (d.card + 3u) >> 2;
        if (!is_uint_type(t)) soff = a.aoff[key];
      }
    }
    acc += flat_pass<PASS_SEARCH, KIND, W, UNROLL>(a, S, uc, us, d.off, d.card, lane, meta);
    acc += flat_pass<PASS_PROBE, KIND, W, UNROLL>(a, S, pc, ps, d.off, d.w0, lane, meta);
    if constexpr (KIND == KIND_BITMAP)
      acc += flat_pass<PASS_AND, KIND, W, UNROLL>(a, S, ac, as, d.off, d.w0, lane, meta);
    acc += flat_pass<PASS_STREAM, KIND, W, UNROLL>(a, S, sc, 0u, soff, 0u, lane, meta);
    if constexpr (BLK && KIND == KIND_BITMAP) {
      acc += flat_pass<PASS_BLOCK, KIND, W, UNROLL>(a, S, bc, bs, 0u, 0u, lane, meta);
      acc += flat_pass<PASS_NTSTREAM, KIND, W, UNROLL>(a, S, nc, 0u, noff, 0u, lane, meta);
    }
  }
  return acc;
}

__device__ __forceinline__ uint32_t hash_lg(uint32_t n) {
  uint32_t lg = 4;
  while ((1u << lg) < 2u * n) lg++;
  return lg;
}

// LDS hash table over a sorted list (linear probing, multiplicative hash)
// (at most half the 2^lg slots are used; an insert that finds no free slot in a full walk
// sets `bad` instead of spinning)
__device__ __forceinline__ void build_hash(const uint32_t *list, uint32_t n, uint32_t *ht, uint32_t lg,
                                           uint32_t lane, uint32_t &bad) {
  const uint32_t cap = 1u << lg, mask = cap - 1u;
  for (uint32_t w = lane; w < cap; w += 64) ht[w] = HASH_EMPTY;
  wave_sync();
  for (uint32_t k = lane; k < n; k += 64) {
    const uint32_t x = list[k];
    uint32_t s = ((x * 2654435761u) >> (32u - lg)) & mask;
    uint32_t it = 0;
    while (it <= mask && atomicCAS(&ht[s], HASH_EMPTY, x) != HASH_EMPTY) {
      s = (s + 1u) & mask;
      it++;
    }
    bad |= (uint32_t)(it > mask);
  }
  // wave-uniform: a table with a failed insert is never probed
  bad = __ballot(bad != 0u) ? 1u : 0u;
  wave_sync();
}

// stage a tuned source row: list + (bitmap | hash | nothing)
__device__ __forceinline__ uint32_t stage_tuned(const CountArgs &a, uint32_t kind, uint32_t i, uint32_t lane,
                                                uint32_t *L, AState &A) {
  if (kind != KIND_HASH) return stage_a(a, kind, i, lane, L, A);
  const uint32_t used = stage_a(a, KIND_LIST, i, lane, L, A);
  const uint32_t lg = hash_lg(A.n);
  A.ht = L + used;
  A.hlg = lg;
  A.bad = 0u;
  build_hash(A.list, A.n, A.ht, lg, lane, A.bad);
  return used + (1u << lg);
}

template <uint32_t KIND, bool BLK>
__device__ uint64_t task_cliques4_tuned(const CountArgs &a, const AState &A, uint32_t jb, uint32_t je,
                                        uint32_t lane, uint32_t *free_lds, uint32_t *meta) {
  uint64_t acc = 0;
  uint32_t *clist = free_lds;
  uint32_t *cmemb = free_lds + A.n;
  for (uint32_t jc = jb; jc < je; jc += 64) {
    const uint32_t nj = min(64u, je - jc);
    RowDesc dl = {0u, 0u, 0u, 0u};
    if (lane < nj) dl = a.desc[A.list[jc + lane]];
    for (uint32_t t = 0; t < nj; t++) {
      RowDesc dj;
      dj.card = rdl(dl.card, t);
      dj.tinfo = rdl(dl.tinfo, t);
      dj.w0 = rdl(dl.w0, t);
      dj.off = rdl(dl.off, t);
      const uint32_t nc = rfl(isect_write<KIND>(A, a, dj, lane, clist));
      if (nc == 0) continue;
      AState C;
      if constexpr (KIND == KIND_BITMAP) {
        make_cstate<KIND_BITMAP>(A, clist, nc, cmemb, lane, C);
        acc += sum_isect<KIND_BITMAP, false, 1, BLK>(a, C, clist, 0, nc, lane, meta);
      } else if constexpr (KIND == KIND_HASH) {
        make_cstate<KIND_LIST>(A, clist, nc, cmemb, lane, C);
        C.hlg = hash_lg(nc);
        C.ht = cmemb;
        C.bad = 0u;
        build_hash(clist, nc, cmemb, C.hlg, lane, C.bad);
        if (C.bad) A.bad = 1u;  // wave-uniform: C's table is not probed
        else acc += sum_isect<KIND_HASH, false>(a, C, clist, 0, nc, lane, meta);
      } else {
        make_cstate<KIND_LIST>(A, clist, nc, cmemb, lane, C);
        acc += sum_isect<KIND_LIST, false>(a, C, clist, 0, nc, lane, meta);
      }
      wave_sync();
    }
  }
  return acc;
}

template <int K, bool W, bool BLK>
__device__ __forceinline__ uint64_t run_task_tuned(const CountArgs &a, uint32_t kind, uint32_t i, uint32_t jb,
                                                   uint32_t je, uint32_t lane, uint32_t *L) {
  uint32_t *meta = L;
  uint32_t *Ls = L + META_WORDS;
  AState A;
  const uint32_t used = stage_tuned(a, kind, i, lane, Ls, A);
  if (A.bad) {  // row(i)'s hash table (wave-uniform flag): no free slot, nothing is probed
    if (lane == 0) atomicOr(a.err, ERR_HASH_FULL);
    return 0;
  }
  uint64_t r;
  if constexpr (K == 3) {
    switch (kind) {
      case KIND_BITMAP: return sum_isect<KIND_BITMAP, W, 4, BLK>(a, A, A.list, jb, je, lane, meta);
      case KIND_HASH: r = sum_isect<KIND_HASH, W, 4>(a, A, A.list, jb, je, lane, meta); break;
      case KIND_LIST: return sum_isect<KIND_LIST, W, 4>(a, A, A.list, jb, je, lane, meta);
      // nothing staged: B's elements are searched in row(i) in HBM, the pairs flattened
      // like the staged kinds (the wave's slice holds the pass metadata only)
      default: return sum_isect<KIND_GLOBAL, W, 4>(a, A, A.list, jb, je, lane, meta);
    }
  } else {
    switch (kind) {
      case KIND_BITMAP: r = task_cliques4_tuned<KIND_BITMAP, BLK>(a, A, jb, je, lane, Ls + used, meta); break;
      case KIND_HASH: r = task_cliques4_tuned<KIND_HASH, BLK>(a, A, jb, je, lane, Ls + used, meta); break;
      default: r = task_cliques4_tuned<KIND_LIST, BLK>(a, A, jb, je, lane, Ls + used, meta); break;
    }
  }
  // a 4-clique task's C table that found no free slot
  if (A.bad && lane == 0) atomicOr(a.err, ERR_HASH_FULL);
  return r;
}

template <int K, bool W, bool TUNED, bool BLK>
__device__ __forceinline__ uint64_t run_task(const CountArgs &a, uint32_t kind, uint32_t i, uint32_t jb,
                                             uint32_t je, uint32_t lane, uint32_t *L) {
  if constexpr (TUNED) return run_task_tuned<K, W, BLK>(a, kind, i, jb, je, lane, L);
  AState A;
  const uint32_t used = stage_a(a, kind, i, lane, L, A);
  if constexpr (K == 3) {
    switch (kind) {
      case KIND_BITMAP: return task_triangles<KIND_BITMAP, W>(a, A, jb, je, lane);
      case KIND_LIST: return task_triangles<KIND_LIST, W>(a, A, jb, je, lane);
      case KIND_PSDIR: return task_triangles<KIND_PSDIR, W>(a, A, jb, je, lane);
      default: return task_triangles<KIND_GLOBAL, W>(a, A, jb, je, lane);
    }
  } else {
    switch (kind) {
      case KIND_BITMAP: return task_cliques4<KIND_BITMAP>(a, A, jb, je, lane, L + used);
      case KIND_LIST: return task_cliques4<KIND_LIST>(a, A, jb, je, lane, L + used);
      case KIND_PSDIR: return task_cliques4<KIND_PSDIR>(a, A, jb, je, lane, L + used);
      default: return task_cliques4<KIND_GLOBAL>(a, A, jb, je, lane, L + used);
    }
  }
}

// ---------------------------------------------------------------- small source rows
//
// Rows with few kept entries are counted edge-centrically: for an oriented
// matrix (every row j only holds ids below j) the triangles closed at row i are
// the pairs a < b of row(i) with row(i)[a] in row(row(i)[b]), and the
// 4-cliques are those pairs extended by every c > b whose row holds both.  The
// pairs of 64 rows are flattened over the lanes; each is a handful of
// membership tests against rows in HBM, with no LDS staging.

// galloping membership: probe positions 1, 2, 4, ... until one holds a value >= x,
// then binary-search the last doubling interval (the exponential search the
// reference's set_intersect_galloping runs over 128-element blocks)
__device__ __forceinline__ uint32_t gallop_contains(const uint32_t *R, uint32_t n, uint32_t x) {
  if (n == 0) return 0;
  if (R[0] >= x) return R[0] == x;
  uint32_t b = 1;
  while (b < n && R[b] < x) b <<= 1;
  uint32_t lo = (b >> 1) + 1, hi = min(b, n - 1) + 1;  // answer in [lo, hi)
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (R[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < n && R[lo] == x;
}

// membership of x in row(j) as stored in HBM
__device__ __forceinline__ uint32_t row_contains(const CountArgs &a, const RowDesc &d, uint32_t x) {
  if (d.card == 0) return 0;
  const uint32_t t = desc_type(d);
  if (t == EH_BS) {
    const uint32_t w = (x >> 6) - d.w0;
    return (w < desc_aux(d)) ? (uint32_t)((a.bsw[d.off + w] >> (x & 63u)) & 1ull) : 0u;
  }
  if (t == EH_U || is_block_type(t)) return sorted_contains(a.unbr + (uint64_t)d.off * 4, d.card, x);
  if (t == EH_UG) return gallop_contains(a.unbr + (uint64_t)d.off * 4, d.card, x);
  const uint16_t *S = a.psd + (uint64_t)d.off * 4;
  const uint32_t len = desc_aux(d), hx = x >> 16;
  uint32_t pos = 0;
  while (pos < len) {
    const uint32_t pre = S[pos], cnt = (uint32_t)S[pos + 1] + 1u;
    if (pre == hx) return sorted_contains(S + pos + 2, cnt, x & 0xFFFFu);
    if (pre > hx) return 0;
    pos += cnt + 2u;
  }
  return 0;
}

// r -> (p, q) with 0 <= p < q, r = q(q-1)/2 + p
__device__ __forceinline__ void tri_decode(uint32_t r, uint32_t *p, uint32_t *q) {
  uint32_t b = (uint32_t)sqrtf(2.0f * (float)r);
  while ((b * (b + 1u)) / 2u <= r) b++;
  while (b > 1u && (b * (b - 1u)) / 2u > r) b--;
  if (b < 1u) b = 1u;
  *q = b;
  *p = r - (b * (b - 1u)) / 2u;
}

// One membership query x in row(j) of the two-query interleave below: a bitset row or an
// empty row answers at once, a pshort row walks its partitions, every other row (uint,
// and the canonical list of a block row) is searched in the sorted list.
struct MemberQuery {
  const uint32_t *L;
  uint32_t base, len, x, hit;
};

__device__ __forceinline__ void mq_start(const CountArgs &a, const RowDesc &d, uint32_t x, MemberQuery &q) {
  q.x = x;
  q.base = 0;
  q.len = 0;
  q.hit = 0;
  q.L = a.unbr;
  if (d.card == 0) return;
  const uint32_t t = desc_type(d);
  if (t == EH_BS || t == EH_PS) {
    q.hit = row_contains(a, d, x);
    return;
  }
  q.L = a.unbr + (uint64_t)d.off * 4;
  q.len = d.card;
}

// advance NQ searches together: every probe is issued before any is used, so a lane keeps
// NQ dependent load chains in flight; lanes whose search is over probe their base
template <int NQ>
__device__ __forceinline__ void mq_finish(MemberQuery (&q)[NQ]) {
  for (;;) {
    bool more = false;
#pragma unroll
    for (int k = 0; k < NQ; k++) more |= q[k].len > 1;
    if (!more) break;
    uint32_t v[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) v[k] = q[k].L[q[k].base + (q[k].len >> 1)];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
      const uint32_t h = q[k].len >> 1;
      if (q[k].len > 1) {
        q[k].base = (v[k] <= q[k].x) ? q[k].base + h : q[k].base;
        q[k].len -= h;
      }
    }
  }
  uint32_t v[NQ];
#pragma unroll
  for (int k = 0; k < NQ; k++) v[k] = q[k].L[q[k].base];
#pragma unroll
  for (int k = 0; k < NQ; k++)
    if (q[k].len == 1) q[k].hit = v[k] == q[k].x;
}

constexpr int SMALL_NQ = 4;  // pairs per lane whose searches run interleaved
// the XCD queues of a short-row list: 9 bounds, then 8 heads on lines of their own
constexpr uint32_t XQ_HEADS = 64, XQ_STRIDE = 64, XQ_WORDS = XQ_HEADS + 8 * XQ_STRIDE;

// Add a wave's per-row counts c (rows of one chunk, valid lanes only) into the per-graph
// counters.  The chunk's rows run in row order, so their graph ids form a few runs: one
// atomic per distinct graph id instead of one per row (these atomics are performed at the
// memory side and serialise per line).
__device__ __forceinline__ void add_graph_counts(unsigned long long *counts, uint32_t c, uint32_t gid, uint32_t lane) {
  uint64_t pend = __ballot(c != 0u);
  while (pend) {
    const uint32_t g = rdl(gid, (uint32_t)__builtin_ctzll(pend));
    const bool mine = c != 0u && gid == g;
    const uint64_t sum = wave_sum_u64(mine ? (uint64_t)c : 0ull);
    if (lane == 0) atomicAdd(&counts[g], (unsigned long long)sum);
    pend &= ~__ballot(mine);
  }
}

// The next (at most) 64 rows of a short-row list for a wave: [c0, c0 + nr); false once
// every queue is drained.  Without XCD queues the list is one queue whose head is a.next.
// With them it is 8 ranges of equal work, a.xq[q] .. a.xq[q + 1], with heads
// a.xq[XQ_HEADS + XQ_STRIDE * q]:
// workgroups are dispatched to the 8 XCDs in turn, so blockIdx.x % 8 names the block's
// XCD, and a wave takes from that XCD's range first, then from the next ranges in turn
// (qs, wave-uniform, counts the ranges it has found drained).  Which range a block takes
// changes only what its L2 holds, never which rows are counted.
__device__ __forceinline__ bool take_rows(const CountArgs &a, uint32_t nrows, uint32_t lane, uint32_t &qs,
                                          uint32_t &c0, uint32_t &nr) {
  if (!a.xq) {
    uint32_t c = 0;
    if (lane == 0) c = atomicAdd(a.next, 64u);
    c = rfl(c);
    if (c >= nrows) return false;
    c0 = c;
    nr = min(64u, nrows - c);
    return true;
  }
  for (; qs < 8u; qs++) {
    const uint32_t q = (blockIdx.x + qs) & 7u;
    const uint32_t b = a.xq[q], e = a.xq[q + 1u];
    if (b >= e) continue;
    uint32_t c = 0;
    if (lane == 0) c = atomicAdd(a.xq + XQ_HEADS + XQ_STRIDE * q, 64u);
    c = rfl(c);
    if (c < e - b) {
      c0 = b + c;
      nr = min(64u, e - b - c);
      return true;
    }
  }
  return false;
}

template <int K, bool W>
__global__ void __launch_bounds__(256) k_small_rows(CountArgs a, const uint32_t *__restrict__ rows, uint32_t nrows) {
  static_assert(K == 3, "4-cliques of short rows run in k_small_cliques4");
  extern __shared__ uint32_t lds_all[];
  const uint32_t lane = lane_id();
  uint32_t *meta = lds_all + (threadIdx.x >> 6) * SMALL_META_WORDS;
  uint32_t *mbase = meta, *moff = meta + 65, *rcnt = meta + 130, *mdeg = meta + 194;
  const bool per_graph = a.graph_of != nullptr;
  uint64_t acc = 0;
  uint32_t qs = 0, c0 = 0, nr = 0;
  while (take_rows(a, nrows, lane, qs, c0, nr)) {
    uint32_t items = 0, off = 0, gid = 0, d = 0, skip = 0;
    if (lane < nr) {
      const uint32_t i = rows[c0 + lane];
      d = a.card[i];
      off = a.aoff[i];
      if (per_graph) gid = a.graph_of[i];
      // wedges (p, q), p < q, of the row; those with q below jstart belong to k_bstat
      const uint32_t js = a.jstart ? a.jstart[i] : 0u;
      skip = (js * (js - (js ? 1u : 0u))) / 2u;
      items = (d * (d - 1u)) / 2u - skip;
    }
    uint32_t tot;
    const uint32_t base = wave_excl_scan(items, lane, &tot);
    tot = rfl(tot);
    mbase[lane] = base;
    moff[lane] = off;
    mdeg[lane] = skip;
    rcnt[lane] = 0;
    if (lane == 0) mbase[64] = tot;
    wave_sync();
    {
      // SMALL_NQ pairs per lane, t + 64 k, their searches interleaved.  A lane's slots only
      // grow, so its row index moves forward from the previous slot's (a row holds tens to
      // thousands of wedges: usually no step, where a search over the 64 bases took six
      // dependent LDS reads); mbase[64] = tot stops it
      uint32_t qc = 0;
      for (uint32_t t = lane; t < tot; t += 64 * SMALL_NQ) {
        uint32_t q[SMALL_NQ], x[SMALL_NQ], j[SMALL_NQ];
#pragma unroll
        for (int k = 0; k < SMALL_NQ; k++) {
          const uint32_t tk = t + 64u * k;
          while (qc < 63u && mbase[qc + 1u] <= tk) qc++;
          q[k] = qc;
          x[k] = 0;
          j[k] = 0;
          if (tk < tot) {
            uint32_t p, sq;
            tri_decode(tk - mbase[q[k]] + mdeg[q[k]], &p, &sq);
            const uint32_t *R = a.unbr + (uint64_t)moff[q[k]] * 4;
            x[k] = R[p];
            j[k] = R[sq];
          }
        }
        MemberQuery m[SMALL_NQ];
#pragma unroll
        for (int k = 0; k < SMALL_NQ; k++) {
          const RowDesc d = (t + 64u * k < tot) ? a.desc[j[k]] : RowDesc{0u, 0u, 0u, 0u};
          mq_start(a, d, x[k], m[k]);
        }
        mq_finish<SMALL_NQ>(m);
#pragma unroll
        for (int k = 0; k < SMALL_NQ; k++) {
          uint64_t h;
          if constexpr (W) h = m[k].hit ? (uint64_t)a.card[x[k]] : 0ull;
          else h = m[k].hit;
          if (per_graph) {
            if (h) atomicAdd(&rcnt[q[k]], (uint32_t)h);
          } else {
            acc += h;
          }
        }
      }
    }
    wave_sync();
    if (per_graph) add_graph_counts(a.counts, lane < nr ? rcnt[lane] : 0u, gid, lane);
    wave_sync();
  }
  const uint64_t s = wave_sum_u64(acc);
  if (lane == 0 && s) atomicAdd(&a.counts[0], (unsigned long long)s);
}

// 4-cliques of short rows through each row's local graph.  For a source row i with
// sorted entries R[0..d), the pairs p < c with R[p] in row(R[c]) are the edges of
// the graph induced on row(i); a 4-clique {i, R[a], R[b], R[c]} (a < b < c) is a
// triangle of that local graph.  Pass 1 runs the triangle pass's membership tests
// (four interleaved per lane) and records each hit as bit p of the LDS bit-row
// adj(c) (ceil(d/32) words per local vertex); pass 2 visits the pairs again and, for
// every local edge (b, c), adds popcount(adj(b) & adj(c)) -- the a < b adjacent to
// both.  The extension over later entries is then LDS work, not more HBM searches.
// A wave takes 64 rows at a time and runs them in sub-batches whose bit-rows fit
// its SMALL_ADJ_WORDS (a row of up to 64 entries needs at most 128 words).
// With a.counts3 set the kernel is the triangle pass of the same rows too: every hit of
// pass 1 is a triangle {i, R[p], R[c]}, so each row's hits go to its graph's triangle
// counter, and rows of two entries (one wedge, no 4-clique) run pass 1 as well.
constexpr uint32_t SMALL4_META_WORDS = 456;

// Pass 1 of k_small_cliques4 in chunks: the pairs (p, c), p < c, of a row of d entries as
// chunks of up to eight p's of one column c (p0 = 0, 8, ...): column c has ceil(c / 8) of
// them, the columns 8(b-1)+1 .. 8b have b each, so with D = d - 1 = 8F + R the row holds
// 4F(F+1) + R(F+1).  A chunk reads its B row's descriptor once and tests its eight ids
// together (one aligned 32-byte read of the row), where one query per pair re-read both.
__device__ __forceinline__ uint32_t chunk_count(uint32_t d) {
  if (d < 2u) return 0u;
  const uint32_t D = d - 1u, F = D >> 3, R = D & 7u;
  return 4u * F * (F + 1u) + R * (F + 1u);
}

// chunk r of a row -> (column c, first p): block b holds the chunks 4b(b-1) .. 4b(b+1) - 1
__device__ __forceinline__ void chunk_decode(uint32_t r, uint32_t *c, uint32_t *p0) {
  uint32_t b = (uint32_t)((1.0f + sqrtf(1.0f + (float)r)) * 0.5f);
  if (b < 1u) b = 1u;
  while (4u * b * (b + 1u) <= r) b++;
  while (b > 1u && 4u * b * (b - 1u) > r) b--;
  const uint32_t o = r - 4u * b * (b - 1u);
  *c = 8u * (b - 1u) + 1u + o / b;
  *p0 = (o % b) * 8u;
}

// bit u set when x[u] (u < np) is in row(j) (descriptor d); the ids are tested together
__device__ __forceinline__ uint32_t chunk_hits(const CountArgs &a, const RowDesc &d, const uint32_t (&x)[8],
                                               uint32_t np) {
  if (d.card == 0) return 0u;
  const uint32_t t = desc_type(d);
  uint32_t h = 0;
  if (t == EH_BS) {
    uint32_t v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t w = (x[u] >> 6) - d.w0;
      v[u] = ((uint32_t)u < np && w < desc_aux(d))
                 ? reinterpret_cast<const uint32_t *>(a.bsw)[2ull * (d.off + w) + ((x[u] >> 5) & 1u)]
                 : 0u;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) h |= ((v[u] >> (x[u] & 31u)) & 1u) << u;
    return h;
  }
  if (t == EH_PS) {
    for (uint32_t u = 0; u < np; u++) h |= row_contains(a, d, x[u]) << u;
    return h;
  }
  // uint, uint searched and the block rows' canonical lists: up to eight searches of one sorted
  // list in lockstep (one length for all); the slots past np issue no loads
  const uint32_t *L = a.unbr + (uint64_t)d.off * 4;
  uint32_t b[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  uint32_t len = d.card;
  while (len > 1u) {
    const uint32_t half = len >> 1;
    uint32_t v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = ((uint32_t)u < np) ? L[b[u] + half] : 0xFFFFFFFFu;
#pragma unroll
    for (int u = 0; u < 8; u++) b[u] = (v[u] <= x[u]) ? b[u] + half : b[u];
    len -= half;
  }
#pragma unroll
  for (int u = 0; u < 8; u++) h |= ((uint32_t)u < np && L[b[u]] == x[u]) ? (1u << u) : 0u;
  return h;
}
constexpr uint32_t LOCAL4_META = 196;  // k_local4: pair scan bases (65), offsets (64), columns (64)
constexpr uint32_t LOCAL4_MID = 128, LOCAL4_MAX = 512;  // the two k_local4 row classes
constexpr uint32_t SMALL_ADJ_WORDS = 1024;

__global__ void __launch_bounds__(256) k_small_cliques4(CountArgs a, const uint32_t *__restrict__ rows,
                                                        uint32_t nrows) {
  extern __shared__ uint32_t lds_all[];
  const uint32_t lane = lane_id();
  uint32_t *meta = lds_all + (threadIdx.x >> 6) * (SMALL4_META_WORDS + SMALL_ADJ_WORDS);
  uint32_t *mbase = meta, *moff = meta + 65, *rcnt = meta + 130, *mdeg = meta + 194, *mwb = meta + 258;
  uint32_t *rcnt3 = meta + 322, *cbase = meta + 386;
  uint32_t *adj = meta + SMALL4_META_WORDS;
  const bool per_graph = a.graph_of != nullptr;
  const bool tri = a.counts3 != nullptr;
  uint64_t acc = 0, acc3 = 0;
  uint32_t qs = 0, c0 = 0, nr = 0;
  while (take_rows(a, nrows, lane, qs, c0, nr)) {
    uint32_t off = 0, gid = 0, d = 0;
    if (lane < nr) {
      const uint32_t i = rows[c0 + lane];
      d = a.card[i];
      off = a.aoff[i];
      if (per_graph) gid = a.graph_of[i];
    }
    rcnt[lane] = 0;
    rcnt3[lane] = 0;
    // fewer than 3 entries close no 4-clique; 2 close a triangle
    bool pend = lane < nr && d >= (tri ? 2u : 3u);
    while (__ballot(pend)) {
      // the pending rows whose bit-rows fit form a prefix of the lanes (at least one row)
      const uint32_t wpr = (d + 31u) >> 5;
      const uint32_t words = pend ? d * wpr : 0u;
      uint32_t wtot;
      const uint32_t wb = wave_excl_scan(words, lane, &wtot);
      const bool sel = pend && wb + words <= SMALL_ADJ_WORDS;
      // pass 2 visits every column c >= 1 of a row: d - 1 of them
      const uint32_t items = sel ? d - 1u : 0u;
      uint32_t tot, ctot;
      const uint32_t base = wave_excl_scan(items, lane, &tot);
      tot = rfl(tot);
      const uint32_t cb = wave_excl_scan(sel ? chunk_count(d) : 0u, lane, &ctot);
      ctot = rfl(ctot);
      mbase[lane] = base;
      cbase[lane] = cb;
      moff[lane] = off;
      mdeg[lane] = sel ? d : 0u;
      mwb[lane] = wb;
      if (lane == 0) {
        mbase[64] = tot;
        cbase[64] = ctot;
      }
      for (uint32_t w = lane; w < SMALL_ADJ_WORDS; w += 64) adj[w] = 0u;
      wave_sync();
      // pass 1: local edges, a chunk of up to eight pairs of one column per lane (a lane's
      // row index moves forward with its chunks, as in k_small_rows); a chunk's hits are one
      // byte of adj(c): p0 is a multiple of 8, so its bits never cross a word
      uint32_t qc = 0;
      for (uint32_t t = lane; t < ctot; t += 64) {
        while (qc < 63u && cbase[qc + 1u] <= t) qc++;
        const uint32_t q = qc;
        uint32_t c, p0;
        chunk_decode(t - cbase[q], &c, &p0);
        // the row's entries p0 .. p0 + 7 (aligned: rows start on 16 bytes, p0 is a multiple of
        // 8); those past c are not tested, and a read past the row's end stays in the arena
        const uint32_t *R = a.unbr + (uint64_t)moff[q] * 4;
        const uint4 xa = *reinterpret_cast<const uint4 *>(R + p0), xb = *reinterpret_cast<const uint4 *>(R + p0 + 4);
        const uint32_t x[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
        const RowDesc dd = a.desc[R[c]];
        const uint32_t h = chunk_hits(a, dd, x, min(8u, c - p0));
        if (h) {
          const uint32_t wq = (mdeg[q] + 31u) >> 5;
          atomicOr(&adj[mwb[q] + c * wq + (p0 >> 5)], h << (p0 & 31u));
          if (tri) {
            if (per_graph) atomicAdd(&rcnt3[q], (uint32_t)__popc(h));
            else acc3 += (uint32_t)__popc(h);
          }
        }
      }
      wave_sync();
      // pass 2: every local edge (b, c) adds |adj(b) & adj(c)|; a lane takes a column c and
      // walks the set bits b of adj(c), its local edges, instead of every pair (b, c)
      qc = 0;
      for (uint32_t t = lane; t < tot; t += 64) {
        while (qc < 63u && mbase[qc + 1u] <= t) qc++;
        const uint32_t q = qc;
        const uint32_t c = t - mbase[q] + 1u;
        const uint32_t wq = (mdeg[q] + 31u) >> 5;
        const uint32_t *ac = adj + mwb[q] + c * wq;
        uint32_t h = 0;
        // adj(c) only holds bits below c (c < 64: at most two words)
        for (uint32_t wc = 0; wc <= (c >> 5); wc++) {
          uint32_t bits = ac[wc];
          while (bits) {
            const uint32_t b = 32u * wc + (uint32_t)__builtin_ctz(bits);
            bits &= bits - 1u;
            const uint32_t *ab = adj + mwb[q] + b * wq;
            // adj(b) only holds bits below b: words past b's word are zero
            for (uint32_t w = 0; w <= (b >> 5); w++) h += __popc(ab[w] & ac[w]);
          }
        }
        if (per_graph) {
          if (h) atomicAdd(&rcnt[q], h);
        } else {
          acc += h;
        }
      }
      wave_sync();
      pend = pend && !sel;
    }
    wave_sync();
    if (per_graph) add_graph_counts(a.counts, lane < nr ? rcnt[lane] : 0u, gid, lane);
    if (per_graph && tri) add_graph_counts(a.counts3, lane < nr ? rcnt3[lane] : 0u, gid, lane);
    wave_sync();
  }
  const uint64_t s = wave_sum_u64(acc);
  if (lane == 0 && s) atomicAdd(&a.counts[0], (unsigned long long)s);
  if (tri) {
    const uint64_t s3 = wave_sum_u64(acc3);
    if (lane == 0 && s3) atomicAdd(&a.counts3[0], (unsigned long long)s3);
  }
}

// B-stationary triangle pairs.  On an oriented matrix row(j) only holds ids below j,
// so for a pair (i, j) with j at position p of row(i), |row(i) & row(j)| =
// |row(i)[0, p) & row(j)|.  A task {j, b, e} stages row(j) once as an LDS bitmap over
// the ids [0, split) and streams the prefixes row(i)[0, p) of its in-pairs [b, e)
// through it, 16-byte vectors of eight 16-bit ids flattened over the lanes (every id of
// such a prefix is below j < split <= 65536; the vector's tail holds ids >= j or the
// padding split - 1, whose bits are zero in any row(j) with j < split).  Each prefix is
// read contiguously at two bytes an id, where probing every element of it into row(j)'s
// bitset in HBM touches about one cache line per element: the hub rows with low ids are
// probed by most rows.
__global__ void __launch_bounds__(256) k_bstat(CountArgs a, const uint4 *__restrict__ btasks, uint64_t nbt,
                                               const uint2 *__restrict__ in_pairs, uint32_t bm_words) {
  extern __shared__ uint32_t lds_all[];
  const uint32_t lane = lane_id();
  uint32_t *meta = lds_all + (threadIdx.x >> 6) * a.lds_words;
  uint32_t *bm = meta + META_WORDS;
  uint64_t acc = 0;
  uint64_t t = 0, tend = 0;
  AState S;
  S.bm = bm;
  S.base32 = S.aw0 = S.lo = 0;
  S.anw = bm_words >> 1;
  S.span = 0xFFFFFFFFu;
  S.list = nullptr;
  S.n = 0;
  S.P = 0;
  S.dpre = S.dstart = nullptr;
  S.ht = nullptr;
  S.bad = 0u;
  S.hlg = 0;
  for (;;) {
    if (t >= tend) {
      uint32_t t0 = 0;
      if (lane == 0) t0 = atomicAdd(a.next, a.grab);
      t = rfl(t0);
      if (t >= nbt) break;
      tend = min(t + a.grab, nbt);
    }
    // {j | |row(j)| << 16, first in-pair, end, row(j)'s 16-byte offset}: the row is found
    // without another dependent load, and the first 64 in-pairs load while the bitmap builds
    const uint4 task = btasks[t++];
    const uint32_t dj = rfl(task.x) >> 16, b = rfl(task.y), e = rfl(task.z);
    const uint32_t *G = a.unbr + (uint64_t)rfl(task.w) * 4;
    uint2 pr = (b + lane < e) ? in_pairs[b + lane] : make_uint2(0u, 0u);
    // the whole bitmap is cleared: a prefix vector's tail may test ids up to split - 1
    for (uint32_t w = lane; w < bm_words; w += 64) bm[w] = 0u;
    wave_sync();
    for (uint32_t k = lane; k < dj; k += 64) {
      const uint32_t x = G[k];
      atomicOr(&bm[x >> 5], 1u << (x & 31u));
    }
    wave_sync();
    for (uint32_t c0 = b; c0 < e; c0 += 64) {
      const uint32_t nv = (c0 + lane < e) ? (pr.y + 7u) >> 3 : 0u, off = pr.x;
      // the next 64 in-pairs load during this pass
      const uint2 prn = (c0 + 64u + lane < e) ? in_pairs[c0 + 64u + lane] : make_uint2(0u, 0u);
      acc += flat_pass<PASS_STREAM16, KIND_BITMAP, false, 4>(a, S, nv, 0u, off, 0u, lane, meta);
      pr = prn;
    }
    wave_sync();
  }
  const uint64_t sum = wave_sum_u64(acc);
  if (lane == 0 && sum) atomicAdd(&a.counts[0], (unsigned long long)sum);
}

// 4-cliques of mid-size rows through each row's local graph (the short-row form of
// k_small_cliques4 for rows of up to LMAX entries, one row per block).  Row i's list R is
// staged in LDS with a hash table from id to position; for every c the B row row(R[c]) is
// streamed (16-byte vectors; each wave flattens the items of its own 64 c's over its lanes,
// four loads per lane in flight) and each of its ids found in R at position p (always p < c:
// row(R[c]) lies below R[c]) sets bit p of the bit-row adj(c).  Then every local edge (b, c)
// adds popcount(adj(b) & adj(c)).  This replaces, per pair (i, j), the materialised
// C = A & row(j) and the stream of row(k) for every k in C: on rows inside dense communities
// that is |C| rows per pair instead of one.  GROUP_BLOCK: the block's four waves share one
// row and its LDS (long rows); otherwise each wave takes its own row (rows up to 128).
template <uint32_t LMAX>
constexpr uint32_t local4_slice(bool group_block) {
  // the task slot, the scan metadata of each wave of the group, the row's list, hash keys and
  // positions (2 * LMAX each) and its bit-rows (LMAX x LMAX / 32 words)
  return 4 + LOCAL4_META * (group_block ? 4 : 1) + 5 * LMAX + LMAX * (LMAX / 32);
}

template <uint32_t LMAX, bool GROUP_BLOCK>
__global__ void __launch_bounds__(256) k_local4(CountArgs a, const uint32_t *__restrict__ rows, uint32_t nrows) {
  constexpr uint32_t CAP = 2 * LMAX;
  extern __shared__ uint32_t lds_all[];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  // the group that shares one row: the block, or one wave
  const uint32_t gwave = GROUP_BLOCK ? wave : 0u, gwaves = GROUP_BLOCK ? (blockDim.x >> 6) : 1u;
  const uint32_t gtid = GROUP_BLOCK ? threadIdx.x : lane, gsize = GROUP_BLOCK ? blockDim.x : 64u;
  uint32_t *gbase = lds_all + (GROUP_BLOCK ? 0u : wave * local4_slice<LMAX>(false));
  uint32_t *slot = gbase;
  uint32_t *mbase = gbase + 4 + gwave * LOCAL4_META, *moff = mbase + 65, *mcol = mbase + 130;
  uint32_t *list = gbase + 4 + gwaves * LOCAL4_META, *hkey = list + LMAX, *hpos = hkey + CAP, *adj = hpos + CAP;
  auto gsync = [] {
    if constexpr (GROUP_BLOCK) __syncthreads();
    else wave_sync();
  };
  const bool per_graph = a.graph_of != nullptr;
  uint64_t acc = 0;
  uint32_t bad = 0;  // a hash insert or lookup walked the whole table (reported at the end)
  for (;;) {
    if (gtid == 0) slot[0] = atomicAdd(a.next, 1u);
    gsync();
    const uint32_t r = rfl(slot[0]);
    if (r >= nrows) break;
    const uint32_t i = rows[r];
    const uint32_t d = rfl(a.card[i]);
    const uint32_t *R = a.unbr + (uint64_t)a.aoff[i] * 4;
    const uint32_t wr = (d + 31u) >> 5;  // bit-row stride in words
    uint32_t lg = 4;
    while ((1u << lg) < 2u * d) lg++;
    const uint32_t cap = 1u << lg, mask = cap - 1u;
    for (uint32_t k = gtid; k < d; k += gsize) list[k] = R[k];
    for (uint32_t w = gtid; w < cap; w += gsize) hkey[w] = HASH_EMPTY;
    for (uint32_t w = gtid; w < d * wr; w += gsize) adj[w] = 0u;
    if (gtid == 0) slot[1] = 0u;  // set when an insert finds no free slot
    gsync();
    for (uint32_t k = gtid; k < d; k += gsize) {
      const uint32_t x = list[k];
      uint32_t h = ((x * 2654435761u) >> (32u - lg)) & mask, it = 0;
      while (it <= mask && atomicCAS(&hkey[h], HASH_EMPTY, x) != HASH_EMPTY) {
        h = (h + 1u) & mask;
        it++;
      }
      if (it > mask) slot[1] = 1u;
      else hpos[h] = k;
    }
    gsync();
    // a table with a failed insert is never probed (the lookups below rely on a free slot)
    if (slot[1]) {
      bad = 1u;
      gsync();
      continue;
    }
    // local edges: stream row(R[c]) for c in [1, d), wave w of the group taking every
    // gwaves-th run of 64
    for (uint32_t c0 = 1 + 64u * gwave; c0 < d; c0 += 64u * gwaves) {
      const uint32_t c = c0 + lane;
      uint32_t items = 0, off = 0;
      if (c < d) {
        const uint32_t j = list[c];
        items = (a.card[j] + 3u) >> 2;
        off = a.aoff[j];
      }
      uint32_t tot;
      const uint32_t bse = wave_excl_scan(items, lane, &tot);
      tot = rfl(tot);
      mbase[lane] = bse;
      moff[lane] = off;
      mcol[lane] = c;
      if (lane == 0) mbase[64] = tot;
      wave_sync();
      for (uint32_t t0 = lane; t0 < tot; t0 += 64u * 4u) {
        uint4 v[4];
        uint32_t col[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const uint32_t t = min(t0 + 64u * u, tot - 1u);
          uint32_t q = 0;
#pragma unroll
          for (uint32_t step = 32; step; step >>= 1)
            if (mbase[q + step] <= t) q += step;
          v[u] = reinterpret_cast<const uint4 *>(a.unbr)[(uint64_t)moff[q] + (t - mbase[q])];
          col[u] = mcol[q];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (t0 + 64u * u >= tot) continue;
          const uint32_t xs[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const uint32_t x = xs[e];
            if (x == HASH_EMPTY) continue;  // the list padding
            // at most d ids in 2^lg >= 2d slots, every insert placed: the walk meets a free slot
            uint32_t h = ((x * 2654435761u) >> (32u - lg)) & mask;
            for (;;) {
              const uint32_t key = hkey[h];
              if (key == x) {
                const uint32_t pp = hpos[h];
                atomicOr(&adj[col[u] * wr + (pp >> 5)], 1u << (pp & 31u));
                break;
              }
              if (key == HASH_EMPTY) break;
              h = (h + 1u) & mask;
            }
          }
        }
      }
      wave_sync();
    }
    gsync();
    // triangles of the local graph: every (c, word of adj(c)) unit walks its bits b
    uint64_t rc = 0;
    for (uint32_t u = gtid; u < d * wr; u += gsize) {
      const uint32_t c = u / wr, w = u - c * wr;
      uint32_t bits = adj[c * wr + w];
      const uint32_t *ac = adj + c * wr;
      while (bits) {
        const uint32_t b = w * 32u + (uint32_t)__builtin_ctz(bits);
        bits &= bits - 1u;
        const uint32_t *ab = adj + b * wr;
        for (uint32_t ww = 0; ww <= (b >> 5); ww++) rc += (uint32_t)__popc(ab[ww] & ac[ww]);
      }
    }
    if (per_graph) {
      const uint64_t sr = wave_sum_u64(rc);
      if (lane == 0 && sr) atomicAdd(&a.counts[a.graph_of[i]], (unsigned long long)sr);
    } else {
      acc += rc;
    }
    // every wave is done with the row's LDS (and has read the slot) before the next row
    gsync();
  }
  if (bad) atomicOr(a.err, ERR_HASH_FULL);
  const uint64_t sum = wave_sum_u64(acc);
  if (lane == 0 && sum) atomicAdd(&a.counts[0], (unsigned long long)sum);
}

// B-stationary pairs above the split: the pairs (i, j) with j at or above the split, for rows
// i whose B rows there all have at most BSH_MAX_D entries.  Every id of row(i) below the
// split is below j too, so the pair's count is |LO(i) & LO(j)| + |HI(i, j) & HI(j)|, with LO a
// row's ids below the split, HI(j) row(j)'s ids at or above it and HI(i, j) those of row(i)
// before j.  A task {bin, b, e} takes the bin's B rows one at a time: LO(j)'s bits go into
// the wave's bitmap over [0, split) and HI(j) into an LDS list; the 16-bit copies of LO(i)
// of the row's in-pairs in [b, e) stream through the bitmap (as in k_bstat) and the vectors
// of row(i) from its first id at or above the split to j's through the list (binary search,
// four ids in lockstep); then LO(j)'s words are cleared again, so a B row costs its own
// length instead of the bitmap's.  The rows are pipelined: a lane keeps ids k, k + 64, ...,
// k + 64 (R - 1) of row(j) in registers for setting and clearing them, and the next row and
// its first 64 in-pairs load during the current row's passes.  The 16-bit copies are padded
// with split - 1, a bit never set here: a pair whose LO(i) and LO(j) both end with split - 1
// counts that id apart.  R is 4 when no B row of the schedule holds more than 256 ids, else 8
// (BSH_MAX_D = 512); the LDS list holds 64 R ids.
template <int R>
__global__ void __launch_bounds__(256) k_bshl(CountArgs a, const uint4 *__restrict__ htasks, uint64_t nht,
                                              const uint4 *__restrict__ hpairs, const uint32_t *__restrict__ hj,
                                              const uint32_t *__restrict__ hjlo, const uint64_t *__restrict__ hoff,
                                              const uint32_t *__restrict__ bstart, uint32_t bm_words, uint32_t pad) {
  static_assert(R >= 1 && 64 * R <= (int)BSH_MAX_D, "row(j) register rounds");
  extern __shared__ uint32_t lds_all[];
  const uint32_t lane = lane_id();
  uint32_t *meta = lds_all + (threadIdx.x >> 6) * a.lds_words;
  uint32_t *bm = meta + META_WORDS, *hl = bm + bm_words;
  const uint16_t *lo16s = reinterpret_cast<const uint16_t *>(a.lo16);
  AState S;  // the bitmap (LO(j)) for the 16-bit streams
  S.bm = bm;
  S.base32 = S.aw0 = S.lo = 0;
  S.anw = bm_words >> 1;
  S.span = 0xFFFFFFFFu;
  S.list = nullptr;
  S.n = 0;
  S.P = 0;
  S.dpre = S.dstart = nullptr;
  S.ht = nullptr;
  S.bad = 0u;
  S.hlg = 0;
  AState H = S;  // the list (HI(j)) for the full-width streams
  H.bm = nullptr;
  H.list = hl;
  for (uint32_t w = lane; w < bm_words; w += 64) bm[w] = 0u;
  wave_sync();
  uint64_t acc = 0;
  uint64_t t = 0, tend = 0;
  for (;;) {
    if (t >= tend) {
      uint32_t t0 = 0;
      if (lane == 0) t0 = atomicAdd(a.next, a.grab);
      t = rfl(t0);
      if (t >= nht) break;
      tend = min(t + a.grab, nht);
    }
    const uint4 task = htasks[t++];
    const uint32_t bin = rfl(task.x), pb = rfl(task.y), pe = rfl(task.z);
    const uint32_t h0 = rfl(bstart[bin]), h1 = rfl(bstart[bin + 1]);
    for (uint32_t hc = h0; hc < h1; hc += 64) {
      const uint32_t nr = min(64u, h1 - hc);
      // the metadata of the bin's B rows (at most BSH_GROUP), one row per lane, loaded together:
      // a row's fetch below is then one dependent load (its ids) instead of a chain of three
      // (row id -> length and offset -> ids), which the one-row-ahead prefetch could not hide
      // behind the few in-pairs most rows have
      uint32_t mq0 = 0, mq1 = 0, md = 0, mlo = 0, mao = 0;
      if (lane < nr) {
        const uint32_t h = hc + lane;
        mq0 = max((uint32_t)hoff[h], pb);
        mq1 = min((uint32_t)hoff[h + 1], pe);
        const uint32_t j = hj[h];
        md = (mq0 < mq1) ? a.card[j] : 0u;  // row(j) has at most 64 R ids
        mlo = hjlo[h];
        mao = a.aoff[j];
      }
      uint32_t xn[R], xc[R];
      uint4 prn = make_uint4(0u, 0u, 0u, 0u);
      uint32_t dn = 0, ln = 0, q0n = 0, q1n = 0;
      auto fetch = [&](uint32_t r) {
        q0n = rdl(mq0, r);
        q1n = rdl(mq1, r);
        dn = rdl(md, r);
        ln = rdl(mlo, r);
        const uint32_t *G = a.unbr + (uint64_t)rdl(mao, r) * 4;
#pragma unroll
        for (int u = 0; u < R; u++) xn[u] = (lane + 64u * u < dn) ? G[lane + 64u * u] : pad;
        prn = (dn && q0n + lane < q1n) ? hpairs[q0n + lane] : make_uint4(0u, 0u, 0u, 0u);
      };
      fetch(0);
      for (uint32_t r = 0; r < nr; r++) {
        const uint32_t dj = dn, dlo = ln, q0 = q0n, q1 = q1n;
        uint4 pr = prn;
#pragma unroll
        for (int u = 0; u < R; u++) xc[u] = xn[u];
        if (r + 1 < nr) fetch(r + 1);
        if (dj == 0) continue;  // no pairs of this row in the task
        const uint32_t dhi = dj - dlo;
        bool has_pad = false;
        if (dlo) {
          // LO(j) is sorted: it ends with the pad id exactly when its last id is split - 1
          const uint32_t qd = (dlo - 1u) >> 6;  // wave-uniform: which register holds it
          uint32_t xl = xc[0];
#pragma unroll
          for (int u = 1; u < R; u++)
            if (qd == (uint32_t)u) xl = xc[u];
          has_pad = __shfl(xl, (int)((dlo - 1u) & 63u), 64) == pad;
        }
#pragma unroll
        for (int u = 0; u < R; u++) {
          const uint32_t k = lane + 64u * u;
          if (k < dlo && xc[u] != pad) atomicOr(&bm[xc[u] >> 5], 1u << (xc[u] & 31u));
          else if (k >= dlo && k < dj) hl[k - dlo] = xc[u];
        }
        wave_sync();
        if (dhi) {
          H.n = dhi;
          H.lo = hl[0];
          H.span = hl[dhi - 1u] - H.lo;
        }
        for (uint32_t c0 = q0; c0 < q1; c0 += 64) {
          const uint4 pnext = (c0 + 64u + lane < q1) ? hpairs[c0 + 64u + lane] : make_uint4(0u, 0u, 0u, 0u);
          if (dlo)
            acc += flat_pass<PASS_STREAM16, KIND_BITMAP, false, 4>(a, S, (pr.y + 7u) >> 3, 0u, pr.x, 0u, lane, meta);
          if (has_pad && pr.y && lo16s[(uint64_t)pr.x * 8 + pr.y - 1u] == pad) acc++;
          if (dhi) {
            // row(i) from the vector of its first id at or above the split to j's vector: the
            // ids below the split and from j on are in no HI(j)
            const uint32_t v0 = pr.y >> 2, nv = (pr.z | pr.w) ? ((pr.w + 3u) >> 2) - v0 : 0u;
            acc += flat_pass<PASS_STREAM, KIND_LIST, false, 4>(a, H, nv, 0u, pr.z + v0, 0u, lane, meta);
          }
          pr = pnext;
        }
        wave_sync();
#pragma unroll
        for (int u = 0; u < R; u++)
          if (lane + 64u * u < dlo && xc[u] != pad) bm[xc[u] >> 5] = 0u;
        wave_sync();
      }
    }
  }
  const uint64_t sum = wave_sum_u64(acc);
  if (lane == 0 && sum) atomicAdd(&a.counts[0], (unsigned long long)sum);
}

template <int K, bool W, bool TUNED, bool BLK>
__global__ void __launch_bounds__(256) k_count(CountArgs a) {
  extern __shared__ uint32_t lds_all[];
  const uint32_t lane = lane_id();
  uint32_t *L = lds_all + (threadIdx.x >> 6) * a.lds_words;
  uint64_t acc = 0;
  uint32_t cur = 0;
  uint64_t t = 0, tend = 0;
  for (;;) {
    if (t >= tend) {
      uint32_t t0 = 0;
      if (lane == 0) t0 = atomicAdd(a.next, a.grab);
      t = rfl(t0);
      if (t >= a.ntasks) break;
      tend = min(t + a.grab, a.ntasks);
    }
    const uint4 task = a.tasks[t];
    const uint32_t i = rfl(task.x), jb = rfl(task.y), je = rfl(task.z), w = rfl(task.w);
    const uint32_t gid = w & GID_MASK, kind = w >> KIND_SHIFT;
    if (gid != cur) {
      const uint64_t s = wave_sum_u64(acc);
      if (lane == 0 && s) atomicAdd(&a.counts[cur], (unsigned long long)s);
      acc = 0;
      cur = gid;
    }
    if (a.task_cycles) {
      const uint64_t c0 = clock64();
      acc += run_task<K, W, TUNED, BLK>(a, kind, i, jb, je, lane, L);
      const uint64_t c1 = clock64();
      if (lane == 0) a.task_cycles[a.task_base + t] = c1 - c0;
    } else {
      acc += run_task<K, W, TUNED, BLK>(a, kind, i, jb, je, lane, L);
    }
    t++;
  }
  const uint64_t s = wave_sum_u64(acc);
  if (lane == 0 && s) atomicAdd(&a.counts[cur], (unsigned long long)s);
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}
__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// ---------------------------------------------------------------- schedule kernels

struct SchedArgs {
  const uint32_t *card;
  const uint32_t *aoff;
  const uint32_t *unbr;
  const RowDesc *desc;
  const uint32_t *graph_of;
  uint32_t n;
  int tuned;
  int k;
  uint32_t lds_small, lds_big;
  uint64_t task_work;
  uint32_t small_deg;     // rows with 1..small_deg entries go to the edge-centric kernel
  uint32_t local_max;     // 4-cliques: rows above small_deg and up to this go to k_local4
  uint32_t bstat_j;       // task rows hand their pairs with j < bstat_j to the B-stationary pass
  uint32_t bsh_max;       // rows whose B rows at or above bstat_j all have at most this many
                          // entries hand those pairs to the B-stationary pass above the split (0: off)
  uint32_t *bpairs;       // out (bstat): row i's B-stationary in-pairs (entries k in [1, below)
                          // with a non-empty row(j)), one word per row
  uint32_t *hcount;       // out (bsh): such rows containing j, per j >= bstat_j (index j - bstat_j)
  uint32_t *jstart;       // out: first entry of row i its A-side tasks cover
  uint32_t *jbelow;       // out: entries of row i below bstat_j
  uint32_t *is_small;     // out: 1 for those rows, 2 / 3 for the k_local4 classes
  uint64_t *sortkey;      // out: (i << 32) | (class << 30) | order, in row order (a stable sort on
                          // the low 32 bits orders the rows by class, then order, then i)
  uint32_t *ntask;        // out: tasks of row i
  uint32_t *rowinfo;      // out: kind << 28 | graph id
  unsigned long long *stats;  // [0..3] pair types, [4] work units, [5] wedges
};

// one wave per row: per-row work and pair statistics
__global__ void __launch_bounds__(256) k_row_work(SchedArgs s) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  unsigned long long loc[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t i = wid; i < s.n; i += nw) {
    const uint32_t dA = s.card[i];
    uint64_t work = 0;
    uint32_t lo = 0, hi = 0, below = 0, bad = 0, bp = 0;
    bool full = false;
    // with the B-stationary pass on, every row's pairs with j < bstat_j (a prefix of the
    // sorted row) leave its A-side tasks or its edge-centric wedges
    if (dA) {
      const uint32_t *G = s.unbr + (uint64_t)s.aoff[i] * 4;
      lo = G[0];
      hi = G[dA - 1];
      for (uint32_t k = lane; k < dA; k += 64) {
        const uint32_t jv = G[k];
        const RowDesc d = s.desc[jv];
        const uint32_t t = desc_type(d);
        uint64_t units;
        if (d.card == 0) { loc[3]++; units = 0; }
        else if (is_block_type(t)) { loc[2]++; units = d.card; }  // block rows count with the bitsets
        else if (is_uint_type(t)) { loc[0]++; units = d.card; }
        else if (t == EH_PS) { loc[1]++; units = desc_aux(d); }
        else { loc[2]++; units = desc_aux(d); }
        loc[5] += d.card;
        if (jv < s.bstat_j) {
          below++;
          // the row's first entry has an empty prefix: nothing to count
          bp += (d.card && k > 0) ? 1u : 0u;
          continue;
        }
        if (d.card > s.bsh_max) bad = 1u;
        work += units + 16u;
      }
      work = wave_sum_u64(work);
      below = (uint32_t)wave_sum_u64(below);
      bp = (uint32_t)wave_sum_u64(bp);
      loc[4] += (lane == 0) ? work - 16ull * (dA - below) : 0ull;
      // every pair at or above the split goes B-stationary (k_bshl) when no B row there is longer
      // than bsh_max: the row then has no A-side or edge-centric work left
      full = s.bstat_j && s.bsh_max && __ballot(bad != 0u) == 0ull;
      if (full && below < dA) {
        for (uint32_t k = max(below, 1u) + lane; k < dA; k += 64) {
          const uint32_t jv = G[k];
          if (s.desc[jv].card) atomicAdd(&s.hcount[jv - s.bstat_j], 1u);
        }
      }
    }
    if (lane != 0) continue;
    if (s.bpairs) s.bpairs[i] = bp;
    const uint32_t js = full ? dA : below;
    s.jstart[i] = js;
    s.jbelow[i] = below;
    below = js;  // from here on: the entries before the A-side / edge-centric part
    const bool small = dA > 0 && dA <= s.small_deg && below < dA;
    const bool local = !small && dA > 2 && dA <= s.local_max;
    s.is_small[i] = small ? 1u : local ? (dA <= LOCAL4_MID ? 2u : 3u) : 0u;
    if (dA == 0 || small || local || below == dA) {
      s.sortkey[i] = ((uint64_t)i << 32) | (3u << 30);
      s.ntask[i] = 0;
      s.rowinfo[i] = 0;
      continue;
    }
    // staging kind and LDS need
    const uint32_t anw = (hi >> 6) - (lo >> 6) + 1u;
    const RowDesc da = s.desc[i];
    const uint32_t ta = desc_type(da);
    const uint32_t extra = (s.k >= 4) ? 1u : 0u;  // the 4-clique level set needs room too
    const uint64_t need_bitmap = 2ull * anw + dA + extra * (2ull * anw + dA);
    const uint64_t need_list = (uint64_t)dA + extra * dA;
    const uint64_t need_psdir = 3ull * dA + 1 + extra * dA;
    uint32_t kind, cls;
    uint64_t need;
    if (s.tuned) {
      // flattened schedule: LDS holds the pair metadata, the list and a bitmap or
      // hash table of the source row (twice for the 4-clique level set)
      const uint64_t cap = 1ull << hash_lg(dA);
      const uint64_t rep = 1ull + extra;
      const uint64_t nb = META_WORDS + rep * (2ull * anw + dA);
      const uint64_t nh = META_WORDS + rep * (cap + dA);
      const uint64_t nl = META_WORDS + rep * dA;
      // everything that can stays in the small slice (the large-LDS class runs at
      // a fraction of the occupancy): O(1) membership first, then a staged list,
      // then binary search in HBM for triangles
      if (nb <= s.lds_small && 2ull * anw <= 4ull * cap) { kind = KIND_BITMAP; need = nb; }
      else if (nh <= s.lds_small) { kind = KIND_HASH; need = nh; }
      else if (nb <= s.lds_small) { kind = KIND_BITMAP; need = nb; }
      else if (nl <= s.lds_small) { kind = KIND_LIST; need = nl; }
      else if (s.k == 3) { kind = KIND_GLOBAL; need = 0; }
      else if (nh <= s.lds_big) { kind = KIND_HASH; need = nh; }
      else { kind = KIND_LIST; need = nl; }
    } else {
      if (ta == EH_BS || is_block_type(ta)) { kind = KIND_BITMAP; need = need_bitmap; }
      else if (ta == EH_PS) { kind = KIND_PSDIR; need = need_psdir; }
      else { kind = KIND_LIST; need = need_list; }
    }
    if (need <= s.lds_small) cls = 1;
    else if (need <= s.lds_big) cls = 0;
    else if (s.k == 3) { kind = KIND_GLOBAL; cls = 1; }
    else { kind = KIND_GLOBAL; cls = 2; }  // unsupported: flagged by the host
    const uint32_t gid = s.graph_of ? s.graph_of[i] : 0u;
    s.rowinfo[i] = (kind << KIND_SHIFT) | gid;
    uint32_t nt = 1;
    if (s.tuned && s.task_work) {
      uint64_t q = (work + s.task_work - 1) / s.task_work;
      nt = (uint32_t)min<uint64_t>(max<uint64_t>(q, 1ull), (uint64_t)(dA - below));
    }
    s.ntask[i] = nt;
    const uint64_t order = s.tuned ? (0x3FFFFFFFull - min<uint64_t>(work, 0x3FFFFFFFull)) : 0ull;
    s.sortkey[i] = ((uint64_t)i << 32) | ((uint64_t)cls << 30) | order;
  }
  // the statistics: summed over the block's waves in LDS, then one atomic per block and value
  // (one per wave had every wave of the grid queue on the same six words)
  __shared__ unsigned long long part[4][6];
  const uint32_t wv = threadIdx.x >> 6;
  for (int q = 0; q < 6; q++) {
    const unsigned long long v = wave_sum_u64(loc[q]);
    if (lane == 0) part[wv][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    unsigned long long v = 0;
    for (uint32_t w = 0; w < (blockDim.x >> 6); w++) v += part[w][threadIdx.x];
    if (v) atomicAdd(&s.stats[threadIdx.x], v);
  }
}

__global__ void k_gather_ntask(const uint64_t *sorted, const uint32_t *ntask, uint32_t n, uint32_t *nt_sorted,
                               uint32_t *nbig_rows) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const uint64_t key = sorted[q];
    nt_sorted[q] = ntask[(uint32_t)(key >> 32)];
    const uint32_t cls = ((uint32_t)key >> 30) & 3u;
    if (cls == 0) atomicAdd(nbig_rows, 1u);
    if (cls == 2) atomicAdd(nbig_rows + 1, 1u);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) nt_sorted[n] = 0;
}

__global__ void k_expand_tasks(const uint64_t *sorted, const uint32_t *nt_sorted, const uint64_t *toff,
                               const uint32_t *card, const uint32_t *jstart, const uint32_t *rowinfo, uint32_t n,
                               uint4 *tasks) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const uint32_t nt = nt_sorted[q];
    if (!nt) continue;
    const uint32_t i = (uint32_t)(sorted[q] >> 32);
    const uint32_t js = jstart[i];
    const uint32_t d = card[i] - js;
    const uint64_t base = toff[q];
    for (uint32_t p = 0; p < nt; p++) {
      const uint32_t jb = js + (uint32_t)(((uint64_t)p * d) / nt);
      const uint32_t je = js + (uint32_t)(((uint64_t)(p + 1) * d) / nt);
      tasks[base + p] = make_uint4(i, jb, je, rowinfo[i]);
    }
  }
}

// B-stationary in-pairs: every task row writes (its offset / 4, position) into the list
// of each non-empty B row j < bstat_j it holds (one wave per row, positions [0, jstart))
// every row writes the keys of its B-stationary in-pairs at its own offset (rowpos, the scan of
// the per-row counts): no atomics, so the hub rows' in-lists do not serialise on counters; the
// radix sort that follows groups the keys by j
__global__ void __launch_bounds__(256) k_bstat_fill(const uint32_t *aoff, const uint32_t *unbr, const RowDesc *desc,
                                                    const uint32_t *jbelow, uint32_t n, const uint64_t *rowpos,
                                                    const uint64_t *lo16off, uint64_t *keys) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t i = wid; i < n; i += nw) {
    const uint32_t jb = jbelow[i];
    if (jb <= 1) continue;
    const uint32_t *G = unbr + (uint64_t)aoff[i] * 4;
    const uint64_t pre = (uint64_t)(uint32_t)lo16off[i];
    uint64_t pos = rowpos[i];
    for (uint32_t k0 = 0; k0 < jb; k0 += 64) {
      const uint32_t k = k0 + lane;
      uint32_t j = 0;
      bool q = false;
      if (k < jb && k > 0) {
        j = G[k];
        q = desc[j].card != 0;
      }
      const uint64_t m = __ballot(q);
      // p (16 bits, p < j < 2^16) | j (16 bits) | row i's prefix vector offset (32 bits),
      // written in row order: a stable sort on j's bits alone leaves every in-list in row order
      if (q) keys[pos + mbcnt(m)] = ((uint64_t)k << 48) | ((uint64_t)j << 32) | pre;
      pos += (uint64_t)__popcll(m);
    }
  }
}

// in-list bounds from the sorted keys: inoff[j] = the first key of j, incount[j] = how many
// (incount starts at zero; j without keys keep count 0)
__global__ void k_bstat_bounds(const uint64_t *keys, uint64_t np, uint64_t *inoff, uint32_t *inend) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < np; q += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = (uint32_t)(keys[q] >> 32) & 0xFFFFu;
    if (q == 0 || ((uint32_t)(keys[q - 1] >> 32) & 0xFFFFu) != j) inoff[j] = q;
    if (q + 1 == np || ((uint32_t)(keys[q + 1] >> 32) & 0xFFFFu) != j) inend[j] = (uint32_t)(q + 1);
  }
}

__global__ void k_bstat_counts(const uint64_t *inoff, const uint32_t *inend, uint32_t nj, uint32_t *incount) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += gridDim.x * blockDim.x)
    incount[j] = inend[j] ? (uint32_t)(inend[j] - inoff[j]) : 0u;
}

__global__ void k_bstat_unpack(const uint64_t *keys, uint64_t np, uint2 *in_pairs) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < np; q += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = keys[q];
    in_pairs[q] = make_uint2((uint32_t)x, (uint32_t)(x >> 48));
  }
}

// task order keys: the first in-pair's prefix offset (row order) | task index
__global__ void k_bstat_task_keys(const uint4 *btasks, uint64_t nbt, const uint2 *in_pairs, uint64_t *tkeys) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < nbt; t += (uint64_t)gridDim.x * blockDim.x)
    tkeys[t] = ((uint64_t)in_pairs[btasks[t].y].x << 32) | t;
}

__global__ void k_bstat_task_gather(const uint4 *btasks, uint64_t nbt, const uint64_t *tkeys, uint4 *out) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < nbt; t += (uint64_t)gridDim.x * blockDim.x)
    out[t] = btasks[tkeys[t] & 0xFFFFFFFFu];
}

// ---- the B-stationary schedule above the split: B rows with in-pairs, their bins, in-pairs, tasks
__global__ void k_nonzero_flag(const uint32_t *c, uint32_t n, uint32_t *flag) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    flag[q] = c[q] ? 1u : 0u;
}

// compacted B rows (in row order), their pair counts and table-word weights
__global__ void k_bsh_rows(const uint32_t *hcount, const uint64_t *hpos, uint32_t nh, uint32_t split,
                           const uint32_t *card, const uint32_t *jbelow, uint32_t *hj, uint32_t *hjlo, uint32_t *hc,
                           uint32_t *hw, uint32_t *longest) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nh; q += gridDim.x * blockDim.x) {
    if (!hcount[q]) continue;
    const uint64_t h = hpos[q];
    const uint32_t j = split + q, d = card[j] - jbelow[j];  // the row's ids at or above the split
    atomicMax(longest, card[j]);
    uint32_t lg = 4;
    while ((1u << lg) < 2u * d) lg++;
    hj[h] = j;
    hjlo[h] = jbelow[j];
    hc[h] = hcount[q];
    hw[h] = (1u << lg) + BSH_BIN_WEIGHT / BSH_GROUP;
  }
}

__global__ void k_bsh_bin_flag(const uint64_t *pw, uint32_t nhr, uint32_t *flag) {
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nhr; h += gridDim.x * blockDim.x)
    flag[h] = (h == 0 || pw[h] / BSH_BIN_WEIGHT != pw[h - 1] / BSH_BIN_WEIGHT) ? 1u : 0u;
}

__global__ void k_bsh_bin_start(const uint32_t *flag, const uint64_t *bpos, uint32_t nhr, uint32_t *bstart) {
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nhr; h += gridDim.x * blockDim.x)
    if (flag[h]) bstart[bpos[h]] = h;
}

// every row whose pairs at or above the split all go B-stationary (jstart == card)
// writes (its offset / 16 bytes, position) into the list of each non-empty B row there
__global__ void __launch_bounds__(256) k_bsh_fill(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr,
                                                  const RowDesc *desc, const uint32_t *jstart,
                                                  const uint32_t *jbelow, uint32_t n, uint32_t split,
                                                  const uint64_t *hpos, const uint64_t *hoff, uint32_t *hfill,
                                                  const uint64_t *lo16off, uint4 *hpairs) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t i = wid; i < n; i += nw) {
    const uint32_t d = card[i], jb = jbelow[i];
    if (jstart[i] != d || jb >= d) continue;
    const uint32_t *G = unbr + (uint64_t)aoff[i] * 4;
    for (uint32_t k = max(jb, 1u) + lane; k < d; k += 64) {
      const uint32_t j = G[k];
      if (desc[j].card == 0) continue;
      const uint64_t h = hpos[j - split];
      const uint32_t slot = atomicAdd(&hfill[h], 1u);
      hpairs[hoff[h] + slot] = make_uint4((uint32_t)lo16off[i], jb, aoff[i], k);
    }
  }
}

__global__ void k_bsh_bin_ntask(const uint32_t *bstart, const uint64_t *hoff, uint32_t nb, uint32_t chunk,
                                uint32_t *ntb) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
    const uint64_t np = hoff[bstart[b + 1]] - hoff[bstart[b]];
    ntb[b] = (uint32_t)((np + chunk - 1) / chunk);
  }
}

__global__ void k_bsh_bin_expand(const uint32_t *bstart, const uint64_t *hoff, const uint64_t *boff, uint32_t nb,
                                 uint32_t chunk, uint4 *htasks) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
    const uint64_t p0 = hoff[bstart[b]], p1 = hoff[bstart[b + 1]], t0 = boff[b];
    for (uint64_t q = 0; p0 + q * chunk < p1; q++) {
      const uint32_t pb = (uint32_t)(p0 + q * chunk), pe = (uint32_t)min(p1, p0 + (q + 1) * chunk);
      htasks[t0 + q] = make_uint4(b, pb, pe, 0u);
    }
  }
}

// 16-bit copies of the B-stationary prefixes: row i's first jstart[i] entries (all below
// the split) in ceil(jstart / 8) 16-byte vectors, the tail padded with split - 1
// 16-bit prefix copy layout: rows with at least LO16_ALIGN_MIN ids below the split first, each
// starting on a 128-byte line (their prefixes are read by many pairs, and a prefix that starts
// mid-line costs one more L1->L2 request: 12% fewer lines on RMAT-22), then the shorter rows
// packed at 16 bytes
constexpr uint32_t LO16_ALIGN_MIN = 16;

__global__ void k_lo16_count(const uint32_t *jstart, uint32_t n, uint32_t *nbig, uint32_t *nsmall) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t js = jstart[i], v = (js + 7u) >> 3;
    const bool big = js >= LO16_ALIGN_MIN;
    nbig[i] = big ? (v + 7u) & ~7u : 0u;
    nsmall[i] = big ? 0u : v;
  }
}

// vector offset of every row's copy; entry n is the total
__global__ void k_lo16_place(const uint32_t *jstart, uint32_t n, const uint64_t *obig, const uint64_t *osmall,
                             uint64_t *lo16off) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x)
    lo16off[i] = (i < n && jstart[i] >= LO16_ALIGN_MIN) ? obig[i] : obig[n] + osmall[i];
}

__global__ void __launch_bounds__(256) k_lo16_fill(const uint32_t *aoff, const uint32_t *unbr, const uint32_t *jstart,
                                                   uint32_t n, const uint64_t *lo16off, uint32_t pad,
                                                   uint16_t *lo16) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t i = wid; i < n; i += nw) {
    const uint32_t js = jstart[i];
    if (js == 0) continue;
    const uint32_t *G = unbr + (uint64_t)aoff[i] * 4;
    uint16_t *D = lo16 + lo16off[i] * 8;
    const uint32_t len = ((js + 7u) >> 3) * 8u;
    for (uint32_t k = lane; k < len; k += 64) D[k] = (uint16_t)(k < js ? G[k] : pad);
  }
}

// B-stationary tasks: the in-list of row j cut into pieces of at most `chunk` pairs
__global__ void k_bstat_ntask(const uint32_t *incount, uint32_t nj, uint32_t chunk, uint32_t *ntb) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += gridDim.x * blockDim.x)
    ntb[j] = (incount[j] + chunk - 1u) / chunk;
}

// task {j | |row(j)| << 16, b, e, row(j)'s offset in unbr}: j < 2^16 and row(j) holds ids
// below j, so both fit in the first word
__global__ void k_bstat_expand(const uint32_t *incount, const uint64_t *inoff, const uint64_t *boff, uint32_t nj,
                               uint32_t chunk, const uint32_t *card, const uint32_t *aoff, uint4 *btasks) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += gridDim.x * blockDim.x) {
    const uint32_t c = incount[j];
    const uint64_t b0 = inoff[j], t0 = boff[j];
    const uint32_t head = j | (card[j] << 16), ao = aoff[j];
    for (uint32_t p = 0; p * chunk < c; p++) {
      const uint32_t b = (uint32_t)(b0 + (uint64_t)p * chunk);
      const uint32_t e = (uint32_t)(b0 + min<uint64_t>((uint64_t)(p + 1) * chunk, c));
      btasks[t0 + p] = make_uint4(head, b, e, ao);
    }
  }
}

__global__ void k_class_flag(const uint32_t *cls, uint32_t n, uint32_t c, uint32_t *flag) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    flag[i] = cls[i] == c ? 1u : 0u;
}

// per short row of a batch arena's list, its edge-centric items C(d, 2) (the wedges
// k_small_rows tests, the pairs of k_small_cliques4's local graph): the XCD queues' weights
__global__ void k_short_items(const uint32_t *rows, uint32_t nr, const uint32_t *card, uint32_t *items) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < nr; p += gridDim.x * blockDim.x) {
    const uint32_t d = card[rows[p]];
    items[p] = (d * (d - 1u)) / 2u;
  }
}

// the bounds of the 8 XCD queues over a short-row list of nr rows: queue q starts at the
// first row whose exclusive item prefix reaches q/8 of the total (pre holds nr + 1 prefixes)
__global__ void k_short_cuts(const uint64_t *pre, uint32_t nr, uint32_t *xq) {
  const uint32_t q = threadIdx.x;
  if (q > 8u) return;
  const uint64_t target = (pre[nr] / 8u) * q + ((pre[nr] % 8u) * q) / 8u;
  uint32_t lo = 0, hi = nr;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (pre[mid] < target) lo = mid + 1u; else hi = mid;
  }
  xq[q] = (q == 8u) ? nr : lo;
}

__global__ void k_scatter_small(const uint32_t *is_small, const uint64_t *pos, uint32_t n, uint32_t *rows) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (is_small[i]) rows[pos[i]] = i;
}

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = ceil_div(work ? work : 1, block);
  return (unsigned)(g > 65536 ? 65536 : g);
}

struct SchedStats {
  unsigned long long v[6];
};

}  // namespace

// ---------------------------------------------------------------- host side

TuneParams tune_defaults(eh_matrix *m, int sched, int k, const eh_run_opts *opts) {
  TuneParams tp;
  tp.lds_small = 2048;
  tp.lds_big = 16384;
  tp.big_waves = 1;
  tp.waves_per_block = sched == EH_SCHED_TUNED ? 2 : 4;
  tp.lanes = 64;
  if (sched == EH_SCHED_TUNED) {
    uint64_t waves = (uint64_t)(m->ctx->num_cus ? m->ctx->num_cus : 256) * 16;
    uint64_t tw = m->wedges / (waves * 8 + 1);
    tp.task_work = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(tw, 2048), 1u << 26);
  } else {
    tp.task_work = 0;
  }
  // rows this short are cheaper as explicit wedges (triangles) or triples (4-cliques)
  tp.small_deg = (sched == EH_SCHED_TUNED && m->filter == EH_FILTER_LOWER) ? 32u : 0u;
  // B rows below 64Ki ids B-stationary (schedule_build applies it to tuned triangle passes
  // over oriented single-graph matrices only)
  tp.bstat_j = (sched == EH_SCHED_TUNED && k == 3) ? 65536u : 0u;
  tp.bsh_max = tp.bstat_j ? BSH_MAX_D : 0u;
  // triangle tasks are many and light: two per queue atomic; 4-clique tasks are few and
  // heavy, where pulling two at once unbalances the end of the pass
  tp.grab = (k == 4) ? 1u : 2u;
  tp.local_max = (sched == EH_SCHED_TUNED && k == 4 && m->filter == EH_FILTER_LOWER) ? LOCAL4_MAX : 0u;
  if (opts) {
    if (opts->small_degree < 0) tp.small_deg = 0;
    else if (opts->small_degree > 0 && tp.small_deg) tp.small_deg = (uint32_t)std::min(opts->small_degree, k == 4 ? 64 : 256);
    if (opts->task_work) tp.task_work = opts->task_work;
    if (opts->lds_words) tp.lds_small = opts->lds_words;
    if (opts->waves_per_block > 0) tp.waves_per_block = opts->waves_per_block;
    if (opts->lanes_per_row > 0) tp.lanes = opts->lanes_per_row;
    if (opts->stream_ratio < 0) tp.stream_ratio = 0;
    else if (opts->stream_ratio > 0) tp.stream_ratio = (uint32_t)opts->stream_ratio;
    if (opts->bstat_split < 0) tp.bstat_j = tp.bsh_max = 0;
    if (opts->bsh_max < 0) tp.bsh_max = 0;
    else if (opts->bsh_max > 0 && tp.bsh_max) tp.bsh_max = (uint32_t)std::min<int32_t>(opts->bsh_max, (int32_t)BSH_MAX_D);
    if (opts->local_max < 0) tp.local_max = 0;
    else if (opts->local_max > 0 && tp.local_max)
      tp.local_max = (uint32_t)std::min<int32_t>(opts->local_max, (int32_t)LOCAL4_MAX);
    if (opts->bstat_split > 0) tp.bstat_j = (uint32_t)std::min(std::max(opts->bstat_split, 64), 65536);
    if (opts->short_queues < 0) tp.short_queues = false;
  }
  // the task kernels are built for at most 256 threads (4 waves) per block, and a
  // wave's slice holds at most 16384 words (64 KB)
  // the flattened passes keep META_WORDS of pair metadata in every slice
  if (tp.lds_small < 512) tp.lds_small = 512;
  if (tp.lds_small > 16384) tp.lds_small = 16384;
  if (tp.waves_per_block < 1) tp.waves_per_block = 1;
  if (tp.waves_per_block > 4) tp.waves_per_block = 4;
  while (tp.waves_per_block > 1 && (uint64_t)tp.lds_small * tp.waves_per_block * 4 > 64 * 1024) tp.waves_per_block--;
  if (tp.lds_big < tp.lds_small) tp.lds_big = tp.lds_small;
  return tp;
}

eh_status schedule_build(eh_matrix *m, int sched, int k, const TuneParams &tp, Schedule &s) {
  eh_ctx *ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t n = m->n;
  DevBuf key, key2, ntask, rowinfo, stats, ntsorted, toff, cnt2, is_small, spos, bpairs;
  DevBuf &jstart = s.jstart;
  // B-stationary pairs: tuned triangles on an oriented single-graph matrix only (the
  // prefix identity needs every row(j) below j)
  const uint32_t bstat_j = (sched == EH_SCHED_TUNED && k == 3 && m->filter == EH_FILTER_LOWER &&
                            !m->dgraph_of.p) ? std::min(tp.bstat_j, n) : 0u;
  EH_CHECK(jstart.alloc(4ull * n + 8));  // kept with the schedule
  DevBuf jbelow, hcount;
  EH_CHECK(jbelow.alloc(4ull * n + 8, &ctx->cache));
  // rows whose B rows at or above the split are all short go wholly B-stationary (hashed)
  const uint32_t bsh_max = (bstat_j && bstat_j < n) ? tp.bsh_max : 0u;
  const uint32_t nh = bsh_max ? n - bstat_j : 0u;
  if (bstat_j) EH_CHECK(bpairs.alloc(4ull * n + 8, &ctx->cache));
  if (nh) {
    EH_CHECK(hcount.alloc(4ull * nh + 8, &ctx->cache));
    k_zero_u32<<<grid_for((uint64_t)nh + 1), 256, 0, st>>>(hcount.as<uint32_t>(), nh + 1);
  }
  EH_CHECK(is_small.alloc(4ull * n + 8, &ctx->cache));
  EH_CHECK(spos.alloc(8ull * n + 16, &ctx->cache));
  k_zero_u32<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(is_small.as<uint32_t>(), n + 1);
  EH_CHECK(key.alloc(8ull * n + 8, &ctx->cache));
  EH_CHECK(key2.alloc(8ull * n + 8, &ctx->cache));
  EH_CHECK(ntask.alloc(4ull * n + 4, &ctx->cache));
  EH_CHECK(rowinfo.alloc(4ull * n + 4, &ctx->cache));
  EH_CHECK(stats.alloc(64, &ctx->cache));
  EH_CHECK(cnt2.alloc(8, &ctx->cache));
  EH_HIP(hipEventRecord(ctx->ev0, st));
  k_zero_u64<<<1, 64, 0, st>>>(stats.as<unsigned long long>(), 8);
  k_zero_u32<<<1, 64, 0, st>>>(cnt2.as<uint32_t>(), 2);
  SchedArgs sa;
  sa.card = m->dcard.as<uint32_t>();
  sa.aoff = m->daoff.as<uint32_t>();
  sa.unbr = m->dunbr.as<uint32_t>();
  sa.desc = m->ddesc.as<RowDesc>();
  sa.graph_of = m->dgraph_of.p ? m->dgraph_of.as<uint32_t>() : nullptr;
  sa.n = n;
  sa.tuned = sched == EH_SCHED_TUNED;
  sa.k = k;
  sa.lds_small = tp.lds_small;
  sa.lds_big = tp.lds_big;
  sa.task_work = tp.task_work;
  sa.small_deg = tp.small_deg;
  const uint32_t local_max = (sched == EH_SCHED_TUNED && k == 4) ? tp.local_max : 0u;
  sa.local_max = local_max;
  sa.bstat_j = bstat_j;
  sa.bsh_max = bsh_max;
  sa.bpairs = bstat_j ? bpairs.as<uint32_t>() : nullptr;
  sa.hcount = nh ? hcount.as<uint32_t>() : nullptr;
  sa.jstart = jstart.as<uint32_t>();
  sa.jbelow = jbelow.as<uint32_t>();
  sa.is_small = is_small.as<uint32_t>();
  sa.sortkey = key.as<uint64_t>();
  sa.ntask = ntask.as<uint32_t>();
  sa.rowinfo = rowinfo.as<uint32_t>();
  sa.stats = stats.as<unsigned long long>();
  // a persistent grid: every block adds its statistics once
  k_row_work<<<std::min(grid_for((uint64_t)n * 64), (unsigned)(ctx->num_cus ? ctx->num_cus : 256) * 8u), 256, 0, st>>>(sa);
  EH_HIP(hipGetLastError());
  // rows of each k_local4 class, in row order
  for (int c = 0; c < 2; c++) {
    s.n_local[c] = 0;
    if (!local_max) continue;
    DevBuf flag;
    EH_CHECK(flag.alloc(4ull * n + 8, &ctx->cache));
    k_class_flag<<<grid_for(n), 256, 0, st>>>(is_small.as<uint32_t>(), n, 2u + c, flag.as<uint32_t>());
    k_zero_u32<<<1, 64, 0, st>>>(flag.as<uint32_t>() + n, 1);
    EH_CHECK(prim_exscan_u32_to_u64(ctx, flag.as<uint32_t>(), spos.as<uint64_t>(), (uint64_t)n + 1));
    uint64_t nl = 0;
    EH_HIP(hipMemcpyAsync(&nl, spos.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    EH_CHECK(s.local_rows[c].alloc(4 * nl + 16));
    if (nl)
      k_scatter_small<<<grid_for(n), 256, 0, st>>>(flag.as<uint32_t>(), spos.as<uint64_t>(), n,
                                                  s.local_rows[c].as<uint32_t>());
    s.n_local[c] = nl;
  }
  k_class_flag<<<grid_for(n), 256, 0, st>>>(is_small.as<uint32_t>(), n, 1u, is_small.as<uint32_t>());
  EH_CHECK(prim_exscan_u32_to_u64(ctx, is_small.as<uint32_t>(), spos.as<uint64_t>(), (uint64_t)n + 1));
  uint64_t nsmall = 0;
  EH_HIP(hipMemcpyAsync(&nsmall, spos.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  EH_CHECK(s.small_rows.alloc(4 * nsmall + 16));
  if (nsmall)
    k_scatter_small<<<grid_for(n), 256, 0, st>>>(is_small.as<uint32_t>(), spos.as<uint64_t>(), n,
                                                s.small_rows.as<uint32_t>());
  s.n_small_rows = nsmall;
  s.small_deg = tp.small_deg;
  // batch arenas: the list (in row order, so graph by graph) cut into 8 ranges of equal
  // items, one queue per XCD; a range then keeps its L2 on the few graphs it is working on
  s.xq.release();
  if (nsmall && m->dgraph_of.p) {
    DevBuf items;
    EH_CHECK(items.alloc(4ull * nsmall + 8, &ctx->cache));
    k_short_items<<<grid_for(nsmall), 256, 0, st>>>(s.small_rows.as<uint32_t>(), (uint32_t)nsmall,
                                                    m->dcard.as<uint32_t>(), items.as<uint32_t>());
    k_zero_u32<<<1, 64, 0, st>>>(items.as<uint32_t>() + nsmall, 1);
    EH_CHECK(prim_exscan_u32_to_u64(ctx, items.as<uint32_t>(), spos.as<uint64_t>(), nsmall + 1));
    EH_CHECK(s.xq.alloc(4 * XQ_WORDS));
    k_zero_u32<<<1, 256, 0, st>>>(s.xq.as<uint32_t>(), XQ_WORDS);
    k_short_cuts<<<1, 64, 0, st>>>(spos.as<uint64_t>(), (uint32_t)nsmall, s.xq.as<uint32_t>());
    EH_HIP(hipGetLastError());
  }
  EH_CHECK(prim_sort_u64(ctx, key.as<uint64_t>(), key2.as<uint64_t>(), n, 32));
  EH_CHECK(ntsorted.alloc(4ull * n + 8, &ctx->cache));
  EH_CHECK(toff.alloc(8ull * n + 16, &ctx->cache));
  k_gather_ntask<<<grid_for(n), 256, 0, st>>>(key.as<uint64_t>(), ntask.as<uint32_t>(), n,
                                             ntsorted.as<uint32_t>(), cnt2.as<uint32_t>());
  EH_CHECK(prim_exscan_u32_to_u64(ctx, ntsorted.as<uint32_t>(), toff.as<uint64_t>(), (uint64_t)n + 1));
  uint64_t ntasks = 0;
  uint32_t hcnt[2] = {0, 0};
  EH_HIP(hipMemcpyAsync(&ntasks, toff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st));
  EH_HIP(hipMemcpyAsync(hcnt, cnt2.p, 8, hipMemcpyDeviceToHost, st));
  SchedStats hs;
  EH_HIP(hipMemcpyAsync(hs.v, stats.p, 48, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  if (hcnt[1]) {
    eh_set_error("%u rows need more LDS than the %u-word budget for a %d-clique pass", hcnt[1], tp.lds_big, k);
    return EH_ERR_CAPACITY;
  }
  uint64_t nbig = 0;
  if (hcnt[0]) EH_HIP(hipMemcpy(&nbig, toff.as<uint64_t>() + hcnt[0], 8, hipMemcpyDeviceToHost));
  EH_CHECK(s.tasks.alloc(16 * ntasks + 16));
  k_expand_tasks<<<grid_for(n), 256, 0, st>>>(key.as<uint64_t>(), ntsorted.as<uint32_t>(), toff.as<uint64_t>(),
                                              m->dcard.as<uint32_t>(), jstart.as<uint32_t>(),
                                              rowinfo.as<uint32_t>(), n, s.tasks.as<uint4>());
  EH_HIP(hipGetLastError());
  s.bstat_j = 0;
  s.n_btasks = s.n_in_pairs = 0;
  DevBuf lo16off;  // per row, its 16-bit copy's first vector (the B-stationary and hashed pairs)
  if (bstat_j) {
    // in-lists of the B rows below bstat_j, then tasks of at most BSTAT_CHUNK pairs each
    // (each task clears and fills an 8 KB bitmap: on RMAT-22 256 pairs per task 3.74 ms,
    // 512 3.54, 1024 3.47, 2048 3.55)
    constexpr uint32_t BSTAT_CHUNK = 1024;
    DevBuf inoff, inend, incount, rowpos, ntb, boff;
    EH_CHECK(rowpos.alloc(8ull * n + 16, &ctx->cache));
    k_zero_u32<<<1, 64, 0, st>>>(bpairs.as<uint32_t>() + n, 1);
    EH_CHECK(prim_exscan_u32_to_u64(ctx, bpairs.as<uint32_t>(), rowpos.as<uint64_t>(), (uint64_t)n + 1));
    uint64_t npairs = 0;
    EH_HIP(hipMemcpyAsync(&npairs, rowpos.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (npairs >= (1ull << 32)) {
      eh_set_error("%llu B-stationary pairs exceed the 32-bit pair index", (unsigned long long)npairs);
      return EH_ERR_UNSUPPORTED;
    }
    EH_CHECK(s.in_pairs.alloc(8 * npairs + 16));
    DevBuf pkeys, pkeys2;
    EH_CHECK(pkeys.alloc(8 * npairs + 16, &ctx->cache));
    EH_CHECK(pkeys2.alloc(8 * npairs + 16, &ctx->cache));
    EH_CHECK(inoff.alloc(8ull * bstat_j + 16, &ctx->cache));
    EH_CHECK(inend.alloc(4ull * bstat_j + 8, &ctx->cache));
    EH_CHECK(incount.alloc(4ull * bstat_j + 8, &ctx->cache));
    EH_CHECK(ntb.alloc(4ull * bstat_j + 8, &ctx->cache));
    EH_CHECK(boff.alloc(8ull * bstat_j + 16, &ctx->cache));
    k_zero_u32<<<grid_for(bstat_j + 1), 256, 0, st>>>(inend.as<uint32_t>(), bstat_j + 1);
    // 16-bit prefixes
    DevBuf nvec;
    EH_CHECK(nvec.alloc(4ull * n + 8, &ctx->cache));
    EH_CHECK(lo16off.alloc(8ull * n + 16, &ctx->cache));
    {
      DevBuf nvec2, obig, osmall;
      EH_CHECK(nvec2.alloc(4ull * n + 8, &ctx->cache));
      EH_CHECK(obig.alloc(8ull * n + 16, &ctx->cache));
      EH_CHECK(osmall.alloc(8ull * n + 16, &ctx->cache));
      k_lo16_count<<<grid_for(n), 256, 0, st>>>(jbelow.as<uint32_t>(), n, nvec.as<uint32_t>(), nvec2.as<uint32_t>());
      k_zero_u32<<<1, 64, 0, st>>>(nvec.as<uint32_t>() + n, 1);
      k_zero_u32<<<1, 64, 0, st>>>(nvec2.as<uint32_t>() + n, 1);
      EH_CHECK(prim_exscan_u32_to_u64(ctx, nvec.as<uint32_t>(), obig.as<uint64_t>(), (uint64_t)n + 1));
      EH_CHECK(prim_exscan_u32_to_u64(ctx, nvec2.as<uint32_t>(), osmall.as<uint64_t>(), (uint64_t)n + 1));
      k_lo16_place<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(jbelow.as<uint32_t>(), n, obig.as<uint64_t>(),
                                                            osmall.as<uint64_t>(), lo16off.as<uint64_t>());
      EH_HIP(hipStreamSynchronize(st));
    }
    uint64_t nv16 = 0;
    EH_HIP(hipMemcpyAsync(&nv16, lo16off.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (nv16 >= (1ull << 32)) {
      eh_set_error("B-stationary prefixes exceed the 32-bit vector index");
      return EH_ERR_UNSUPPORTED;
    }
    EH_CHECK(s.lo16.alloc(16 * nv16 + 16));
    k_lo16_fill<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(),
                                                          jbelow.as<uint32_t>(), n, lo16off.as<uint64_t>(),
                                                          bstat_j - 1u, s.lo16.as<uint16_t>());
    k_bstat_fill<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(),
                                                           m->ddesc.as<RowDesc>(), jbelow.as<uint32_t>(), n,
                                                           rowpos.as<uint64_t>(), lo16off.as<uint64_t>(),
                                                           pkeys.as<uint64_t>());
    EH_CHECK(prim_sort_u64_range(ctx, pkeys.as<uint64_t>(), pkeys2.as<uint64_t>(), npairs, 32, 48));
    k_bstat_unpack<<<grid_for(npairs), 256, 0, st>>>(pkeys.as<uint64_t>(), npairs, s.in_pairs.as<uint2>());
    if (npairs)
      k_bstat_bounds<<<grid_for(npairs), 256, 0, st>>>(pkeys.as<uint64_t>(), npairs, inoff.as<uint64_t>(),
                                                       inend.as<uint32_t>());
    k_bstat_counts<<<grid_for(bstat_j + 1), 256, 0, st>>>(inoff.as<uint64_t>(), inend.as<uint32_t>(), bstat_j + 1,
                                                          incount.as<uint32_t>());
    EH_HIP(hipStreamSynchronize(st));
    pkeys.release();
    pkeys2.release();
    k_bstat_ntask<<<grid_for(bstat_j + 1), 256, 0, st>>>(incount.as<uint32_t>(), bstat_j + 1, BSTAT_CHUNK,
                                                        ntb.as<uint32_t>());
    EH_CHECK(prim_exscan_u32_to_u64(ctx, ntb.as<uint32_t>(), boff.as<uint64_t>(), (uint64_t)bstat_j + 1));
    uint64_t nbt = 0;
    EH_HIP(hipMemcpyAsync(&nbt, boff.as<uint64_t>() + bstat_j, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    // the in-list bounds come from sorted keys: every list non-empty has at least one pair per
    // task, so more tasks than pairs would mean unsorted keys -- refuse rather than launch them
    if (nbt > npairs) {
      eh_set_error("B-stationary in-lists inconsistent (%llu tasks for %llu pairs)", (unsigned long long)nbt,
                   (unsigned long long)npairs);
      return EH_ERR_INVALID;
    }
    EH_CHECK(s.btasks.alloc(16 * nbt + 16));
    k_bstat_expand<<<grid_for(bstat_j), 256, 0, st>>>(incount.as<uint32_t>(), inoff.as<uint64_t>(),
                                                     boff.as<uint64_t>(), bstat_j, BSTAT_CHUNK,
                                                     m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                     s.btasks.as<uint4>());
    {
      // tasks in the order of their first pair's prefix offset: the tasks running at one time
      // read the prefixes of nearby rows (with 256-pair tasks the L2 hit rate on RMAT-22 rose
      // from 16% to 38%; with 1024-pair tasks the order is worth 0.5% on RMAT-24 only)
      DevBuf tk, tk2, bt2;
      EH_CHECK(tk.alloc(8 * nbt + 8, &ctx->cache));
      EH_CHECK(tk2.alloc(8 * nbt + 8, &ctx->cache));
      EH_CHECK(bt2.alloc(16 * nbt + 16, &ctx->cache));
      k_bstat_task_keys<<<grid_for(nbt), 256, 0, st>>>(s.btasks.as<uint4>(), nbt, s.in_pairs.as<uint2>(),
                                                       tk.as<uint64_t>());
      EH_CHECK(prim_sort_u64(ctx, tk.as<uint64_t>(), tk2.as<uint64_t>(), nbt, 64));
      k_bstat_task_gather<<<grid_for(nbt), 256, 0, st>>>(s.btasks.as<uint4>(), nbt, tk.as<uint64_t>(),
                                                         bt2.as<uint4>());
      EH_HIP(hipMemcpyAsync(s.btasks.p, bt2.p, 16 * nbt, hipMemcpyDeviceToDevice, st));
      EH_HIP(hipStreamSynchronize(st));
    }
    EH_HIP(hipGetLastError());
    s.bstat_j = bstat_j;
    s.n_btasks = nbt;
    s.n_in_pairs = npairs;
  }
  s.n_htasks = 0;
  s.bsh_longest = 0;
  if (nh) {
    // hashed part: the B rows at or above the split with in-pairs (compacted, row order),
    // bins of them by table words, the in-pairs of each, tasks of at most BSH_CHUNK pairs
    constexpr uint32_t BSH_CHUNK = 256;
    DevBuf hflag, hpos, hc, hw, pw, bflag, bpos, hfill, ntb, boff;
    EH_CHECK(hflag.alloc(4ull * nh + 8, &ctx->cache));
    EH_CHECK(hpos.alloc(8ull * nh + 16, &ctx->cache));
    k_nonzero_flag<<<grid_for(nh), 256, 0, st>>>(hcount.as<uint32_t>(), nh, hflag.as<uint32_t>());
    k_zero_u32<<<1, 64, 0, st>>>(hflag.as<uint32_t>() + nh, 1);
    EH_CHECK(prim_exscan_u32_to_u64(ctx, hflag.as<uint32_t>(), hpos.as<uint64_t>(), (uint64_t)nh + 1));
    uint64_t nhr = 0;
    EH_HIP(hipMemcpyAsync(&nhr, hpos.as<uint64_t>() + nh, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (nhr) {
      EH_CHECK(s.hj.alloc(4 * nhr + 16));
      EH_CHECK(s.hjlo.alloc(4 * nhr + 16));
      EH_CHECK(hc.alloc(4 * nhr + 16, &ctx->cache));
      EH_CHECK(hw.alloc(4 * nhr + 16, &ctx->cache));
      EH_CHECK(s.hoff.alloc(8 * nhr + 16));
      EH_CHECK(pw.alloc(8 * nhr + 16, &ctx->cache));
      EH_CHECK(bflag.alloc(4 * nhr + 16, &ctx->cache));
      EH_CHECK(bpos.alloc(8 * nhr + 16, &ctx->cache));
      EH_CHECK(hfill.alloc(4 * nhr + 16, &ctx->cache));
      k_zero_u32<<<1, 64, 0, st>>>(cnt2.as<uint32_t>(), 1);
      k_bsh_rows<<<grid_for(nh), 256, 0, st>>>(hcount.as<uint32_t>(), hpos.as<uint64_t>(), nh, bstat_j,
                                               m->dcard.as<uint32_t>(), jbelow.as<uint32_t>(), s.hj.as<uint32_t>(),
                                               s.hjlo.as<uint32_t>(), hc.as<uint32_t>(), hw.as<uint32_t>(),
                                               cnt2.as<uint32_t>());
      k_zero_u32<<<1, 64, 0, st>>>(hc.as<uint32_t>() + nhr, 1);
      EH_CHECK(prim_exscan_u32_to_u64(ctx, hc.as<uint32_t>(), s.hoff.as<uint64_t>(), nhr + 1));
      EH_CHECK(prim_exscan_u32_to_u64(ctx, hw.as<uint32_t>(), pw.as<uint64_t>(), nhr));
      k_bsh_bin_flag<<<grid_for(nhr), 256, 0, st>>>(pw.as<uint64_t>(), (uint32_t)nhr, bflag.as<uint32_t>());
      k_zero_u32<<<1, 64, 0, st>>>(bflag.as<uint32_t>() + nhr, 1);
      EH_CHECK(prim_exscan_u32_to_u64(ctx, bflag.as<uint32_t>(), bpos.as<uint64_t>(), nhr + 1));
      uint64_t nb = 0, nhp = 0;
      uint32_t longest = 0;
      EH_HIP(hipMemcpyAsync(&nb, bpos.as<uint64_t>() + nhr, 8, hipMemcpyDeviceToHost, st));
      EH_HIP(hipMemcpyAsync(&nhp, s.hoff.as<uint64_t>() + nhr, 8, hipMemcpyDeviceToHost, st));
      EH_HIP(hipMemcpyAsync(&longest, cnt2.p, 4, hipMemcpyDeviceToHost, st));
      EH_HIP(hipStreamSynchronize(st));
      if (longest > BSH_MAX_D) {
        eh_set_error("a B row of %u ids above the split exceeds the %u the B-stationary pass holds", longest,
                     BSH_MAX_D);
        return EH_ERR_INVALID;
      }
      s.bsh_longest = longest;
      if (nhp >= (1ull << 32)) {
        eh_set_error("%llu B-stationary pairs above the split exceed the 32-bit pair index", (unsigned long long)nhp);
        return EH_ERR_UNSUPPORTED;
      }
      EH_CHECK(s.bstart.alloc(4 * nb + 16));
      k_bsh_bin_start<<<grid_for(nhr), 256, 0, st>>>(bflag.as<uint32_t>(), bpos.as<uint64_t>(), (uint32_t)nhr,
                                                     s.bstart.as<uint32_t>());
      const uint32_t last = (uint32_t)nhr;
      EH_HIP(hipStreamSynchronize(st));
      EH_HIP(hipMemcpy(s.bstart.as<uint32_t>() + nb, &last, 4, hipMemcpyHostToDevice));
      EH_CHECK(s.hpairs.alloc(16 * nhp + 16));
      k_zero_u32<<<grid_for(nhr + 1), 256, 0, st>>>(hfill.as<uint32_t>(), (uint32_t)(nhr + 1));
      k_bsh_fill<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(
          m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(), m->ddesc.as<RowDesc>(),
          jstart.as<uint32_t>(), jbelow.as<uint32_t>(), n, bstat_j, hpos.as<uint64_t>(), s.hoff.as<uint64_t>(),
          hfill.as<uint32_t>(), lo16off.as<uint64_t>(), s.hpairs.as<uint4>());
      EH_CHECK(ntb.alloc(4 * nb + 16, &ctx->cache));
      EH_CHECK(boff.alloc(8 * nb + 16, &ctx->cache));
      k_bsh_bin_ntask<<<grid_for(nb), 256, 0, st>>>(s.bstart.as<uint32_t>(), s.hoff.as<uint64_t>(), (uint32_t)nb,
                                                    BSH_CHUNK, ntb.as<uint32_t>());
      k_zero_u32<<<1, 64, 0, st>>>(ntb.as<uint32_t>() + nb, 1);
      EH_CHECK(prim_exscan_u32_to_u64(ctx, ntb.as<uint32_t>(), boff.as<uint64_t>(), nb + 1));
      uint64_t nht = 0;
      EH_HIP(hipMemcpyAsync(&nht, boff.as<uint64_t>() + nb, 8, hipMemcpyDeviceToHost, st));
      EH_HIP(hipStreamSynchronize(st));
      EH_CHECK(s.htasks.alloc(16 * nht + 16));
      k_bsh_bin_expand<<<grid_for(nb), 256, 0, st>>>(s.bstart.as<uint32_t>(), s.hoff.as<uint64_t>(),
                                                     boff.as<uint64_t>(), (uint32_t)nb, BSH_CHUNK,
                                                     s.htasks.as<uint4>());
      EH_HIP(hipGetLastError());
      s.n_htasks = nht;
    }
  }
  EH_HIP(hipEventRecord(ctx->ev1, st));
  EH_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  EH_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  s.ms = ms;
  s.n_big = nbig;
  s.n_small = ntasks - nbig;
  s.valid = true;
  s.sched = sched;
  s.task_work = tp.task_work;
  s.lds_small = tp.lds_small;
  s.lds_big = tp.lds_big;
  for (int q = 0; q < 6; q++) s.stats[q] = hs.v[q];
  return EH_OK;
}

namespace {

// blocks of `kern` one CU holds at this shape; the query is cached per (kernel, shape)
// because per-graph batch launches would otherwise repeat it thousands of times
int blocks_per_cu(const void *kern, int threads, size_t lds_bytes) {
  struct Entry {
    const void *k;
    int threads;
    size_t lds;
    int per_cu;
  };
  static std::mutex mu;
  static std::vector<Entry> cache;
  std::lock_guard<std::mutex> lock(mu);
  for (const Entry &e : cache)
    if (e.k == kern && e.threads == threads && e.lds == lds_bytes) return e.per_cu;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds_bytes) != hipSuccess || per_cu < 1)
    per_cu = 1;
  cache.push_back({kern, threads, lds_bytes, per_cu});
  return per_cu;
}

// device-time marks around one kernel class of a pass (eh_run_opts.time_parts)
inline void part_begin(eh_ctx *ctx, PartEv *pe) {
  if (pe) (void)hipEventRecord(pe->b, ctx->stream);
}
inline void part_end(eh_ctx *ctx, PartEv *pe) {
  if (pe) {
    (void)hipEventRecord(pe->e, ctx->stream);
    pe->used = true;
  }
}

template <int K, bool W>
eh_status launch_count(eh_ctx *ctx, const CountArgs &base, const uint4 *tasks, uint64_t ntasks,
                       uint32_t lds_words, int waves, uint32_t *d_next, uint64_t task_base, PartEv *pe) {
  auto kern = !base.tuned ? k_count<K, W, false, false>
                          : base.blocks ? k_count<K, W, true, true> : k_count<K, W, true, false>;
  if (ntasks == 0) return EH_OK;
  CountArgs a = base;
  a.task_base = task_base;
  a.tasks = tasks;
  a.ntasks = ntasks;
  a.next = d_next;
  a.lds_words = lds_words;
  const size_t lds_bytes = (size_t)lds_words * 4 * waves;
  const int threads = 64 * waves;
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(kern), threads, lds_bytes);
  uint64_t blocks = (uint64_t)per_cu * (ctx->num_cus ? ctx->num_cus : 256);
  const uint64_t need = ceil_div(ntasks, (uint64_t)waves);
  if (blocks > need) blocks = need;
  k_zero_u32<<<1, 64, 0, ctx->stream>>>(d_next, 1);
  part_begin(ctx, pe);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(threads), lds_bytes, ctx->stream, a);
  EH_HIP(hipGetLastError());
  part_end(ctx, pe);
  return EH_OK;
}

template <int K, bool W>
eh_status launch_small(eh_ctx *ctx, const CountArgs &base, const uint32_t *rows, uint64_t nrows, uint32_t *d_next,
                       uint32_t *xq, PartEv *pe) {
  if (nrows == 0) return EH_OK;
  CountArgs a = base;
  a.next = d_next;
  a.xq = xq;
  const int waves = 4, threads = 256;
  const size_t lds_bytes =
      (size_t)(K == 4 ? SMALL4_META_WORDS + SMALL_ADJ_WORDS : SMALL_META_WORDS) * 4 * waves;
  auto kern = (K == 4) ? k_small_cliques4 : k_small_rows<3, W>;
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(kern), threads, lds_bytes);
  uint64_t blocks = (uint64_t)per_cu * (ctx->num_cus ? ctx->num_cus : 256);
  const uint64_t need = ceil_div(ceil_div(nrows, 64), (uint64_t)waves);
  if (blocks > need) blocks = need;
  if (xq) k_zero_u32<<<1, 256, 0, ctx->stream>>>(xq + XQ_HEADS, 8 * XQ_STRIDE);
  else k_zero_u32<<<1, 64, 0, ctx->stream>>>(d_next, 1);
  part_begin(ctx, pe);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(threads), lds_bytes, ctx->stream, a, rows, (uint32_t)nrows);
  EH_HIP(hipGetLastError());
  part_end(ctx, pe);
  return EH_OK;
}

eh_status launch_bstat(eh_ctx *ctx, const CountArgs &base, const uint4 *btasks, uint64_t nbt,
                       const uint2 *in_pairs, uint32_t bstat_j, uint32_t *d_next, PartEv *pe) {
  if (nbt == 0) return EH_OK;
  CountArgs a = base;
  a.next = d_next;
  // a wave's slice: the pass metadata and a bitmap over the ids [0, bstat_j)
  const uint32_t bm_words = 2u * (((bstat_j - 1u) >> 6) + 1u);
  a.lds_words = META_WORDS + bm_words;
  int waves = 4;
  while (waves > 1 && (size_t)a.lds_words * 4 * waves > 64 * 1024) waves--;
  const size_t lds_bytes = (size_t)a.lds_words * 4 * waves;
  const int threads = 64 * waves;
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(k_bstat), threads, lds_bytes);
  uint64_t blocks = (uint64_t)per_cu * (ctx->num_cus ? ctx->num_cus : 256);
  const uint64_t need = ceil_div(ceil_div(nbt, (uint64_t)a.grab), (uint64_t)waves);
  if (blocks > need) blocks = need;
  k_zero_u32<<<1, 64, 0, ctx->stream>>>(d_next, 1);
  part_begin(ctx, pe);
  hipLaunchKernelGGL(k_bstat, dim3((unsigned)blocks), dim3(threads), lds_bytes, ctx->stream, a, btasks, nbt,
                     in_pairs, bm_words);
  EH_HIP(hipGetLastError());
  part_end(ctx, pe);
  return EH_OK;
}

template <uint32_t LMAX, bool GROUP_BLOCK>
eh_status launch_local(eh_ctx *ctx, const CountArgs &base, const uint32_t *rows, uint64_t nrows, uint32_t *d_next,
                       PartEv *pe) {
  if (nrows == 0) return EH_OK;
  CountArgs a = base;
  a.next = d_next;
  const int waves = 4, threads = 256;
  const size_t lds_bytes = (size_t)local4_slice<LMAX>(GROUP_BLOCK) * 4 * (GROUP_BLOCK ? 1 : waves);
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(k_local4<LMAX, GROUP_BLOCK>), threads, lds_bytes);
  uint64_t blocks = (uint64_t)per_cu * (ctx->num_cus ? ctx->num_cus : 256);
  const uint64_t need = GROUP_BLOCK ? nrows : ceil_div(nrows, (uint64_t)waves);
  if (blocks > need) blocks = need;
  k_zero_u32<<<1, 64, 0, ctx->stream>>>(d_next, 1);
  part_begin(ctx, pe);
  hipLaunchKernelGGL((k_local4<LMAX, GROUP_BLOCK>), dim3((unsigned)blocks), dim3(threads), lds_bytes, ctx->stream, a,
                     rows, (uint32_t)nrows);
  EH_HIP(hipGetLastError());
  part_end(ctx, pe);
  return EH_OK;
}

eh_status launch_bshl(eh_ctx *ctx, const CountArgs &base, const uint4 *htasks, uint64_t nht, const uint4 *hpairs,
                      const uint32_t *hj, const uint32_t *hjlo, const uint64_t *hoff, const uint32_t *bstart,
                      uint32_t bstat_j, uint32_t longest, uint32_t *d_next, PartEv *pe) {
  if (nht == 0) return EH_OK;
  CountArgs a = base;
  a.next = d_next;
  // row(j) in 4 register rounds (256 ids) unless a B row is longer, then 8 (512)
  const bool wide = longest > 256u;
  auto kern = wide ? k_bshl<8> : k_bshl<4>;
  // a wave's slice: the pass metadata, a bitmap over the ids [0, bstat_j) and a B row's ids
  // at or above the split
  const uint32_t bm_words = 2u * (((bstat_j - 1u) >> 6) + 1u);
  a.lds_words = META_WORDS + bm_words + (wide ? 512u : 256u);
  // two waves per block: the 10.75 KB slices then fill 14 waves per CU (four per block fit 12;
  // RMAT-22 0.93 against 1.01 ms)
  int waves = 2;
  while (waves > 1 && (size_t)a.lds_words * 4 * waves > 64 * 1024) waves--;
  const size_t lds_bytes = (size_t)a.lds_words * 4 * waves;
  const int threads = 64 * waves;
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(kern), threads, lds_bytes);
  uint64_t blocks = (uint64_t)per_cu * (ctx->num_cus ? ctx->num_cus : 256);
  const uint64_t need = ceil_div(ceil_div(nht, (uint64_t)a.grab), (uint64_t)waves);
  if (blocks > need) blocks = need;
  k_zero_u32<<<1, 64, 0, ctx->stream>>>(d_next, 1);
  part_begin(ctx, pe);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(threads), lds_bytes, ctx->stream, a, htasks, nht, hpairs,
                     hj, hjlo, hoff, bstart, bm_words, bstat_j - 1u);
  EH_HIP(hipGetLastError());
  part_end(ctx, pe);
  return EH_OK;
}

}  // namespace

eh_status launch_parts(eh_matrix *m, int k, int sched, const TuneParams &tp, const uint32_t *small_rows,
                       uint64_t n_small_rows, const uint4 *tasks_big, uint64_t n_big, const uint4 *tasks_small,
                       uint64_t n_small, uint32_t lds_small, uint32_t lds_big, uint64_t *d_counts, uint32_t *d_next,
                       uint64_t *d_task_cycles, const uint4 *btasks, uint64_t n_btasks, const uint2 *in_pairs,
                       uint32_t bstat_j, const uint32_t *jstart, const uint4 *lo16, const uint32_t *local_rows0,
                       uint64_t n_local0, const uint32_t *local_rows1, uint64_t n_local1, const Schedule *hs,
                       PartEv *pev, uint64_t *d_counts3) {
  // pev (may be null): one pair of events per kernel class, indexed by EH_PART_*
  auto pe = [pev](int q) -> PartEv * { return pev ? pev + q : nullptr; };
  if (d_counts3 && (k != 4 || n_big || n_small || n_local0 || n_local1)) {
    eh_set_error("the triangle pass is fused into a 4-clique pass of short rows only");
    return EH_ERR_INVALID;
  }
  CountArgs a;
  a.counts3 = reinterpret_cast<unsigned long long *>(d_counts3);
  a.grab = tp.grab ? tp.grab : 1u;
  // d_next[7] is the pass's error word (the queue heads use d_next[0..4]); the caller
  // clears it before the pass and reads it after
  a.err = d_next + 7;
  a.xq = nullptr;
  // a batch arena's short rows from its per-XCD queues (the schedule's own list only)
  uint32_t *xq = (hs && tp.short_queues && hs->xq.p && small_rows == hs->small_rows.as<uint32_t>())
                     ? hs->xq.as<uint32_t>() : nullptr;
  a.jstart = jstart;
  a.lo16 = lo16;
  a.task_cycles = reinterpret_cast<unsigned long long *>(d_task_cycles);
  a.task_base = 0;
  a.card = m->dcard.as<uint32_t>();
  a.aoff = m->daoff.as<uint32_t>();
  a.unbr = m->dunbr.as<uint32_t>();
  a.desc = m->ddesc.as<RowDesc>();
  a.bsw = m->dbsw.as<uint64_t>();
  a.psd = m->dpsd.as<uint16_t>();
  a.bid = m->dbid.as<uint32_t>();
  a.bwd = m->dbwd.as<uint4>();
  a.ntu = m->dntu.as<uint32_t>();
  a.blocks = m->blk_slots ? 1u : 0u;
  a.counts = reinterpret_cast<unsigned long long *>(d_counts);
  a.tuned = sched == EH_SCHED_TUNED ? 1u : 0u;
  a.graph_of = m->dgraph_of.p ? m->dgraph_of.as<uint32_t>() : nullptr;
  a.stream_ratio = tp.stream_ratio;
  const bool weighted = k == -3;
  if ((n_btasks || (hs && hs->n_htasks)) && k != 3) {
    eh_set_error("B-stationary schedules serve the plain triangle count only");
    return EH_ERR_INVALID;
  }
  if (k == 4) {
    EH_CHECK((launch_small<4, false>(m->ctx, a, small_rows, n_small_rows, d_next + 2, xq, pe(EH_PART_SHORT_ROWS))));
    EH_CHECK((launch_local<LOCAL4_MID, false>(m->ctx, a, local_rows0, n_local0, d_next + 3, pe(EH_PART_LOCAL_MID))));
    EH_CHECK((launch_local<LOCAL4_MAX, true>(m->ctx, a, local_rows1, n_local1, d_next + 4, pe(EH_PART_LOCAL_BIG))));
    EH_CHECK((launch_count<4, false>(m->ctx, a, tasks_big, n_big, lds_big, tp.big_waves, d_next, 0,
                                     pe(EH_PART_TASKS_BIG))));
    EH_CHECK((launch_count<4, false>(m->ctx, a, tasks_small, n_small, lds_small, tp.waves_per_block, d_next + 1,
                                     n_big, pe(EH_PART_TASKS))));
  } else if (weighted) {
    EH_CHECK((launch_small<3, true>(m->ctx, a, small_rows, n_small_rows, d_next + 2, xq, pe(EH_PART_SHORT_ROWS))));
    EH_CHECK((launch_count<3, true>(m->ctx, a, tasks_big, n_big, lds_big, tp.big_waves, d_next, 0,
                                    pe(EH_PART_TASKS_BIG))));
    EH_CHECK((launch_count<3, true>(m->ctx, a, tasks_small, n_small, lds_small, tp.waves_per_block, d_next + 1,
                                    n_big, pe(EH_PART_TASKS))));
  } else {
    EH_CHECK((launch_small<3, false>(m->ctx, a, small_rows, n_small_rows, d_next + 2, xq, pe(EH_PART_SHORT_ROWS))));
    EH_CHECK(launch_bstat(m->ctx, a, btasks, n_btasks, in_pairs, bstat_j, d_next + 3, pe(EH_PART_BSTAT)));
    if (hs && hs->n_htasks) {
      EH_CHECK(launch_bshl(m->ctx, a, hs->htasks.as<uint4>(), hs->n_htasks, hs->hpairs.as<uint4>(),
                           hs->hj.as<uint32_t>(), hs->hjlo.as<uint32_t>(), hs->hoff.as<uint64_t>(),
                           hs->bstart.as<uint32_t>(), bstat_j, hs->bsh_longest, d_next + 5,
                           pe(EH_PART_BSTAT_ABOVE)));
    }
    EH_CHECK((launch_count<3, false>(m->ctx, a, tasks_big, n_big, lds_big, tp.big_waves, d_next, 0,
                                     pe(EH_PART_TASKS_BIG))));
    EH_CHECK((launch_count<3, false>(m->ctx, a, tasks_small, n_small, lds_small, tp.waves_per_block, d_next + 1,
                                     n_big, pe(EH_PART_TASKS))));
  }
  return EH_OK;
}

eh_status schedule_launch(eh_matrix *m, int k, int sched, const TuneParams &tp, Schedule &s,
                          uint64_t *d_counts, uint32_t *d_next, uint64_t *d_task_cycles, PartEv *pev,
                          uint64_t *d_counts3) {
  const uint4 *t = s.tasks.as<uint4>();
  return launch_parts(m, k, sched, tp, s.small_rows.as<uint32_t>(), s.n_small_rows, t, s.n_big, t + s.n_big,
                      s.n_small, s.lds_small, s.lds_big, d_counts, d_next, d_task_cycles, s.btasks.as<uint4>(),
                      s.n_btasks, s.in_pairs.as<uint2>(), s.bstat_j,
                      s.bstat_j ? s.jstart.as<uint32_t>() : nullptr, s.bstat_j ? s.lo16.as<uint4>() : nullptr,
                      s.local_rows[0].as<uint32_t>(), s.n_local[0], s.local_rows[1].as<uint32_t>(), s.n_local[1],
                      &s, pev, d_counts3);
}

// Sum over the kept pairs (i, j) of sum over x in row(i) & row(j) of card(x): over the
// triangles i > j > x of an oriented matrix, the card(x) term of the 4-clique byte model
// (the weighted triangle pass, without the B-stationary kernels, which count only).
// Computed once per matrix (summed over a batch arena's graphs, `ncounts` of them).
eh_status matrix_triangle_weight(eh_matrix *m, uint32_t ncounts, uint64_t *out) {
  if (m->tw_valid) { *out = m->tri_weight; return EH_OK; }
  eh_ctx *ctx = m->ctx;
  hipStream_t st = ctx->stream;
  eh_run_opts o = {};
  o.schedule = EH_SCHED_TUNED;
  o.bstat_split = -1;
  const TuneParams tp = tune_defaults(m, EH_SCHED_TUNED, 3, &o);
  Schedule s;
  EH_CHECK(schedule_build(m, EH_SCHED_TUNED, 3, tp, s));
  if (ncounts == 0) ncounts = 1;
  DevBuf cnt, next;
  EH_CHECK(cnt.alloc(8ull * ncounts));
  EH_CHECK(next.alloc(32));
  k_zero_u64<<<grid_for(ncounts), 256, 0, st>>>(cnt.as<unsigned long long>(), ncounts);
  k_zero_u32<<<1, 64, 0, st>>>(next.as<uint32_t>() + 7, 1);
  EH_CHECK(schedule_launch(m, -3, EH_SCHED_TUNED, tp, s, cnt.as<uint64_t>(), next.as<uint32_t>(), nullptr, nullptr));
  uint32_t dev_err = 0;
  std::vector<uint64_t> h(ncounts);
  EH_HIP(hipMemcpy(h.data(), cnt.p, 8ull * ncounts, hipMemcpyDeviceToHost));
  EH_HIP(hipMemcpy(&dev_err, next.as<uint32_t>() + 7, 4, hipMemcpyDeviceToHost));
  if (dev_err) { eh_set_error("weighted triangle pass: an LDS hash table had no free slot"); return EH_ERR_INVALID; }
  uint64_t sum = 0;
  for (uint64_t v : h) sum += v;
  m->tri_weight = sum;
  m->tw_valid = true;
  *out = sum;
  return EH_OK;
}

// k: 3 triangles, 4 four-cliques, -3 weighted triangle pass (sum of card(x) over
// common x; the per-triangle byte term of the 4-clique model)
eh_status schedule_ensure(eh_matrix *m, int k, const eh_run_opts *opts, double *sched_ms) {
  const int sched = opts ? opts->schedule : EH_SCHED_TUNED;
  if (sched != EH_SCHED_PLAIN && sched != EH_SCHED_TUNED) { eh_set_error("bad schedule %d", sched); return EH_ERR_INVALID; }
  const int kk = (k == 4) ? 4 : 3;
  const TuneParams tp = tune_defaults(m, sched, kk, opts);
  Schedule &s = m->sched[kk == 4 ? 1 : 0][sched];
  const int skey = kk * 8 + sched;
  if (sched_ms) *sched_ms = 0;
  if (!s.valid || s.task_work != tp.task_work || s.lds_small != tp.lds_small || s.lds_big != tp.lds_big ||
      (int)s.sched != skey || s.small_deg != tp.small_deg || s.bstat_req != tp.bstat_j || s.bsh_req != tp.bsh_max ||
      s.local_req != tp.local_max) {
    s = Schedule();
    EH_CHECK(schedule_build(m, sched, kk, tp, s));
    s.sched = skey;
    s.bstat_req = tp.bstat_j;
    s.bsh_req = tp.bsh_max;
    s.local_req = tp.local_max;
    if (sched_ms) *sched_ms = s.ms;
  }
  return EH_OK;
}

eh_status count_run(eh_matrix *m, int k, const eh_run_opts *opts, uint64_t *d_counts, uint32_t ncounts,
                    uint64_t *count, eh_run_stats *stats, uint64_t *d_task_cycles) {
  eh_ctx *ctx = m->ctx;
  hipStream_t st = ctx->stream;
  const int sched = opts ? opts->schedule : EH_SCHED_TUNED;
  const int kk = (k == 4) ? 4 : 3;
  double sched_ms = 0;
  EH_CHECK(schedule_ensure(m, k, opts, &sched_ms));
  const TuneParams tp = tune_defaults(m, sched, kk, opts);
  Schedule &s = m->sched[kk == 4 ? 1 : 0][sched];
  // the pass's queue heads and error word (32 words) and, when the caller passes no device
  // counters, its counts: the context's buffer, grown when needed
  const uint64_t own = d_counts ? 0ull : (ncounts ? ncounts : 1);
  DevBuf &pw = ctx->pass_words;
  if (pw.bytes < 128 + 8 * own) EH_CHECK(pw.alloc(128 + 8 * own + 1024));
  uint32_t *const nx = pw.as<uint32_t>();
  unsigned long long *dc = d_counts ? reinterpret_cast<unsigned long long *>(d_counts)
                                    : reinterpret_cast<unsigned long long *>(pw.as<uint8_t>() + 128);
  const int repeat = (opts && opts->repeat > 1) ? opts->repeat : 1;
  PartEv *pev = nullptr;
  if (opts && opts->time_parts) {
    EH_CHECK(ctx_part_events(ctx));
    pev = ctx->parts;
  }
  double part_ms[EH_PARTS] = {};
  float total_ms = 0;
  for (int r = 0; r < repeat; r++) {
    k_zero_u64<<<grid_for(ncounts ? ncounts : 1), 256, 0, st>>>(dc, ncounts ? ncounts : 1);
    k_zero_u32<<<1, 64, 0, st>>>(nx + 7, 1);  // the pass's error word
    if (pev)
      for (int q = 0; q < EH_PARTS; q++) pev[q].used = false;
    EH_HIP(hipEventRecord(ctx->ev0, st));
    EH_CHECK(schedule_launch(m, k, sched, tp, s, reinterpret_cast<uint64_t *>(dc), nx,
                             d_task_cycles, pev));
    EH_HIP(hipEventRecord(ctx->ev1, st));
    EH_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    EH_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    total_ms += ms;
    if (pev)
      for (int q = 0; q < EH_PARTS; q++)
        if (pev[q].used) {
          float pm = 0;
          EH_HIP(hipEventElapsedTime(&pm, pev[q].b, pev[q].e));
          part_ms[q] += pm;
        }
  }
  uint32_t dev_err = 0;
  std::vector<uint64_t> h(ncounts ? ncounts : 1);
  if (!d_counts && count) {
    // the error word and the counts sit in one buffer: one copy brings both
    std::vector<uint64_t> words(16 + h.size());
    EH_HIP(hipMemcpy(words.data(), pw.p, 8 * words.size(), hipMemcpyDeviceToHost));
    dev_err = reinterpret_cast<const uint32_t *>(words.data())[7];
    std::copy(words.begin() + 16, words.end(), h.begin());
  } else {
    EH_HIP(hipMemcpy(&dev_err, nx + 7, 4, hipMemcpyDeviceToHost));
    if (count) EH_HIP(hipMemcpy(h.data(), dc, 8 * h.size(), hipMemcpyDeviceToHost));
  }
  if (dev_err) {
    eh_set_error("counting pass: an LDS hash table had no free slot (device error word %#x); the count is "
                 "not valid", dev_err);
    return EH_ERR_INVALID;
  }
  if (count)
    for (size_t q = 0; q < h.size(); q++) count[q] = h[q];
  if (stats) {
    stats->kernel_ms = total_ms / repeat;
    stats->sched_ms = sched_ms;
    stats->tasks = s.n_big + s.n_small;
    stats->alg_bytes = 4ull * (m->card + m->wedges) + 8ull * ((uint64_t)m->n + 1);
    if (k == 4) {
      uint64_t tw = 0;
      EH_CHECK(matrix_triangle_weight(m, ncounts, &tw));
      stats->alg_bytes += 4ull * tw;
    }
    for (int q = 0; q < EH_PARTS; q++) stats->part_ms[q] = part_ms[q] / repeat;
    for (int q = 0; q < 4; q++) stats->pairs[q] = s.stats[q];
    stats->work_units = s.stats[4];
    stats->schedule = sched;
    stats->lanes_per_row = tp.lanes;
    stats->task_work = tp.task_work;
    stats->lds_words = tp.lds_small;
  }
  return EH_OK;
}

extern "C" eh_status eh_triangle_count(eh_matrix *m, const eh_run_opts *opts, uint64_t *count,
                                       eh_run_stats *stats) {
  if (!m || !count) { eh_set_error("eh_triangle_count: null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(m->ctx->device));
  if (stats) *stats = eh_run_stats{};
  return count_run(m, 3, opts, nullptr, 1, count, stats, nullptr);
}

extern "C" eh_status eh_clique_count(eh_matrix *m, uint32_t k, const eh_run_opts *opts, uint64_t *count,
                                     eh_run_stats *stats) {
  if (!m || !count) { eh_set_error("eh_clique_count: null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(m->ctx->device));
  if (stats) *stats = eh_run_stats{};
  if (k == 3 || k == 4) return count_run(m, (int)k, opts, nullptr, 1, count, stats, nullptr);
  // deeper cliques: the general level-set recursion
  double ms = 0;
  EH_CHECK(clique_count_deep(m, k, count, &ms));
  if (stats) {
    stats->kernel_ms = ms;
    stats->schedule = opts ? opts->schedule : EH_SCHED_TUNED;
  }
  return EH_OK;
}

extern "C" eh_status eh_count_profile(eh_matrix *m, uint32_t k, const eh_run_opts *opts, uint32_t *task_rows,
                                      uint32_t *task_kinds, uint64_t *task_cycles, uint64_t cap, uint64_t *ntasks) {
  if (!m || !ntasks) { eh_set_error("eh_count_profile: null argument"); return EH_ERR_INVALID; }
  if (k != 3 && k != 4) { eh_set_error("eh_count_profile: k=%u not supported (3 or 4)", k); return EH_ERR_UNSUPPORTED; }
  EH_HIP(hipSetDevice(m->ctx->device));
  const int sched = opts ? opts->schedule : EH_SCHED_TUNED;
  if (sched != EH_SCHED_PLAIN && sched != EH_SCHED_TUNED) { eh_set_error("bad schedule %d", sched); return EH_ERR_INVALID; }
  uint64_t count = 0;
  // first pass builds (or reuses) the schedule; the second records cycles
  EH_CHECK(count_run(m, (int)k, opts, nullptr, 1, &count, nullptr, nullptr));
  Schedule &s = m->sched[k == 4 ? 1 : 0][sched];
  const uint64_t nt = s.n_big + s.n_small;
  *ntasks = nt;
  if (cap < nt) { eh_set_error("eh_count_profile needs room for %llu tasks", (unsigned long long)nt); return EH_ERR_CAPACITY; }
  DevBuf cyc;
  EH_CHECK(cyc.alloc(8 * nt + 8));
  EH_CHECK(count_run(m, (int)k, opts, nullptr, 1, &count, nullptr, cyc.as<uint64_t>()));
  std::vector<uint4> t(nt);
  if (nt) EH_HIP(hipMemcpy(t.data(), s.tasks.p, 16 * nt, hipMemcpyDeviceToHost));
  if (task_cycles && nt) EH_HIP(hipMemcpy(task_cycles, cyc.p, 8 * nt, hipMemcpyDeviceToHost));
  for (uint64_t q = 0; q < nt; q++) {
    if (task_rows) task_rows[q] = t[q].x;
    if (task_kinds) task_kinds[q] = t[q].w >> KIND_SHIFT;
  }
  return EH_OK;
}
