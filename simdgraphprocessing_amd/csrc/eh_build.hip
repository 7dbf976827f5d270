// eh_build.hip -- device layout builder (SparseMatrix::build on the GPU).
//
// Restates SparseMatrix<T,R>::build -> from_symmetric_noattribute_graph -> pack_data
// (src/SparseMatrix.hpp:249-266,268-290,406-494) with the per-row layout choice of
// hybrid::get_type (src/set/layouts/hybrid.hpp:73-100) and the layout encoders
// uinteger/pshort/bitset (uinteger.hpp:49-59, pshort.hpp:46-91, bitset.hpp:65-109)
// and the block layouts bitset_new/new_type (bitset_new.hpp:66-118,
// new_type.hpp:50-101).  Passes, all on the device:
//   1. filter count   card[i] = kept entries of row i (all / nbr < node / mask)
//   2. pack           rows copied into one 16-byte-aligned uint32 arena
//   3. select         per-row type from (card, first, last, prefix partitions, blocks)
//   4. encode         bitset words, pshort streams and blocks for the rows that use them
//   5. describe       one 16-byte RowDesc per row for the counting kernels
#include <algorithm>

#include "eh_internal.hpp"

namespace {

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = ceil_div(work ? work : 1, block);
  return (unsigned)(g > 65536 ? 65536 : g);
}

__global__ void k_filter_count(const uint64_t *off, const uint32_t *nbr, uint32_t n, int filter,
                               const uint8_t *mask, uint32_t *card) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = off[i], e = off[i + 1];
    uint64_t c;
    if (filter == EH_FILTER_ALL) {
      c = e - b;
    } else if (filter == EH_FILTER_LOWER) {
      // rows are sorted: the kept entries are the prefix below i
      uint64_t lo = b, hi = e;
      while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (nbr[mid] < (uint32_t)i) lo = mid + 1; else hi = mid;
      }
      c = lo - b;
    } else {
      c = 0;
      for (uint64_t k = b; k < e; k++) c += mask[k] != 0;
    }
    card[i] = (uint32_t)c;
  }
}

__global__ void k_pad4(const uint32_t *card, uint32_t n, uint32_t *pc) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x)
    pc[i] = i < n ? ((card[i] + 3u) & ~3u) : 0u;
}

__global__ void k_aoff(const uint64_t *poff, uint32_t n, uint32_t *aoff) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    aoff[i] = (uint32_t)(poff[i] >> 2);
}

// copy the kept entries of every row, pad with the sentinel.  A wave takes 64 consecutive
// rows: a lane copies a row of at most PACK_LANE_MAX entries itself (most rows), then the wave
// copies the longer ones together; a masked row is always copied by the wave (its kept entries
// are found with ballots).
constexpr uint32_t PACK_LANE_MAX = 32;

__global__ void __launch_bounds__(256) k_pack(const uint64_t *off, const uint32_t *nbr, uint32_t n, int filter,
                                              const uint8_t *mask, const uint32_t *card, const uint64_t *poff,
                                              uint32_t *unbr) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = wid * 64; base < n; base += nwv * 64) {
    const uint64_t i = base + lane;
    const bool valid = i < n;
    const uint32_t c = valid ? card[i] : 0u;
    const bool own = valid && filter != EH_FILTER_MASK && c <= PACK_LANE_MAX;
    if (own) {
      const uint32_t *src = nbr + off[i];
      uint32_t *dst = unbr + poff[i];
      for (uint32_t k = 0; k < c; k++) dst[k] = src[k];
      for (uint32_t k = c; k < ((c + 3u) & ~3u); k++) dst[k] = EH_PAD_VALUE;
    }
    uint64_t lm = __ballot(valid && !own);
    while (lm) {
      const uint32_t q = (uint32_t)__builtin_ctzll(lm);
      lm &= lm - 1;
      const uint64_t r = base + q;
      const uint64_t b = off[r], e = off[r + 1];
      const uint32_t cr = __builtin_amdgcn_readlane(c, q);
      uint32_t *dst = unbr + poff[r];
      if (filter != EH_FILTER_MASK) {
        for (uint32_t k = lane; k < cr; k += 64) dst[k] = nbr[b + k];
      } else {
        uint32_t w = 0;
        for (uint64_t k0 = b; k0 < e; k0 += 64) {
          const uint64_t k = k0 + lane;
          const uint32_t keep = (k < e) && mask[k];
          const uint64_t m = __ballot(keep);
          const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
          if (keep) dst[w + below] = nbr[k];
          w += (uint32_t)__popcll(m);
        }
      }
      const uint32_t padded = (cr + 3u) & ~3u;
      for (uint32_t k = cr + lane; k < padded; k += 64) dst[k] = EH_PAD_VALUE;
    }
  }
}

// 16-bit partitions of every row (the pshort layouts' headers): a wave takes 64 consecutive
// rows, a lane counts a row of at most PACK_LANE_MAX ids itself, the wave the longer ones
__global__ void __launch_bounds__(256) k_parts(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr,
                                               uint32_t n, const uint8_t *only_ps, uint32_t *parts) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = wid * 64; base < n; base += nwv * 64) {
    const uint64_t i = base + lane;
    // only_ps (the learned layouts): rows given another layout need no count
    const bool skip = i < n && only_ps && only_ps[i] != EH_PS;
    if (skip) parts[i] = 0u;
    const bool valid = i < n && !skip;
    const uint32_t c = valid ? card[i] : 0u;
    if (valid && c <= PACK_LANE_MAX) {
      const uint32_t *r = unbr + poff[i];
      uint32_t p = c ? 1u : 0u;
      for (uint32_t k = 1; k < c; k++) p += (r[k] >> 16) != (r[k - 1] >> 16);
      parts[i] = p;
    }
    uint64_t lm = __ballot(valid && c > PACK_LANE_MAX);
    while (lm) {
      const uint32_t q = (uint32_t)__builtin_ctzll(lm);
      lm &= lm - 1;
      const uint64_t row = base + q;
      const uint32_t cr = __builtin_amdgcn_readlane(c, q);
      const uint32_t *r = unbr + poff[row];
      uint32_t p = 0;
      for (uint32_t k = 1 + lane; k < cr; k += 64) p += (r[k] >> 16) != (r[k - 1] >> 16);
      for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o, 64);
      if (lane == 0) parts[row] = 1u + p;
    }
  }
}

struct TypeArgs {
  const uint32_t *card;
  const uint64_t *poff;
  const uint32_t *unbr;
  const uint8_t *forced;  // learned per-row types or null
  const uint32_t *parts_in;  // 16-bit partitions per row (k_parts) or null where no row can be pshort
  uint32_t n;
  int policy;
  uint8_t *type;
  uint32_t *parts;
  uint32_t *bs_n;   // words per BS row
  uint32_t *ps_n;   // shorts per PS row, padded to 4
  uint32_t *blk_n;  // bid/bwd slots per BN / NT row (NT: 2 header slots + dense blocks)
  uint32_t *ntu_n;  // ntu words per NT row (uint part padded to 4)
  uint64_t *refb;   // reference image bytes per row
  unsigned long long *type_rows;  // [3] non-empty rows per type
};

__global__ void k_select(TypeArgs a) {
  // wave-uniform grid stride over a bounded grid: each lane counts its rows per type and the
  // wave adds its counts once at the end (one add per wave per row group serialised on the three
  // counters)
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t cnt[3] = {0u, 0u, 0u};
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < a.n; base += stride) {
    const uint64_t i = base + (threadIdx.x & 63u);
    const bool valid = i < a.n;
    uint32_t counted = 3u;  // type index this lane adds to type_rows (3 = none)
    if (valid) {
      const uint32_t c = a.card[i];
      const uint32_t *r = a.unbr + a.poff[i];
      uint32_t lo = c ? r[0] : 0, hi = c ? r[c - 1] : 0;
      const uint32_t nw = c ? (hi >> 6) - (lo >> 6) + 1u : 0u;
      // partitions (k_parts) decide the pshort policies' choice and size every pshort row
      const uint32_t parts = a.parts_in ? a.parts_in[i] : 0u;
      // 128-value blocks: all of them (bitset_new), and new_type's split into blocks with
      // more than 8 elements (density c/128 > 1/16, new_type.hpp:61-63) and a uint part
      uint32_t nblk = 0, ndense = 0, nsparse = 0;
      const bool need_blocks = a.policy >= EH_LAYOUT_COST;
      if (c && need_blocks) {
        uint32_t run = 1;
        for (uint32_t k = 1; k <= c; k++) {
          if (k < c && (r[k] >> 7) == (r[k - 1] >> 7)) { run++; continue; }
          nblk++;
          if (run > 8u) ndense++; else nsparse += run;
          run = 1;
        }
      }
      uint32_t t = EH_U;
      switch (a.policy) {
        case EH_LAYOUT_UINT: t = EH_U; break;
        case EH_LAYOUT_PSHORT: t = EH_PS; break;
        case EH_LAYOUT_BITSET: t = EH_BS; break;
        case EH_LAYOUT_HYBRID:
        case EH_LAYOUT_HYBRID3: {
          // hybrid::get_type: rows of length <= 1 or zero range stay uint
          t = EH_U;
          if (c > 1 && hi - lo > 0) {
            const double density = (double)c / (double)(hi - lo);
            if (a.policy == EH_LAYOUT_HYBRID) {
              t = density > (1.0 / 256.0) ? EH_BS : EH_U;
            } else {
              t = density > (1.0 / 32.0) ? EH_BS : (density > (1.0 / 256.0) ? EH_PS : EH_U);
            }
          }
          break;
        }
        case EH_LAYOUT_COST: {
          // fewest B-side bytes: uint 4c, pshort 2c + 4 parts, bitset 8 words,
          // bitset_new 20 per block
          const uint64_t bu = 4ull * c, bp = 2ull * c + 4ull * parts, bb = 8ull * nw, bn = 20ull * nblk;
          t = EH_U;
          uint64_t best = bu;
          if (c > 1 && bp < best) { t = EH_PS; best = bp; }
          if (c > 1 && bb < best) { t = EH_BS; best = bb; }
          if (c > 1 && bn < best) { t = EH_BN; best = bn; }
          break;
        }
        case EH_LAYOUT_BITSET_NEW: t = EH_BN; break;
        case EH_LAYOUT_NEW_TYPE: t = EH_NT; break;
        case EH_LAYOUT_HYBRID_BLOCK:
          t = EH_U;
          if (c > 1 && hi - lo > 0) t = ((double)c / (double)(hi - lo)) > (1.0 / 256.0) ? EH_BN : EH_U;
          break;
        case EH_LAYOUT_LEARNED:
          t = a.forced ? a.forced[i] : EH_U;
          if (c <= 1) t = EH_U;
          break;
        default: t = EH_U;
      }
      a.type[i] = (uint8_t)t;
      a.parts[i] = parts;
      const uint32_t shorts = c + 2u * parts;
      a.bs_n[i] = (t == EH_BS && c) ? nw : 0u;
      a.ps_n[i] = (t == EH_PS && c) ? ((shorts + 3u) & ~3u) : 0u;
      a.blk_n[i] = !c ? 0u : t == EH_BN ? nblk : t == EH_NT ? EH_NT_HEADER + ndense : 0u;
      a.ntu_n[i] = (t == EH_NT && c) ? ((nsparse + 3u) & ~3u) : 0u;
      // reference row image: optional u64 tag, then the flattened sub-layout
      // (hybrid.hpp:132-166; block rows: [u32 bytes][image], new_type.hpp:90-101)
      const bool tagged = a.policy == EH_LAYOUT_HYBRID || a.policy == EH_LAYOUT_HYBRID3 ||
                          a.policy == EH_LAYOUT_LEARNED || a.policy == EH_LAYOUT_COST ||
                          a.policy == EH_LAYOUT_HYBRID_BLOCK;
      uint64_t bytes = tagged ? 8ull : 0ull;
      if (t == EH_BN) bytes += c ? 4ull + 20ull * nblk : 0ull;
      else if (t == EH_NT) bytes += c ? 4ull + 8ull + 4ull * nsparse + 20ull * ndense : 0ull;
      else if (is_uint_type(t)) bytes += 4ull * c;
      else if (t == EH_BS) bytes += c ? 4ull + 8ull + 8ull * nw : 0ull;
      else bytes += c ? 4ull + 2ull * shorts : 0ull;
      a.refb[i] = bytes;
      // block rows are counted with the bitsets
      if (c) counted = t == EH_UG ? EH_U : is_block_type(t) ? EH_BS : t;
    }
    for (uint32_t q = 0; q < 3; q++) cnt[q] += counted == q ? 1u : 0u;
  }
  for (uint32_t q = 0; q < 3; q++) {
    uint32_t v = cnt[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63u) == 0 && v) atomicAdd(&a.type_rows[q], (unsigned long long)v);
  }
}

// one wave per bitset row, over words cleared beforehand, 64 ids a step: the ids of one word
// are OR-ed together across lanes (a segmented reduction toward the word's first lane, the ids
// being sorted) and that lane ORs the word in, one atomic per word and step instead of one per id
__global__ void __launch_bounds__(256) k_fill_bitset(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr,
                                                     const uint8_t *type, const uint64_t *bsoff, uint32_t n,
                                                     unsigned long long *bsw) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nwv) {
    if (type[i] != EH_BS) continue;
    const uint32_t c = card[i];
    if (!c) continue;
    const uint32_t *r = unbr + poff[i];
    const uint32_t w0 = r[0] >> 6;
    unsigned long long *W = bsw + bsoff[i];
    for (uint32_t k0 = 0; k0 < c; k0 += 64) {
      const uint32_t k = k0 + lane;
      const bool in = k < c;
      const uint32_t x = in ? r[k] : 0u;
      const uint32_t wi = in ? (x >> 6) - w0 : 0xFFFFFFFFu;
      unsigned long long bits = in ? 1ull << (x & 63u) : 0ull;
      for (uint32_t o = 1; o < 64; o <<= 1) {
        const unsigned long long ob = __shfl_down(bits, o, 64);
        const uint32_t ow = __shfl_down(wi, o, 64);
        if (lane + o < 64u && ow == wi) bits |= ob;
      }
      const uint32_t pw = __shfl_up(wi, 1, 64);
      if (in && (lane == 0u || pw != wi)) atomicOr(&W[wi], bits);
    }
  }
}

// pshort rows: [prefix, count-1, low16...] per 16-bit prefix (pshort.hpp:46-78), then zeros to
// the row's padded length.  A wave takes 64 consecutive rows: a lane encodes a row of at most
// lane_max ids itself, then the wave encodes the longer ones together, 64 ids a step (stores
// of neighbouring shorts): an id at position k in partition p goes to k + 2 (p + 1); a
// partition's first id writes its prefix, its last id the count.  lane_max 0: every row by
// the wave.

__global__ void __launch_bounds__(256) k_fill_pshort(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr,
                                                     const uint8_t *type, const uint64_t *psoff, const uint32_t *ps_n,
                                                     uint32_t n, uint32_t lane_max, uint16_t *psd) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = wid * 64; base < n; base += nwv * 64) {
    const uint64_t i = base + lane;
    const bool ps = i < n && type[i] == EH_PS;
    const uint32_t c = ps ? card[i] : 0u;
    if (ps && c && c <= lane_max) {
      const uint32_t *r = unbr + poff[i];
      uint16_t *S = psd + psoff[i];
      uint32_t w = 0, hdr = 0, cnt = 0, cur = 0xFFFFFFFFu;
      for (uint32_t k = 0; k < c; k++) {
        const uint32_t hi = r[k] >> 16;
        if (hi != cur) {
          if (cur != 0xFFFFFFFFu) S[hdr + 1] = (uint16_t)(cnt - 1u);
          hdr = w;
          S[w++] = (uint16_t)hi;
          S[w++] = 0;
          cur = hi;
          cnt = 0;
        }
        S[w++] = (uint16_t)(r[k] & 0xFFFFu);
        cnt++;
      }
      S[hdr + 1] = (uint16_t)(cnt - 1u);
      for (uint32_t k = w; k < ps_n[i]; k++) S[k] = 0;
    }
    uint64_t lm = __ballot(ps && c > lane_max);
    while (lm) {
      const uint32_t q = (uint32_t)__builtin_ctzll(lm);
      lm &= lm - 1;
      const uint64_t row = base + q;
      const uint32_t cr = __builtin_amdgcn_readlane(c, q);
      const uint32_t *r = unbr + poff[row];
      uint16_t *S = psd + psoff[row];
      uint32_t parts = 0, start = 0;  // partitions before this step; the open partition's first position
      for (uint32_t k0 = 0; k0 < cr; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool in = k < cr;
        const uint32_t x = in ? r[k] : 0u;
        const uint32_t hi = x >> 16;
        const bool first = in && (k == 0 || (r[k - 1] >> 16) != hi);
        const bool last = in && (k + 1 == cr || (r[k + 1] >> 16) != hi);
        const uint64_t fm = __ballot(first);
        const uint64_t upto = fm & (lane == 63u ? ~0ull : ((2ull << lane) - 1ull));
        const uint32_t p = parts + (uint32_t)__popcll(upto) - 1u;  // this id's partition
        const uint32_t s = upto ? k0 + 63u - (uint32_t)__builtin_clzll(upto) : start;
        if (in) {
          S[k + 2u * (p + 1u)] = (uint16_t)(x & 0xFFFFu);
          if (first) S[k + 2u * p] = (uint16_t)hi;
          if (last) S[s + 2u * p + 1u] = (uint16_t)(k - s);
        }
        // carry the partition still open at the step's end
        parts += (uint32_t)__popcll(fm);
        if (fm) start = k0 + 63u - (uint32_t)__builtin_clzll(fm);
      }
      const uint32_t w = cr + 2u * parts;
      for (uint32_t k = w + lane; k < ps_n[row]; k += 64) S[k] = 0;
    }
  }
}

// one thread per bitset_new / new_type row: the blocks in id order, each with its two
// words (bit v & 127; word (v >> 6) & 1), and new_type's sparse elements in ntu.  A
// new_type row's first two slots hold [uint part count, uint part offset / 4].
__global__ void k_fill_blocks(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr, const uint8_t *type,
                              const uint64_t *blkoff, const uint64_t *ntuoff, const uint32_t *ntu_n, uint32_t n,
                              uint32_t *bid, uint64_t *bwd, uint32_t *ntu) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t = type[i];
    const uint32_t c = card[i];
    if (!is_block_type(t) || !c) continue;
    const uint32_t *r = unbr + poff[i];
    uint64_t slot = blkoff[i];
    uint32_t *U = ntu + ntuoff[i];
    uint32_t nu = 0;
    if (t == EH_NT) slot += EH_NT_HEADER;
    uint32_t k = 0;
    while (k < c) {
      const uint32_t b = r[k] >> 7;
      uint32_t e = k;
      while (e < c && (r[e] >> 7) == b) e++;
      if (t == EH_BN || e - k > 8u) {
        uint64_t w[2] = {0ull, 0ull};
        for (uint32_t q = k; q < e; q++) w[(r[q] >> 6) & 1u] |= 1ull << (r[q] & 63u);
        bid[slot] = b;
        bwd[2 * slot] = w[0];
        bwd[2 * slot + 1] = w[1];
        slot++;
      } else {
        for (uint32_t q = k; q < e; q++) U[nu++] = r[q];
      }
      k = e;
    }
    if (t == EH_NT) {
      const uint64_t h = blkoff[i];
      bid[h] = nu;
      bid[h + 1] = (uint32_t)(ntuoff[i] >> 2);
      bwd[2 * h] = bwd[2 * h + 1] = bwd[2 * h + 2] = bwd[2 * h + 3] = 0ull;
      for (uint32_t q = nu; q < ntu_n[i]; q++) U[q] = EH_PAD_VALUE;
    }
  }
}

__global__ void k_desc(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr, const uint8_t *type,
                       const uint32_t *parts, const uint32_t *bs_n, const uint32_t *blk_n, const uint64_t *bsoff,
                       const uint64_t *psoff, const uint64_t *blkoff, uint32_t n, RowDesc *desc) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = card[i];
    const uint32_t t = type[i];
    RowDesc d;
    d.card = c;
    d.w0 = 0;
    d.off = (uint32_t)(poff[i] >> 2);
    d.tinfo = t;  // empty rows keep their tag and carry no payload
    if (c && t == EH_BS) {
      d.tinfo = make_tinfo(EH_BS, bs_n[i]);
      d.w0 = unbr[poff[i]] >> 6;
      d.off = (uint32_t)bsoff[i];
    } else if (c && t == EH_PS) {
      const uint32_t shorts = c + 2u * parts[i];
      d.tinfo = make_tinfo(EH_PS, shorts);
      d.w0 = parts[i];
      d.off = (uint32_t)(psoff[i] >> 2);
    } else if (c && is_block_type(t)) {
      // the canonical list stays at off; the blocks start at slot w0
      d.tinfo = make_tinfo(t, blk_n[i] - (t == EH_NT ? EH_NT_HEADER : 0u));
      d.w0 = (uint32_t)blkoff[i];
    }
    desc[i] = d;
  }
}

// sum over every kept entry j of card(j) (the wedges of the triangle pass), flat over the
// padded rows (the pad sentinel is no id): per-lane sums, one add per wave
__global__ void __launch_bounds__(256) k_wedge_sum(const uint32_t *card, const uint32_t *unbr, uint64_t words,
                                                   unsigned long long *out) {
  const uint4 *U = reinterpret_cast<const uint4 *>(unbr);
  const uint64_t nv = words >> 2;
  unsigned long long s = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 x = U[v];
    if (x.x != EH_PAD_VALUE) s += card[x.x];
    if (x.y != EH_PAD_VALUE) s += card[x.y];
    if (x.z != EH_PAD_VALUE) s += card[x.z];
    if (x.w != EH_PAD_VALUE) s += card[x.w];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63u) == 0 && s) atomicAdd(out, s);
}

// caller-given arms, a bitset only within `guard` words per id (else uint), unknown values uint
__global__ void k_guard_types(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr, const uint8_t *in,
                              uint32_t n, uint32_t guard, uint8_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = card[i];
    uint8_t t = in[i];
    if (t > EH_UG || c < 2) t = EH_U;
    if (t == EH_BS) {
      const uint32_t *r = unbr + poff[i];
      if ((uint64_t)(r[c - 1] >> 6) - (r[0] >> 6) + 1ull > (uint64_t)guard * c) t = EH_U;
    }
    out[i] = t;
  }
}

__global__ void k_u32_to_u64_plus(const uint32_t *in, uint32_t n, uint64_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

}  // namespace

eh_status matrix_build_dev(eh_ctx *ctx, uint32_t n, const uint64_t *d_off, const uint32_t *d_nbr, uint64_t entries,
                           eh_edge_filter filter, const uint8_t *d_mask, eh_layout_policy policy,
                           const uint32_t *d_graph_of, bool symmetric, eh_matrix **out,
                           const uint8_t *d_types_in) {
  (void)entries;
  if (policy < EH_LAYOUT_UINT || policy > EH_LAYOUT_HYBRID_BLOCK) { eh_set_error("unknown layout policy %d", (int)policy); return EH_ERR_INVALID; }
  if (filter < EH_FILTER_ALL || filter > EH_FILTER_MASK) { eh_set_error("unknown edge filter %d", (int)filter); return EH_ERR_INVALID; }
  if (filter == EH_FILTER_MASK && !d_mask) { eh_set_error("EH_FILTER_MASK needs an edge mask"); return EH_ERR_INVALID; }
  hipStream_t st = ctx->stream;
  eh_matrix *m = new eh_matrix();
  m->ctx = ctx;
  m->n = n;
  m->policy = policy;
  m->filter = filter;
  eh_status s = EH_OK;
  auto fail = [&](eh_status e) { delete m; return e; };
  DevBuf pc, poff, type, parts, bs_n, ps_n, blk_n, ntu_n, bsoff, psoff, blkoff, ntuoff, refb, stats;
  if ((s = m->dcard.alloc(4ull * n + 4)) != EH_OK) return fail(s);
  if ((s = m->daoff.alloc(4ull * n + 4)) != EH_OK) return fail(s);
  if ((s = pc.alloc(4ull * n + 8, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = poff.alloc(8ull * n + 16, &ctx->cache)) != EH_OK) return fail(s);
  if (hipEventRecord(ctx->ev0, st) != hipSuccess) return fail(EH_ERR_HIP);
  k_filter_count<<<grid_for(n), 256, 0, st>>>(d_off, d_nbr, n, filter, d_mask, m->dcard.as<uint32_t>());
  k_pad4<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(m->dcard.as<uint32_t>(), n, pc.as<uint32_t>());
  if ((s = prim_exscan_u32_to_u64(ctx, pc.as<uint32_t>(), poff.as<uint64_t>(), (uint64_t)n + 1)) != EH_OK) return fail(s);
  uint64_t total = 0;
  if (hipMemcpy(&total, poff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost) != hipSuccess) return fail(EH_ERR_HIP);
  if (total >= (1ull << 34)) { eh_set_error("matrix too large for 32-bit row offsets"); return fail(EH_ERR_UNSUPPORTED); }
  m->unbr_words = total;
  if ((s = m->dunbr.alloc(4 * total + 64)) != EH_OK) return fail(s);
  k_aoff<<<grid_for(n), 256, 0, st>>>(poff.as<uint64_t>(), n, m->daoff.as<uint32_t>());
  k_pack<<<grid_for(n), 256, 0, st>>>(d_off, d_nbr, n, filter, d_mask, m->dcard.as<uint32_t>(),
                                                     poff.as<uint64_t>(), m->dunbr.as<uint32_t>());
  if ((s = prim_reduce_max_u32(ctx, m->dcard.as<uint32_t>(), n, &m->max_row)) != EH_OK) return fail(s);
  // the graph of every row (batch arenas), before the learned selection, which treats an arena
  // as read wholly by A-side and edge-centric passes (every row held by another is a candidate)
  if (d_graph_of) {
    if ((s = m->dgraph_of.alloc(4ull * n + 4)) != EH_OK) return fail(s);
    if (hipMemcpyAsync(m->dgraph_of.p, d_graph_of, 4ull * n, hipMemcpyDeviceToDevice, st) != hipSuccess) return fail(EH_ERR_HIP);
  }
  // learned types are computed from the packed rows before selection
  DevBuf forced;
  if (policy == EH_LAYOUT_LEARNED) {
    if ((s = forced.alloc((size_t)n + 16, &ctx->cache)) != EH_OK) return fail(s);
    m->card = 0;
    if (d_types_in) {
      k_guard_types<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(), m->dunbr.as<uint32_t>(),
                                                 d_types_in, n, ctx->sel_guard, forced.as<uint8_t>());
    } else if ((s = matrix_select_learned(m, forced.as<uint8_t>(), symmetric ? d_off : nullptr)) != EH_OK) {
      return fail(s);
    }
  }
  if ((s = type.alloc((size_t)n + 16, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = parts.alloc(4ull * n + 4, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = bs_n.alloc(4ull * n + 8, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = ps_n.alloc(4ull * n + 8, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = blk_n.alloc(4ull * n + 8, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = ntu_n.alloc(4ull * n + 8, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = refb.alloc(8ull * n + 8, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = stats.alloc(64, &ctx->cache)) != EH_OK) return fail(s);
  if (hipMemsetAsync(stats.p, 0, 64, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemsetAsync(bs_n.p, 0, 4ull * n + 8, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemsetAsync(ps_n.p, 0, 4ull * n + 8, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemsetAsync(blk_n.p, 0, 4ull * n + 8, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemsetAsync(ntu_n.p, 0, 4ull * n + 8, st) != hipSuccess) return fail(EH_ERR_HIP);
  TypeArgs ta;
  ta.card = m->dcard.as<uint32_t>();
  ta.poff = poff.as<uint64_t>();
  ta.unbr = m->dunbr.as<uint32_t>();
  ta.forced = forced.p ? forced.as<uint8_t>() : nullptr;
  // partitions, where a row can be pshort
  const bool any_ps = policy == EH_LAYOUT_PSHORT || policy == EH_LAYOUT_HYBRID3 || policy == EH_LAYOUT_COST ||
                      policy == EH_LAYOUT_LEARNED;
  DevBuf parts_in;
  if (any_ps) {
    if ((s = parts_in.alloc(4ull * n + 4, &ctx->cache)) != EH_OK) return fail(s);
    k_parts<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(), m->dunbr.as<uint32_t>(), n,
                                         policy == EH_LAYOUT_LEARNED ? forced.as<uint8_t>() : nullptr,
                                         parts_in.as<uint32_t>());
  }
  ta.parts_in = any_ps ? parts_in.as<uint32_t>() : nullptr;
  ta.n = n;
  ta.policy = policy;
  ta.type = type.as<uint8_t>();
  ta.parts = parts.as<uint32_t>();
  ta.bs_n = bs_n.as<uint32_t>();
  ta.ps_n = ps_n.as<uint32_t>();
  ta.blk_n = blk_n.as<uint32_t>();
  ta.ntu_n = ntu_n.as<uint32_t>();
  ta.refb = refb.as<uint64_t>();
  ta.type_rows = stats.as<unsigned long long>();
  const unsigned bounded = std::min(grid_for(n), (unsigned)(ctx->num_cus ? ctx->num_cus : 256) * 8u);
  k_select<<<bounded, 256, 0, st>>>(ta);
  if ((s = bsoff.alloc(8ull * n + 16, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = psoff.alloc(8ull * n + 16, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = prim_exscan_u32_to_u64(ctx, bs_n.as<uint32_t>(), bsoff.as<uint64_t>(), (uint64_t)n + 1)) != EH_OK) return fail(s);
  if ((s = prim_exscan_u32_to_u64(ctx, ps_n.as<uint32_t>(), psoff.as<uint64_t>(), (uint64_t)n + 1)) != EH_OK) return fail(s);
  if ((s = blkoff.alloc(8ull * n + 16, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = ntuoff.alloc(8ull * n + 16, &ctx->cache)) != EH_OK) return fail(s);
  if ((s = prim_exscan_u32_to_u64(ctx, blk_n.as<uint32_t>(), blkoff.as<uint64_t>(), (uint64_t)n + 1)) != EH_OK) return fail(s);
  if ((s = prim_exscan_u32_to_u64(ctx, ntu_n.as<uint32_t>(), ntuoff.as<uint64_t>(), (uint64_t)n + 1)) != EH_OK) return fail(s);
  uint64_t hb[4] = {0, 0, 0, 0};
  if (hipMemcpyAsync(&hb[0], bsoff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemcpyAsync(&hb[1], psoff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemcpyAsync(&hb[2], blkoff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemcpyAsync(&hb[3], ntuoff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  unsigned long long trows[3];
  if (hipMemcpyAsync(trows, stats.p, 24, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipStreamSynchronize(st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hb[0] >= (1ull << 32) || (hb[1] >> 2) >= (1ull << 32) || hb[2] >= (1ull << 32) || (hb[3] >> 2) >= (1ull << 32)) {
    eh_set_error("layout payload too large for 32-bit offsets");
    return fail(EH_ERR_UNSUPPORTED);
  }
  m->bs_words = hb[0];
  m->ps_shorts = hb[1];
  m->blk_slots = hb[2];
  m->ntu_words = hb[3];
  for (int q = 0; q < 3; q++) m->rows_by_type[q] = trows[q];
  if ((s = m->dbsw.alloc(8 * hb[0] + 64)) != EH_OK) return fail(s);
  if ((s = m->dpsd.alloc(2 * hb[1] + 64)) != EH_OK) return fail(s);
  if (hipMemsetAsync(m->dbsw.p, 0, 8 * hb[0] + 64, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hb[0])
    k_fill_bitset<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(),
                                                              m->dunbr.as<uint32_t>(), type.as<uint8_t>(),
                                                              bsoff.as<uint64_t>(), n,
                                                              m->dbsw.as<unsigned long long>());
  if (hb[1])
    k_fill_pshort<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(), m->dunbr.as<uint32_t>(),
                                               type.as<uint8_t>(), psoff.as<uint64_t>(), ps_n.as<uint32_t>(), n, 0u,
                                               m->dpsd.as<uint16_t>());
  if ((s = m->dbid.alloc(4 * hb[2] + 64)) != EH_OK) return fail(s);
  if ((s = m->dbwd.alloc(16 * hb[2] + 64)) != EH_OK) return fail(s);
  if ((s = m->dntu.alloc(4 * hb[3] + 64)) != EH_OK) return fail(s);
  if (hb[2])
    k_fill_blocks<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(), m->dunbr.as<uint32_t>(),
                                               type.as<uint8_t>(), blkoff.as<uint64_t>(), ntuoff.as<uint64_t>(),
                                               ntu_n.as<uint32_t>(), n, m->dbid.as<uint32_t>(),
                                               m->dbwd.as<uint64_t>(), m->dntu.as<uint32_t>());
  if ((s = m->ddesc.alloc(16ull * n + 16)) != EH_OK) return fail(s);
  k_desc<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(), m->dunbr.as<uint32_t>(),
                                      type.as<uint8_t>(), parts.as<uint32_t>(), bs_n.as<uint32_t>(), blk_n.as<uint32_t>(),
                                      bsoff.as<uint64_t>(), psoff.as<uint64_t>(), blkoff.as<uint64_t>(), n,
                                      m->ddesc.as<RowDesc>());
  if (hipEventRecord(ctx->ev1, st) != hipSuccess) return fail(EH_ERR_HIP);
  k_wedge_sum<<<bounded, 256, 0, st>>>(m->dcard.as<uint32_t>(), m->dunbr.as<uint32_t>(), m->unbr_words,
                                       stats.as<unsigned long long>() + 4);
  if (hipGetLastError() != hipSuccess) { eh_set_error("build kernel launch failed"); return fail(EH_ERR_HIP); }
  if (hipMemcpy(&m->wedges, stats.as<unsigned long long>() + 4, 8, hipMemcpyDeviceToHost) != hipSuccess)
    return fail(EH_ERR_HIP);
  if ((s = prim_reduce_sum_u64(ctx, refb.as<uint64_t>(), n, &m->ref_row_bytes)) != EH_OK) return fail(s);
  // kept entries = sum of card
  {
    DevBuf c64;
    if ((s = c64.alloc(8ull * n + 8, &ctx->cache)) != EH_OK) return fail(s);
    k_u32_to_u64_plus<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), n, c64.as<uint64_t>());
    if ((s = prim_reduce_sum_u64(ctx, c64.as<uint64_t>(), n, &m->card)) != EH_OK) return fail(s);
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) ms = 0;
  m->build_ms = ms;
  if (hipStreamSynchronize(st) != hipSuccess) return fail(EH_ERR_HIP);
  *out = m;
  return EH_OK;
}

// ---------------------------------------------------------------- C ABI

extern "C" eh_status eh_matrix_build(eh_graph *g, eh_edge_filter filter, const uint8_t *edge_mask,
                                     eh_layout_policy policy, eh_matrix **out) {
  if (!g || !out) { eh_set_error("eh_matrix_build: null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(g->ctx->device));
  DevBuf dmask;
  if (filter == EH_FILTER_MASK) {
    if (!edge_mask) { eh_set_error("EH_FILTER_MASK needs an edge mask"); return EH_ERR_INVALID; }
    EH_CHECK(dmask.alloc(g->entries + 16));
    if (g->entries) EH_HIP(hipMemcpy(dmask.p, edge_mask, g->entries, hipMemcpyHostToDevice));
  }
  return matrix_build_dev(g->ctx, g->n, g->off.as<uint64_t>(), g->nbr.as<uint32_t>(), g->entries, filter,
                          dmask.as<uint8_t>(), policy, nullptr, g->symmetric, out);
}

extern "C" eh_status eh_matrix_build_typed(eh_graph *g, eh_edge_filter filter, const uint8_t *row_types,
                                           eh_matrix **out) {
  if (!g || !out || !row_types) { eh_set_error("eh_matrix_build_typed: null argument"); return EH_ERR_INVALID; }
  if (filter == EH_FILTER_MASK) { eh_set_error("eh_matrix_build_typed: EH_FILTER_MASK is not supported"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(g->ctx->device));
  DevBuf dt;
  EH_CHECK(dt.alloc((size_t)g->n + 16));
  EH_HIP(hipMemcpy(dt.p, row_types, g->n, hipMemcpyHostToDevice));
  return matrix_build_dev(g->ctx, g->n, g->off.as<uint64_t>(), g->nbr.as<uint32_t>(), g->entries, filter, nullptr,
                          EH_LAYOUT_LEARNED, nullptr, g->symmetric, out, dt.as<uint8_t>());
}

extern "C" eh_status eh_matrix_build_host(eh_ctx *ctx, uint32_t n, const uint64_t *row_off, const uint32_t *nbrs,
                                          eh_edge_filter filter, const uint8_t *edge_mask, eh_layout_policy policy,
                                          eh_matrix **out) {
  eh_graph *g = nullptr;
  EH_CHECK(eh_graph_from_csr(ctx, n, row_off, nbrs, &g));
  eh_status s = eh_matrix_build(g, filter, edge_mask, policy, out);
  eh_graph_free(g);
  return s;
}

extern "C" eh_status eh_matrix_info_get(eh_matrix *m, eh_matrix_info *out) {
  if (!m || !out) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  out->n = m->n;
  out->max_row = m->max_row;
  out->cardinality = m->card;
  out->rows_by_type[0] = m->rows_by_type[0];
  out->rows_by_type[1] = m->rows_by_type[1];
  out->rows_by_type[2] = m->rows_by_type[2];
  out->payload_bytes = 8 * m->bs_words + 2 * m->ps_shorts + 4 * m->unbr_words + 20 * m->blk_slots + 4 * m->ntu_words;
  out->ref_row_bytes = m->ref_row_bytes;
  out->wedges = m->wedges;
  out->build_ms = m->build_ms;
  return EH_OK;
}

extern "C" eh_status eh_matrix_row(eh_matrix *m, uint32_t row, void *buf, size_t cap, eh_row_view *view) {
  if (!m || !view || (cap && !buf)) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (row >= m->n) { eh_set_error("row %u out of range", row); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(m->ctx->device));
  RowDesc d;
  EH_HIP(hipMemcpy(&d, m->ddesc.as<RowDesc>() + row, sizeof(d), hipMemcpyDeviceToHost));
  const uint32_t t = desc_type(d);
  const uint32_t nblk = is_block_type(t) ? desc_aux(d) : 0u;
  uint32_t hdr[2] = {0u, 0u};  // new_type: uint part count, uint part offset / 4
  if (d.card && t == EH_NT) EH_HIP(hipMemcpy(hdr, m->dbid.as<uint32_t>() + d.w0, 8, hipMemcpyDeviceToHost));
  uint64_t bytes = 0;
  if (d.card == 0) bytes = 0;
  else if (t == EH_BN) bytes = 20ull * nblk;
  else if (t == EH_NT) bytes = 8ull + 4ull * hdr[0] + 20ull * nblk;
  else if (is_uint_type(t)) bytes = 4ull * d.card;
  else if (t == EH_BS) bytes = 8ull + 8ull * desc_aux(d);
  else bytes = 2ull * desc_aux(d);
  view->data = buf;
  view->cardinality = d.card;
  view->number_of_bytes = bytes;
  view->type = (int32_t)(t == EH_BN ? EH_SET_BITSET_NEW : t == EH_NT ? EH_SET_NEW_TYPE : is_uint_type(t) ? EH_U : t);
  view->reserved = 0;
  if (bytes > cap) { eh_set_error("row needs %llu bytes", (unsigned long long)bytes); return EH_ERR_CAPACITY; }
  if (!bytes) return EH_OK;
  if (is_block_type(t)) {
    // [new_type: u64 uint bytes, uint part][u32 block ids][2 u64 words per block]
    uint8_t *p = (uint8_t *)buf;
    uint64_t first = d.w0;
    if (t == EH_NT) {
      const uint64_t ub = 4ull * hdr[0];
      memcpy(p, &ub, 8);
      if (ub) EH_HIP(hipMemcpy(p + 8, m->dntu.as<uint32_t>() + 4ull * hdr[1], ub, hipMemcpyDeviceToHost));
      p += 8 + ub;
      first += EH_NT_HEADER;
    }
    if (nblk) {
      EH_HIP(hipMemcpy(p, m->dbid.as<uint32_t>() + first, 4ull * nblk, hipMemcpyDeviceToHost));
      EH_HIP(hipMemcpy(p + 4ull * nblk, m->dbwd.as<uint64_t>() + 2 * first, 16ull * nblk, hipMemcpyDeviceToHost));
    }
  } else if (is_uint_type(t)) {
    EH_HIP(hipMemcpy(buf, m->dunbr.as<uint32_t>() + 4ull * d.off, bytes, hipMemcpyDeviceToHost));
  } else if (t == EH_BS) {
    const uint64_t w0 = d.w0;
    memcpy(buf, &w0, 8);
    EH_HIP(hipMemcpy((uint8_t *)buf + 8, m->dbsw.as<uint64_t>() + d.off, bytes - 8, hipMemcpyDeviceToHost));
  } else {
    EH_HIP(hipMemcpy(buf, m->dpsd.as<uint16_t>() + 4ull * d.off, bytes, hipMemcpyDeviceToHost));
  }
  return EH_OK;
}

extern "C" eh_status eh_matrix_download(eh_matrix *m, uint64_t *row_off, uint32_t *nbrs) {
  if (!m || !row_off || (m->card && !nbrs)) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(m->ctx->device));
  std::vector<uint32_t> card(m->n), aoff(m->n);
  std::vector<uint32_t> un(m->unbr_words);
  EH_HIP(hipMemcpy(card.data(), m->dcard.p, 4ull * m->n, hipMemcpyDeviceToHost));
  EH_HIP(hipMemcpy(aoff.data(), m->daoff.p, 4ull * m->n, hipMemcpyDeviceToHost));
  if (m->unbr_words) EH_HIP(hipMemcpy(un.data(), m->dunbr.p, 4ull * m->unbr_words, hipMemcpyDeviceToHost));
  row_off[0] = 0;
  for (uint32_t i = 0; i < m->n; i++) {
    memcpy(nbrs + row_off[i], un.data() + 4ull * aoff[i], 4ull * card[i]);
    row_off[i + 1] = row_off[i] + card[i];
  }
  return EH_OK;
}

extern "C" eh_status eh_matrix_row_types(eh_matrix *m, uint8_t *types) {
  if (!m || !types) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(m->ctx->device));
  std::vector<RowDesc> d(m->n);
  if (m->n) EH_HIP(hipMemcpy(d.data(), m->ddesc.p, 16ull * m->n, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < m->n; i++) {
    const uint32_t t = desc_type(d[i]);
    types[i] = (uint8_t)(t == EH_BN ? EH_SET_BITSET_NEW : t == EH_NT ? EH_SET_NEW_TYPE : t);
  }
  return EH_OK;
}

extern "C" void eh_matrix_free(eh_matrix *m) { delete m; }
