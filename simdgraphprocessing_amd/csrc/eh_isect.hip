// eh_isect.hip -- the set-level operation ops::set_intersect(Set<X>*, const Set<Y>*,
// const Set<Z>*) (src/set/ops/intersection.hpp:829-1336, 2066-2118) on the device, and a
// batched form over rows of one matrix.
//
// eh_intersect takes two host sets in the reference's byte formats
//   uint       : u32[card]
//   bitset     : [u64 first word][u64 words...]         (src/set/layouts/bitset.hpp:65-93)
//   pshort     : u16 stream [prefix, count-1, low16...]  (src/set/layouts/pshort.hpp:46-78)
//   bitset_new : [u32 block ids][2 u64 words per block]  (bitset_new.hpp:66-108)
//   new_type   : [u64 uint bytes][uint part][bitset_new part] (new_type.hpp:50-88)
// uploads them, decodes both to sorted u32 lists, intersects, and (when asked)
// encodes the result in the reference's result layout: uint for any pair with a
// uint side, the AND over the overlapping word span for bitset x bitset, a
// pshort stream (empty partitions dropped) for pshort x pshort and pshort x bitset,
// the matched blocks for bitset_new x bitset_new, and for new_type x new_type the
// reference's split (uint parts against each other and the other side's blocks,
// merged; blocks AND-ed).
// Each call is one small chain of single-workgroup kernels; it serves wrappers and
// tests, not the counting pass.
//
// eh_intersect_rows counts |row(a[p]) & row(b[p])| for many pairs of one matrix with
// one wave per pair over the rows' sorted u32 lists already resident in HBM.
#include <mutex>

#include "eh_internal.hpp"

namespace {

constexpr int ISECT_THREADS = 1024;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t wave_scan_excl(uint32_t v, uint32_t lane, uint32_t *total) {
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// exclusive scan over the whole workgroup; *total is the workgroup sum
__device__ uint32_t block_scan_excl(uint32_t v, uint32_t *sh, uint32_t *total) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint32_t wt;
  const uint32_t x = wave_scan_excl(v, lane, &wt);
  if (lane == 0) sh[wave] = wt;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < nw; w++) {
    const uint32_t s = sh[w];
    before += (w < wave) ? s : 0u;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + x;
}

__device__ __forceinline__ uint64_t ld64u(const uint8_t *p) {
  const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
  return (uint64_t)q[0] | ((uint64_t)q[1] << 32);
}

// decode one set to a sorted u32 list (uint: copy; bitset: expand bits in word order;
// pshort: partitions located by thread 0, then expanded by all waves)
__global__ void __launch_bounds__(ISECT_THREADS)
k_decode_set(int type, const uint8_t *payload, uint64_t bytes, uint64_t card, uint32_t *out, uint32_t *pscratch) {
  // out holds card entries; a bitset with more bits than card keeps its first card
  // elements (the host checks pshort streams before launch)
  __shared__ uint32_t sh[32];
  if (type == EH_U) {
    const uint32_t *u = reinterpret_cast<const uint32_t *>(payload);
    for (uint64_t k = threadIdx.x; k < card; k += blockDim.x) out[k] = u[k];
    return;
  }
  if (type == EH_BS) {
    if (bytes < 8) return;
    const uint64_t w0 = ld64u(payload), nw = (bytes - 8) / 8;
    const uint8_t *W = payload + 8;
    uint32_t base = 0;
    for (uint64_t c0 = 0; c0 < nw; c0 += blockDim.x) {
      const uint64_t w = c0 + threadIdx.x;
      uint64_t word = w < nw ? ld64u(W + 8 * w) : 0ull;
      uint32_t tot;
      uint32_t pos = base + block_scan_excl((uint32_t)__popcll(word), sh, &tot);
      while (word) {
        const uint32_t b = (uint32_t)__builtin_ctzll(word);
        if (pos < card) out[pos] = (uint32_t)((w0 + w) * 64ull + b);
        pos++;
        word &= word - 1;
      }
      base += tot;
    }
    return;
  }
  if (type == EH_SET_BITSET_NEW) {
    // blocks in id order; a block's elements follow its popcount prefix
    const uint64_t nb = bytes / 20;
    const uint8_t *W = payload + 4 * nb;
    uint32_t base = 0;
    for (uint64_t c0 = 0; c0 < nb; c0 += blockDim.x) {
      const uint64_t b = c0 + threadIdx.x;
      uint64_t w0 = 0, w1 = 0;
      uint32_t id = 0;
      if (b < nb) {
        id = reinterpret_cast<const uint32_t *>(payload)[b];
        w0 = ld64u(W + 16 * b);
        w1 = ld64u(W + 16 * b + 8);
      }
      uint32_t tot;
      uint32_t pos = base + block_scan_excl((uint32_t)(__popcll(w0) + __popcll(w1)), sh, &tot);
      for (uint32_t h = 0; h < 2; h++) {
        uint64_t word = h ? w1 : w0;
        while (word) {
          const uint32_t bit = (uint32_t)__builtin_ctzll(word);
          if (pos < card) out[pos] = 128u * id + 64u * h + bit;
          pos++;
          word &= word - 1;
        }
      }
      base += tot;
    }
    return;
  }
  // pshort: pscratch holds per partition {first short, element count, output base}
  const uint16_t *S = reinterpret_cast<const uint16_t *>(payload);
  const uint64_t ns = bytes / 2;
  if (threadIdx.x == 0) {
    uint32_t np = 0, ob = 0;
    for (uint64_t i = 0; i + 1 < ns;) {
      const uint32_t cnt = (uint32_t)S[i + 1] + 1u;
      pscratch[1 + 3 * np] = (uint32_t)i;
      pscratch[2 + 3 * np] = cnt;
      pscratch[3 + 3 * np] = ob;
      ob += cnt;
      np++;
      i += 2ull + cnt;
    }
    pscratch[0] = np;
  }
  __syncthreads();
  const uint32_t np = pscratch[0];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  for (uint32_t p = wave; p < np; p += nwv) {
    const uint32_t s = pscratch[1 + 3 * p], cnt = pscratch[2 + 3 * p], ob = pscratch[3 + 3 * p];
    const uint32_t pre = (uint32_t)S[s] << 16;
    for (uint32_t t = lane; t < cnt; t += 64) out[ob + t] = pre | S[s + 2 + t];
  }
}

__device__ __forceinline__ bool list_contains(const uint32_t *L, uint32_t n, uint32_t x) {
  if (n == 0) return false;
  uint32_t base = 0, len = n;
  while (len > 1) {
    const uint32_t half = len >> 1;
    base = (L[base + half] <= x) ? base + half : base;
    len -= half;
  }
  return L[base] == x;
}

// R = S & L (S the shorter list; order of S is kept, so R is sorted); res[0] = |R|
__global__ void __launch_bounds__(ISECT_THREADS)
k_isect_lists(const uint32_t *S, uint32_t ns, const uint32_t *L, uint32_t nl, uint32_t *R, unsigned long long *res) {
  __shared__ uint32_t sh[32];
  uint32_t base = 0;
  for (uint32_t c0 = 0; c0 < ns; c0 += blockDim.x) {
    const uint32_t k = c0 + threadIdx.x;
    const uint32_t x = k < ns ? S[k] : 0u;
    const uint32_t hit = (k < ns && list_contains(L, nl, x)) ? 1u : 0u;
    uint32_t tot;
    const uint32_t pos = base + block_scan_excl(hit, sh, &tot);
    if (hit && R) R[pos] = x;
    base += tot;
  }
  if (threadIdx.x == 0) res[0] = base;
}

// bitset_new x bitset_new (src/set/ops/intersection.hpp:1364-1413): the blocks present
// on both sides, in id order, each AND-ed; out = [u32 ids x K][2 u64 words x K] with
// K matched blocks (empty ANDs kept); res[0] = |A & B|, res[1] = K
__global__ void __launch_bounds__(ISECT_THREADS)
k_isect_bn_bn(const uint8_t *A, uint32_t na, const uint8_t *B, uint32_t nb, uint8_t *out, unsigned long long *res) {
  __shared__ uint32_t sh[32];
  const uint32_t *Aid = reinterpret_cast<const uint32_t *>(A), *Bid = reinterpret_cast<const uint32_t *>(B);
  const uint8_t *AW = A + 4ull * na, *BW = B + 4ull * nb;
  // pass 1: K
  uint32_t K = 0;
  for (uint32_t c0 = 0; c0 < na; c0 += blockDim.x) {
    const uint32_t k = c0 + threadIdx.x;
    const uint32_t hit = (k < na && list_contains(Bid, nb, Aid[k])) ? 1u : 0u;
    uint32_t tot;
    (void)block_scan_excl(hit, sh, &tot);
    K += tot;
  }
  // pass 2: ids, words and the count
  uint32_t base = 0, count = 0;
  for (uint32_t c0 = 0; c0 < na; c0 += blockDim.x) {
    const uint32_t k = c0 + threadIdx.x;
    uint32_t hit = 0, j = 0;
    if (k < na) {
      const uint32_t x = Aid[k];
      uint32_t lo = 0, len = nb;
      while (len > 1) {
        const uint32_t half = len >> 1;
        lo = (Bid[lo + half] <= x) ? lo + half : lo;
        len -= half;
      }
      hit = (nb && Bid[lo] == x) ? 1u : 0u;
      j = lo;
    }
    uint32_t tot;
    const uint32_t pos = base + block_scan_excl(hit, sh, &tot);
    if (hit) {
      const uint64_t r0 = ld64u(AW + 16ull * k) & ld64u(BW + 16ull * j);
      const uint64_t r1 = ld64u(AW + 16ull * k + 8) & ld64u(BW + 16ull * j + 8);
      count += (uint32_t)(__popcll(r0) + __popcll(r1));
      if (out) {
        uint32_t *o = reinterpret_cast<uint32_t *>(out);
        o[pos] = Aid[k];
        uint32_t *w = reinterpret_cast<uint32_t *>(out + 4ull * K + 16ull * pos);
        w[0] = (uint32_t)r0;
        w[1] = (uint32_t)(r0 >> 32);
        w[2] = (uint32_t)r1;
        w[3] = (uint32_t)(r1 >> 32);
      }
    }
    base += tot;
  }
  uint32_t total;
  (void)block_scan_excl(count, sh, &total);
  if (threadIdx.x == 0) {
    res[0] = total;
    res[1] = K;
  }
}

// merge of three disjoint sorted lists (distinct_merge_three_way,
// src/set/ops/intersection.hpp:1511-1553): every element lands at its index plus the
// number of smaller elements in the other two lists
__device__ __forceinline__ uint32_t count_below(const uint32_t *L, uint32_t n, uint32_t x) {
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (L[lo + half] < x) { lo += half + 1; len -= half + 1; }
    else len = half;
  }
  return lo;
}

__global__ void __launch_bounds__(ISECT_THREADS)
k_merge3(const uint32_t *X, const uint32_t *Y, const uint32_t *Z, const unsigned long long *nx,
         const unsigned long long *ny, const unsigned long long *nz, uint32_t *out) {
  const uint32_t n[3] = {(uint32_t)nx[0], (uint32_t)ny[0], (uint32_t)nz[0]};
  const uint32_t *L[3] = {X, Y, Z};
  for (uint32_t q = 0; q < 3; q++)
    for (uint32_t k = threadIdx.x; k < n[q]; k += blockDim.x) {
      const uint32_t x = L[q][k];
      uint32_t pos = k;
      for (uint32_t o = 0; o < 3; o++)
        if (o != q) pos += count_below(L[o], n[o], x);
      out[pos] = x;
    }
}

// set_union(bitset A, bitset B) over their common words (union.hpp:7-57): A |= B
__global__ void __launch_bounds__(ISECT_THREADS)
k_union_words(uint8_t *A, uint64_t a_off, const uint8_t *B, uint64_t b_off, uint64_t total) {
  for (uint64_t i = threadIdx.x; i < total; i += blockDim.x) {
    uint32_t *a = reinterpret_cast<uint32_t *>(A + 8 * (a_off + i));
    const uint32_t *b = reinterpret_cast<const uint32_t *>(B + 8 * (b_off + i));
    a[0] |= b[0];
    a[1] |= b[1];
  }
}

// set_union(bitset A, uint/pshort B) (union.hpp:58-80): B's elements inside A's words set
// their bits (A: words of the image after its u64 start, first word a0, na words)
__global__ void __launch_bounds__(ISECT_THREADS)
k_union_elems(uint8_t *A, uint64_t a0, uint64_t na, const uint32_t *list, uint64_t n) {
  uint32_t *W = reinterpret_cast<uint32_t *>(A);
  for (uint64_t k = threadIdx.x; k < n; k += blockDim.x) {
    const uint32_t x = list[k];
    const uint64_t w = (uint64_t)(x >> 6) - a0;
    if ((x >> 6) >= a0 && w < na) atomicOr(&W[2 * w + ((x >> 5) & 1u)], 1u << (x & 31u));
  }
}

// set_difference(C, bitset A, bitset B) (difference.hpp:5-84): A's words AND NOT B's over
// their common span, A's other words as they are, expanded in id order; res[0] = |C|
__global__ void __launch_bounds__(ISECT_THREADS)
k_difference(const uint8_t *A, uint64_t a0, uint64_t na, const uint8_t *B, uint64_t b0, uint64_t nb, uint32_t *out,
             unsigned long long *res) {
  __shared__ uint32_t sh[32];
  uint32_t base = 0;
  for (uint64_t c0 = 0; c0 < na; c0 += blockDim.x) {
    const uint64_t i = c0 + threadIdx.x;
    uint64_t r = 0;
    if (i < na) {
      r = ld64u(A + 8 * i);
      const uint64_t w = a0 + i;
      if (w >= b0 && w - b0 < nb) r &= ~ld64u(B + 8 * (w - b0));
    }
    uint32_t tot;
    uint32_t pos = base + block_scan_excl((uint32_t)__popcll(r), sh, &tot);
    while (r) {
      if (out) out[pos] = (uint32_t)(64ull * (a0 + i) + (uint64_t)__builtin_ctzll(r));
      pos++;
      r &= r - 1;
    }
    base += tot;
  }
  if (threadIdx.x == 0) res[0] = base;
}

// bitset x bitset over the overlapping word span [start, start+total)
__global__ void __launch_bounds__(ISECT_THREADS)
k_isect_bs_bs(const uint8_t *A, uint64_t a_off, const uint8_t *B, uint64_t b_off, uint64_t total, uint8_t *out,
              unsigned long long *res) {
  __shared__ uint32_t sh[32];
  uint64_t acc = 0;
  for (uint64_t i = threadIdx.x; i < total; i += blockDim.x) {
    const uint64_t r = ld64u(A + 8 * (a_off + i)) & ld64u(B + 8 * (b_off + i));
    if (out) {
      uint32_t *o = reinterpret_cast<uint32_t *>(out + 8 + 8 * i);
      o[0] = (uint32_t)r;
      o[1] = (uint32_t)(r >> 32);
    }
    acc += (uint64_t)__popcll(r);
  }
  // counts per thread stay far below 2^32 for one set
  uint32_t tot;
  (void)block_scan_excl((uint32_t)acc, sh, &tot);
  if (threadIdx.x == 0) res[0] = tot;
}

// pshort stream of a sorted u32 list: partitions by the high 16 bits,
// [prefix, count-1, low16 x count] each; res[1] = shorts written
__global__ void __launch_bounds__(ISECT_THREADS)
k_encode_pshort(const uint32_t *R, const unsigned long long *res_in, uint16_t *out, uint32_t *pstart,
                unsigned long long *res) {
  __shared__ uint32_t sh[32];
  const uint32_t c = (uint32_t)res_in[0];
  uint32_t np = 0;
  for (uint32_t c0 = 0; c0 < c; c0 += blockDim.x) {
    const uint32_t k = c0 + threadIdx.x;
    const bool start = k < c && (k == 0 || (R[k] >> 16) != (R[k - 1] >> 16));
    uint32_t tot;
    const uint32_t p = np + block_scan_excl(start ? 1u : 0u, sh, &tot);
    // element k belongs to partition p (start) or p-1 (inside)
    if (k < c) {
      const uint32_t part = start ? p : p - 1u;
      out[2u * (part + 1u) + k] = (uint16_t)(R[k] & 0xFFFFu);
      if (start) pstart[p] = k;
    }
    np += tot;
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
    const uint32_t s = pstart[p];
    const uint32_t e = (p + 1 < np) ? pstart[p + 1] : c;
    out[2u * p + s] = (uint16_t)(R[s] >> 16);
    out[2u * p + s + 1] = (uint16_t)(e - s - 1u);
  }
  if (threadIdx.x == 0) res[1] = 2ull * np + c;
}

// one wave per pair: count of the shorter row's elements found in the longer row
__global__ void __launch_bounds__(256)
k_isect_rows(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr, const uint32_t *ra,
             const uint32_t *rb, uint64_t npairs, unsigned long long *counts) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t p = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < npairs; p += nwaves) {
    uint32_t a = ra[p], b = rb[p];
    uint32_t na = card[a], nb = card[b];
    if (na > nb) {
      const uint32_t t = a; a = b; b = t;
      const uint32_t u = na; na = nb; nb = u;
    }
    // row offsets count 16-byte units of the padded uint arena
    const uint32_t *S = unbr + (uint64_t)aoff[a] * 4, *L = unbr + (uint64_t)aoff[b] * 4;
    uint32_t hits = 0;
    for (uint32_t k = lane; k < na; k += 64) hits += list_contains(L, nb, S[k]) ? 1u : 0u;
    for (int o = 32; o; o >>= 1) hits += __shfl_xor(hits, o, 64);
    if (lane == 0) counts[p] = hits;
  }
}

// one wave per request: the elements of the request's shortest row that every other row of
// the request holds, counted (counts[p]) and, with elems, written in order from eoff[p]
__global__ void __launch_bounds__(256)
k_isect_row_sets(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr, const uint32_t *ids,
                 const uint64_t *roff, uint64_t nreq, unsigned long long *counts, const uint64_t *eoff,
                 uint32_t *elems) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t p = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < nreq; p += nwaves) {
    const uint64_t r0 = roff[p];
    const uint32_t nr = (uint32_t)(roff[p + 1] - r0);  // 1..EH_ROW_SET_MAX, checked by the host
    uint32_t rid = 0, rc = 0xFFFFFFFFu;
    if (lane < nr) {
      rid = ids[r0 + lane];
      rc = card[rid];
    }
    // the shortest row (ties: the lowest lane) as the base
    uint64_t key = ((uint64_t)rc << 8) | lane;
    for (int o = 32; o; o >>= 1) {
      const uint64_t y = __shfl_xor(key, o, 64);
      key = y < key ? y : key;
    }
    const uint32_t bl = (uint32_t)(key & 0xFFu);
    const uint32_t nb = (uint32_t)(key >> 8);
    const uint32_t *B = unbr + (uint64_t)aoff[__shfl(rid, (int)bl, 64)] * 4;
    uint64_t wpos = elems ? eoff[p] : 0;
    uint64_t hits = 0;
    for (uint32_t k0 = 0; k0 < nb; k0 += 64) {
      const uint32_t k = k0 + lane;
      bool in = k < nb;
      const uint32_t x = in ? B[k] : 0u;
      for (uint32_t q = 0; q < nr; q++) {
        const uint32_t rq = __shfl(rid, (int)q, 64), cq = __shfl(rc, (int)q, 64);
        if (q != bl && in) in = list_contains(unbr + (uint64_t)aoff[rq] * 4, cq, x);
      }
      const uint64_t mask = __ballot(in);
      if (elems && in)
        elems[wpos + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u))] = x;
      wpos += (uint64_t)__popcll(mask);
      hits += (uint64_t)__popcll(mask);
    }
    if (lane == 0) counts[p] = hits;
  }
}

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = (work + block - 1) / block;
  if (g == 0) g = 1;
  return (unsigned)(g > 65536 ? 65536 : g);
}

// a pshort stream must consist of whole partitions holding exactly card elements
bool pshort_ok(const eh_row_view *v) {
  const uint8_t *p = static_cast<const uint8_t *>(v->data);
  const uint64_t ns = v->number_of_bytes / 2;
  uint64_t i = 0, total = 0;
  while (i + 1 < ns) {
    uint16_t cnt;
    memcpy(&cnt, p + 2 * (i + 1), 2);
    total += (uint64_t)cnt + 1;
    i += 2ull + cnt + 1;
  }
  return i == ns && total == v->cardinality;
}

inline uint32_t ld32h(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
inline uint64_t ld64h(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }

// a bitset_new image: whole 20-byte blocks, ids strictly increasing below 2^25, card bits
bool bn_ok(const uint8_t *p, uint64_t bytes, uint64_t card) {
  if (bytes % 20) return false;
  const uint64_t nb = bytes / 20;
  uint64_t total = 0;
  for (uint64_t k = 0; k < nb; k++) {
    const uint32_t id = ld32h(p + 4 * k);
    if (id >= (1u << 25) || (k && id <= ld32h(p + 4 * (k - 1)))) return false;
    total += (uint64_t)__builtin_popcountll(ld64h(p + 4 * nb + 16 * k)) +
             (uint64_t)__builtin_popcountll(ld64h(p + 4 * nb + 16 * k + 8));
  }
  return total == card;
}

// a new_type image: [u64 uint bytes][strictly increasing uint part outside every block][bitset_new]
bool nt_ok(const eh_row_view *v) {
  const uint8_t *p = static_cast<const uint8_t *>(v->data);
  const uint64_t bytes = v->number_of_bytes;
  if (bytes < 8) return false;
  const uint64_t ub = ld64h(p);
  if (ub % 4 || ub > bytes - 8 || ub / 4 > v->cardinality) return false;
  const uint64_t nu = ub / 4, bb = bytes - 8 - ub;
  const uint8_t *B = p + 8 + ub;
  if (!bn_ok(B, bb, v->cardinality - nu)) return false;
  const uint64_t nb = bb / 20;
  for (uint64_t k = 0; k < nu; k++) {
    const uint32_t x = ld32h(p + 8 + 4 * k);
    if (k && x <= ld32h(p + 8 + 4 * (k - 1))) return false;
    for (uint64_t lo = 0, hi = nb; lo < hi;) {  // x's block must not be one of the blocks
      const uint64_t mid = (lo + hi) / 2;
      const uint32_t id = ld32h(B + 4 * mid);
      if (id == (x >> 7)) return false;
      if (id < (x >> 7)) lo = mid + 1; else hi = mid;
    }
  }
  return true;
}

bool view_ok(const eh_row_view *v) {
  if (v->cardinality == 0) return true;
  if (!v->data || v->cardinality > 0xFFFFFFFFull) return false;
  if (v->type == EH_U) return true;
  if (v->type == EH_BS) return v->number_of_bytes >= 16 && (v->number_of_bytes - 8) % 8 == 0;
  if (v->type == EH_PS) return v->number_of_bytes >= 6 && v->number_of_bytes % 2 == 0 && pshort_ok(v);
  if (v->type == EH_SET_BITSET_NEW)
    return v->number_of_bytes >= 20 && bn_ok(static_cast<const uint8_t *>(v->data), v->number_of_bytes, v->cardinality);
  if (v->type == EH_SET_NEW_TYPE) return nt_ok(v);
  return false;
}

bool is_block_set(int t) { return t == EH_SET_BITSET_NEW || t == EH_SET_NEW_TYPE; }

uint64_t view_bytes(const eh_row_view *v) {
  return v->type == EH_U ? 4ull * v->cardinality : v->number_of_bytes;
}

// S & L of two device lists (S the shorter one keeps its order): R (capacity |S|), *res = count
void isect_lists(hipStream_t st, const uint32_t *x, uint32_t nx, const uint32_t *y, uint32_t ny, uint32_t *R,
                 unsigned long long *res) {
  const bool xs = nx <= ny;
  k_isect_lists<<<1, ISECT_THREADS, 0, st>>>(xs ? x : y, xs ? nx : ny, xs ? y : x, xs ? ny : nx, R, res);
}

// the block pairs the reference overloads: bitset_new x bitset_new
// (intersection.hpp:1364-1413), uint x bitset_new (:1476-1509), new_type x new_type
// (:1555-1649)
eh_status intersect_blocks(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, int materialize, void *out_buf,
                           size_t out_cap, eh_row_view *out) {
  hipStream_t st = ctx->stream;
  const int ta = a->type, tb = b->type;
  const bool bnbn = ta == EH_SET_BITSET_NEW && tb == EH_SET_BITSET_NEW;
  const bool ubn = (ta == EH_U && tb == EH_SET_BITSET_NEW) || (ta == EH_SET_BITSET_NEW && tb == EH_U);
  const bool ntnt = ta == EH_SET_NEW_TYPE && tb == EH_SET_NEW_TYPE;
  if (!bnbn && !ubn && !ntnt) {
    eh_set_error("no set_intersect overload for set types %d x %d", ta, tb);
    return EH_ERR_INVALID;
  }
  const int rt = bnbn ? EH_SET_BITSET_NEW : ntnt ? EH_SET_NEW_TYPE : EH_U;
  if (out) {
    out->data = out_buf;
    out->cardinality = 0;
    out->number_of_bytes = 0;
    out->type = rt;
    out->reserved = 0;
  }
  // these overloads return an empty uint set when either side has no bytes
  if (a->cardinality == 0 || b->cardinality == 0) {
    if (out) out->type = EH_U;
    return EH_OK;
  }
  const uint64_t ba = view_bytes(a), bb = view_bytes(b);
  DevBuf pa, pb, res;
  EH_CHECK(pa.alloc(ba + 16));
  EH_CHECK(pb.alloc(bb + 16));
  EH_CHECK(res.alloc(128));
  EH_HIP(hipMemsetAsync(res.p, 0, 128, st));
  EH_HIP(hipMemcpyAsync(pa.p, a->data, ba, hipMemcpyHostToDevice, st));
  EH_HIP(hipMemcpyAsync(pb.p, b->data, bb, hipMemcpyHostToDevice, st));
  unsigned long long *R = res.as<unsigned long long>();
  unsigned long long h[8] = {0, 0, 0, 0, 0, 0, 0, 0};

  if (bnbn) {
    const uint32_t na = (uint32_t)(ba / 20), nb = (uint32_t)(bb / 20);
    DevBuf o;
    if (materialize) EH_CHECK(o.alloc(20ull * (na < nb ? na : nb) + 16));
    k_isect_bn_bn<<<1, ISECT_THREADS, 0, st>>>(pa.as<uint8_t>(), na, pb.as<uint8_t>(), nb,
                                               materialize ? o.as<uint8_t>() : nullptr, R);
    EH_HIP(hipGetLastError());
    EH_HIP(hipMemcpyAsync(h, R, 16, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (out) out->cardinality = h[0];
    if (!materialize) return EH_OK;
    const uint64_t need = 20ull * h[1];
    if (need > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)need); return EH_ERR_CAPACITY; }
    if (need) EH_HIP(hipMemcpy(out_buf, o.p, need, hipMemcpyDeviceToHost));
    out->number_of_bytes = need;
    return EH_OK;
  }

  if (ubn) {
    const bool a_bn = ta == EH_SET_BITSET_NEW;
    const uint64_t nbn = a_bn ? a->cardinality : b->cardinality, nu = a_bn ? b->cardinality : a->cardinality;
    DevBuf lbn, r;
    EH_CHECK(lbn.alloc(4 * nbn + 16));
    EH_CHECK(r.alloc(4 * (nu < nbn ? nu : nbn) + 16));
    k_decode_set<<<1, ISECT_THREADS, 0, st>>>(EH_SET_BITSET_NEW, a_bn ? pa.as<uint8_t>() : pb.as<uint8_t>(),
                                              a_bn ? ba : bb, nbn, lbn.as<uint32_t>(), nullptr);
    EH_HIP(hipGetLastError());
    isect_lists(st, a_bn ? pb.as<uint32_t>() : pa.as<uint32_t>(), (uint32_t)nu, lbn.as<uint32_t>(), (uint32_t)nbn,
                r.as<uint32_t>(), R);
    EH_HIP(hipGetLastError());
    EH_HIP(hipMemcpyAsync(h, R, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (out) out->cardinality = h[0];
    if (!materialize) return EH_OK;
    const uint64_t need = 4ull * h[0];
    if (need > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)need); return EH_ERR_CAPACITY; }
    if (need) EH_HIP(hipMemcpy(out_buf, r.p, need, hipMemcpyDeviceToHost));
    out->number_of_bytes = need;
    return EH_OK;
  }

  // new_type x new_type: [u64 uint bytes][uint part][bitset_new part] on both sides
  const uint8_t *ha = static_cast<const uint8_t *>(a->data), *hb = static_cast<const uint8_t *>(b->data);
  const uint64_t uba = ld64h(ha), ubb = ld64h(hb);
  const uint32_t nua = (uint32_t)(uba / 4), nub = (uint32_t)(ubb / 4);
  const uint64_t bba = ba - 8 - uba, bbb = bb - 8 - ubb;
  const uint32_t nba = (uint32_t)(bba / 20), nbb = (uint32_t)(bbb / 20);
  const uint64_t cba = a->cardinality - nua, cbb = b->cardinality - nub;  // elements in blocks
  const uint32_t *AI = reinterpret_cast<const uint32_t *>(pa.as<uint8_t>() + 8);
  const uint32_t *BI = reinterpret_cast<const uint32_t *>(pb.as<uint8_t>() + 8);
  const uint8_t *ABS = pa.as<uint8_t>() + 8 + uba, *BBS = pb.as<uint8_t>() + 8 + ubb;
  const bool all = (uba || ubb) && (bba || bbb);
  const bool uint_only = !bba && !bbb;
  DevBuf r1, r2, r3, da, db, u, o;
  EH_CHECK(r1.alloc(4ull * (nua < nub ? nua : nub) + 16));
  EH_CHECK(r2.alloc(4ull * nua + 16));
  EH_CHECK(r3.alloc(4ull * nub + 16));
  EH_CHECK(u.alloc(4ull * (nua + nub) + 16));
  isect_lists(st, AI, nua, BI, nub, r1.as<uint32_t>(), R + 0);
  EH_HIP(hipGetLastError());
  if (all) {
    // A's uint part against B's blocks, B's uint part against A's blocks
    if (nua && bbb) {
      EH_CHECK(db.alloc(4 * cbb + 16));
      k_decode_set<<<1, ISECT_THREADS, 0, st>>>(EH_SET_BITSET_NEW, BBS, bbb, cbb, db.as<uint32_t>(), nullptr);
      isect_lists(st, AI, nua, db.as<uint32_t>(), (uint32_t)cbb, r2.as<uint32_t>(), R + 2);
    }
    if (nub && bba) {
      EH_CHECK(da.alloc(4 * cba + 16));
      k_decode_set<<<1, ISECT_THREADS, 0, st>>>(EH_SET_BITSET_NEW, ABS, bba, cba, da.as<uint32_t>(), nullptr);
      isect_lists(st, BI, nub, da.as<uint32_t>(), (uint32_t)cba, r3.as<uint32_t>(), R + 4);
    }
    EH_HIP(hipGetLastError());
    if (materialize)
      k_merge3<<<1, ISECT_THREADS, 0, st>>>(r1.as<uint32_t>(), r2.as<uint32_t>(), r3.as<uint32_t>(), R + 0, R + 2,
                                            R + 4, u.as<uint32_t>());
    EH_HIP(hipGetLastError());
  }
  if (!uint_only && bba && bbb) {
    if (materialize) EH_CHECK(o.alloc(20ull * (nba < nbb ? nba : nbb) + 16));
    k_isect_bn_bn<<<1, ISECT_THREADS, 0, st>>>(ABS, nba, BBS, nbb, materialize ? o.as<uint8_t>() : nullptr, R + 6);
    EH_HIP(hipGetLastError());
  }
  EH_HIP(hipMemcpyAsync(h, R, 64, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  const uint64_t nu = h[0] + h[2] + h[4];
  if (out) out->cardinality = nu + h[6];
  if (!materialize) return EH_OK;
  const uint64_t need = 8 + 4 * nu + 20 * h[7];
  if (need > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)need); return EH_ERR_CAPACITY; }
  uint8_t *ob = static_cast<uint8_t *>(out_buf);
  const uint64_t hdr = 4 * nu;
  memcpy(ob, &hdr, 8);
  if (nu) EH_HIP(hipMemcpy(ob + 8, all ? u.p : r1.p, 4 * nu, hipMemcpyDeviceToHost));
  if (h[7]) EH_HIP(hipMemcpy(ob + 8 + 4 * nu, o.p, 20 * h[7], hipMemcpyDeviceToHost));
  out->number_of_bytes = need;
  return EH_OK;
}

}  // namespace

extern "C" eh_status eh_intersect(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, int materialize,
                                  void *out_buf, size_t out_cap, eh_row_view *out) {
  if (!ctx || !a || !b || (materialize && (!out_buf || !out))) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (!view_ok(a) || !view_ok(b)) { eh_set_error("malformed set view (type or byte count)"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(ctx->device));
  if (is_block_set(a->type) || is_block_set(b->type)) return intersect_blocks(ctx, a, b, materialize, out_buf, out_cap, out);
  hipStream_t st = ctx->stream;
  const int ta = a->type, tb = b->type;
  int rt = EH_U;
  if (ta == EH_BS && tb == EH_BS) rt = EH_BS;
  else if ((ta == EH_PS && tb == EH_PS) || (ta == EH_PS && tb == EH_BS) || (ta == EH_BS && tb == EH_PS)) rt = EH_PS;
  if (out) {
    out->data = out_buf;
    out->cardinality = 0;
    out->number_of_bytes = 0;
    out->type = rt;
    out->reserved = 0;
  }
  // the reference returns an empty result as soon as either side is empty
  // (src/set/ops/intersection.hpp:2068-2072)
  if (a->cardinality == 0 || b->cardinality == 0) return EH_OK;

  const uint64_t ba = view_bytes(a), bb = view_bytes(b);
  DevBuf pa, pb, res;
  EH_CHECK(pa.alloc(ba + 16));
  EH_CHECK(pb.alloc(bb + 16));
  EH_CHECK(res.alloc(64));
  EH_HIP(hipMemcpyAsync(pa.p, a->data, ba, hipMemcpyHostToDevice, st));
  EH_HIP(hipMemcpyAsync(pb.p, b->data, bb, hipMemcpyHostToDevice, st));
  unsigned long long hres[2] = {0ull, 0ull};

  if (rt == EH_BS) {
    uint64_t a0 = 0, b0 = 0;
    memcpy(&a0, a->data, 8);
    memcpy(&b0, b->data, 8);
    const uint64_t sa = (a->number_of_bytes - 8) / 8, sb = (b->number_of_bytes - 8) / 8;
    const uint64_t start = a0 > b0 ? a0 : b0;
    const uint64_t end = (a0 + sa) < (b0 + sb) ? (a0 + sa) : (b0 + sb);
    const uint64_t total = start > end ? 0 : end - start;
    const uint64_t need = 8 + 8 * total;
    DevBuf o;
    if (materialize) {
      if (need > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)need); return EH_ERR_CAPACITY; }
      EH_CHECK(o.alloc(need + 16));
    }
    k_isect_bs_bs<<<1, ISECT_THREADS, 0, st>>>(pa.as<uint8_t>() + 8, total ? start - a0 : 0, pb.as<uint8_t>() + 8,
                                               total ? start - b0 : 0, total, materialize ? o.as<uint8_t>() : nullptr,
                                               res.as<unsigned long long>());
    EH_HIP(hipGetLastError());
    EH_HIP(hipMemcpyAsync(hres, res.p, 8, hipMemcpyDeviceToHost, st));
    if (materialize && total) EH_HIP(hipMemcpyAsync((uint8_t *)out_buf + 8, o.as<uint8_t>() + 8, 8 * total, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (materialize) {
      memcpy(out_buf, &start, 8);
      out->number_of_bytes = need;
    }
    if (out) out->cardinality = hres[0];
    return EH_OK;
  }

  // decode both sides to sorted u32 lists
  const uint64_t na = a->cardinality, nb = b->cardinality;
  DevBuf la, lb, ps;
  EH_CHECK(la.alloc(4 * na + 16));
  EH_CHECK(lb.alloc(4 * nb + 16));
  const uint64_t max_parts = (ta == EH_PS ? ba / 6 : 0) + (tb == EH_PS ? bb / 6 : 0) + 2;
  EH_CHECK(ps.alloc(4 * (3 * max_parts + 2)));
  k_decode_set<<<1, ISECT_THREADS, 0, st>>>(ta, pa.as<uint8_t>(), ba, na, la.as<uint32_t>(), ps.as<uint32_t>());
  EH_HIP(hipGetLastError());
  k_decode_set<<<1, ISECT_THREADS, 0, st>>>(tb, pb.as<uint8_t>(), bb, nb, lb.as<uint32_t>(), ps.as<uint32_t>());
  EH_HIP(hipGetLastError());
  const bool a_short = na <= nb;
  const uint64_t ns = a_short ? na : nb;
  DevBuf r;
  EH_CHECK(r.alloc(4 * ns + 16));
  k_isect_lists<<<1, ISECT_THREADS, 0, st>>>(a_short ? la.as<uint32_t>() : lb.as<uint32_t>(), (uint32_t)ns,
                                             a_short ? lb.as<uint32_t>() : la.as<uint32_t>(),
                                             (uint32_t)(a_short ? nb : na), r.as<uint32_t>(),
                                             res.as<unsigned long long>());
  EH_HIP(hipGetLastError());
  if (!materialize) {
    EH_HIP(hipMemcpyAsync(hres, res.p, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    if (out) out->cardinality = hres[0];
    return EH_OK;
  }
  if (rt == EH_U) {
    EH_HIP(hipMemcpyAsync(hres, res.p, 8, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
    const uint64_t need = 4ull * hres[0];
    out->cardinality = hres[0];
    if (need > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)need); return EH_ERR_CAPACITY; }
    if (need) EH_HIP(hipMemcpy(out_buf, r.p, need, hipMemcpyDeviceToHost));
    out->number_of_bytes = need;
    return EH_OK;
  }
  // pshort result: at most 3 shorts per element
  DevBuf o, pst;
  EH_CHECK(o.alloc(6 * ns + 16));
  EH_CHECK(pst.alloc(4 * ns + 16));
  k_encode_pshort<<<1, ISECT_THREADS, 0, st>>>(r.as<uint32_t>(), res.as<unsigned long long>(), o.as<uint16_t>(),
                                               pst.as<uint32_t>(), res.as<unsigned long long>());
  EH_HIP(hipGetLastError());
  EH_HIP(hipMemcpyAsync(hres, res.p, 16, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  const uint64_t need = 2ull * hres[1];
  out->cardinality = hres[0];
  if (need > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)need); return EH_ERR_CAPACITY; }
  if (need) EH_HIP(hipMemcpy(out_buf, o.p, need, hipMemcpyDeviceToHost));
  out->number_of_bytes = need;
  return EH_OK;
}

extern "C" eh_status eh_union(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, void *out_buf, size_t out_cap,
                              eh_row_view *out) {
  if (!ctx || !a || !b || !out || !out_buf) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (a->type != EH_BS || a->number_of_bytes < 16 || (a->number_of_bytes - 8) % 8 || !a->data) {
    eh_set_error("eh_union: A must be a non-empty bitset image");
    return EH_ERR_INVALID;
  }
  if (b->type != EH_BS && b->type != EH_U && b->type != EH_PS) {
    eh_set_error("eh_union: no set_union overload for B of type %d", b->type);
    return EH_ERR_INVALID;
  }
  // a bitset B is OR-ed word by word whatever its cardinality says (the reference's bitset x
  // bitset union reads only the words, union.hpp:6-45, and the sets set_union fills keep the
  // cardinality they had); a uint or pshort B is walked element by element
  const bool b_words = b->type == EH_BS && b->number_of_bytes >= 16;
  if (b_words && ((b->number_of_bytes - 8) % 8 || !b->data)) {
    eh_set_error("malformed set view (bitset byte count)");
    return EH_ERR_INVALID;
  }
  if (!b_words && b->cardinality && !view_ok(b)) {
    eh_set_error("malformed set view (type or byte count)");
    return EH_ERR_INVALID;
  }
  const uint64_t ba = a->number_of_bytes;
  if (ba > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)ba); return EH_ERR_CAPACITY; }
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  uint64_t a0 = 0;
  memcpy(&a0, a->data, 8);
  const uint64_t na = (ba - 8) / 8;
  DevBuf pa, pb, lb, ps, res;
  EH_CHECK(pa.alloc(ba + 16));
  EH_CHECK(res.alloc(64));
  EH_HIP(hipMemcpyAsync(pa.p, a->data, ba, hipMemcpyHostToDevice, st));
  if (b_words || (b->type != EH_BS && b->cardinality)) {
    const uint64_t bb = view_bytes(b);
    EH_CHECK(pb.alloc(bb + 16));
    EH_HIP(hipMemcpyAsync(pb.p, b->data, bb, hipMemcpyHostToDevice, st));
    if (b->type == EH_BS) {
      uint64_t b0 = 0;
      memcpy(&b0, b->data, 8);
      const uint64_t nbw = (bb - 8) / 8;
      const uint64_t start = a0 > b0 ? a0 : b0, end = (a0 + na) < (b0 + nbw) ? (a0 + na) : (b0 + nbw);
      if (start < end)
        k_union_words<<<1, ISECT_THREADS, 0, st>>>(pa.as<uint8_t>() + 8, start - a0, pb.as<uint8_t>() + 8, start - b0,
                                                   end - start);
    } else {
      EH_CHECK(lb.alloc(4 * b->cardinality + 16));
      EH_CHECK(ps.alloc(4 * (3 * (bb / 6 + 2) + 2)));
      k_decode_set<<<1, ISECT_THREADS, 0, st>>>(b->type, pb.as<uint8_t>(), bb, b->cardinality, lb.as<uint32_t>(),
                                                ps.as<uint32_t>());
      k_union_elems<<<1, ISECT_THREADS, 0, st>>>(pa.as<uint8_t>() + 8, a0, na, lb.as<uint32_t>(), b->cardinality);
    }
    EH_HIP(hipGetLastError());
  }
  // the result's element count: A \ {} over A's own words
  k_difference<<<1, ISECT_THREADS, 0, st>>>(pa.as<uint8_t>() + 8, a0, na, nullptr, 0, 0, nullptr,
                                            res.as<unsigned long long>());
  EH_HIP(hipGetLastError());
  unsigned long long c = 0;
  EH_HIP(hipMemcpyAsync(&c, res.p, 8, hipMemcpyDeviceToHost, st));
  EH_HIP(hipMemcpyAsync(out_buf, pa.p, ba, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  out->data = out_buf;
  out->cardinality = c;
  out->number_of_bytes = ba;
  out->type = EH_BS;
  out->reserved = 0;
  return EH_OK;
}

extern "C" eh_status eh_difference(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, void *out_buf,
                                   size_t out_cap, eh_row_view *out) {
  if (!ctx || !a || !b || !out || (out_cap && !out_buf)) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  for (const eh_row_view *v : {a, b})
    if (v->type != EH_BS || (v->number_of_bytes && (v->number_of_bytes < 16 || (v->number_of_bytes - 8) % 8 || !v->data))) {
      eh_set_error("eh_difference: both sets must be bitset images");
      return EH_ERR_INVALID;
    }
  out->data = out_buf;
  out->cardinality = 0;
  out->number_of_bytes = 0;
  out->type = EH_U;
  out->reserved = 0;
  // the reference leaves C untouched unless both sides have bytes (difference.hpp:7)
  if (a->number_of_bytes == 0 || b->number_of_bytes == 0) return EH_OK;
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  uint64_t a0 = 0, b0 = 0;
  memcpy(&a0, a->data, 8);
  memcpy(&b0, b->data, 8);
  const uint64_t na = (a->number_of_bytes - 8) / 8, nb = (b->number_of_bytes - 8) / 8;
  DevBuf pa, pb, o, res;
  EH_CHECK(pa.alloc(a->number_of_bytes + 16));
  EH_CHECK(pb.alloc(b->number_of_bytes + 16));
  EH_CHECK(o.alloc(256 * na + 16));
  EH_CHECK(res.alloc(64));
  EH_HIP(hipMemcpyAsync(pa.p, a->data, a->number_of_bytes, hipMemcpyHostToDevice, st));
  EH_HIP(hipMemcpyAsync(pb.p, b->data, b->number_of_bytes, hipMemcpyHostToDevice, st));
  k_difference<<<1, ISECT_THREADS, 0, st>>>(pa.as<uint8_t>() + 8, a0, na, pb.as<uint8_t>() + 8, b0, nb,
                                            o.as<uint32_t>(), res.as<unsigned long long>());
  EH_HIP(hipGetLastError());
  unsigned long long c = 0;
  EH_HIP(hipMemcpyAsync(&c, res.p, 8, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  out->cardinality = c;
  if (4 * c > out_cap) { eh_set_error("result needs %llu bytes", (unsigned long long)(4 * c)); return EH_ERR_CAPACITY; }
  if (c) EH_HIP(hipMemcpy(out_buf, o.p, 4 * c, hipMemcpyDeviceToHost));
  out->number_of_bytes = 4 * c;
  return EH_OK;
}

namespace {

// Staging of eh_intersect_rows_static: the pairs' row ids and their counts live in this
// file-scope array instead of buffers allocated per call -- EH_STATIC_MAX_PAIRS u32 a-rows,
// then as many u32 b-rows, then as many u64 counts (16 MiB).  Each GPU the library runs on
// holds its own copy until the process ends.  Host code never uses the array's name (it is
// a host-side stand-in, not a device address): eh_ctx_create calls static_scratch_init with
// the context's device current, which keeps that device's address in g_static_ptr.
static __device__ __align__(256) uint8_t scratch[16ull * EH_STATIC_MAX_PAIRS];

constexpr int STATIC_MAX_DEVICES = 64;
std::mutex g_static_mu;
void *g_static_ptr[STATIC_MAX_DEVICES] = {};

// the pairs (device ra, rb) counted into c; a_rows / b_rows / counts are host arrays
eh_status isect_rows_run(eh_matrix *m, const uint32_t *a_rows, const uint32_t *b_rows, uint64_t npairs,
                         uint64_t *counts, uint32_t *ra, uint32_t *rb, unsigned long long *c) {
  hipStream_t st = m->ctx->stream;
  EH_HIP(hipMemcpyAsync(ra, a_rows, 4 * npairs, hipMemcpyHostToDevice, st));
  EH_HIP(hipMemcpyAsync(rb, b_rows, 4 * npairs, hipMemcpyHostToDevice, st));
  k_isect_rows<<<grid_for(npairs * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                      m->dunbr.as<uint32_t>(), ra, rb, npairs, c);
  EH_HIP(hipGetLastError());
  EH_HIP(hipMemcpyAsync(counts, c, 8 * npairs, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  return EH_OK;
}

eh_status isect_rows_check(const eh_matrix *m, const uint32_t *a_rows, const uint32_t *b_rows, uint64_t npairs,
                           const uint64_t *counts) {
  if (!m || (npairs && (!a_rows || !b_rows || !counts))) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  for (uint64_t p = 0; p < npairs; p++)
    if (a_rows[p] >= m->n || b_rows[p] >= m->n) {
      eh_set_error("pair %llu names a row outside [0, %u)", (unsigned long long)p, m->n);
      return EH_ERR_INVALID;
    }
  return EH_OK;
}

}  // namespace

// Called once per context with `device` current: records that device's address of the
// static staging array (hipGetSymbolAddress), once per device.
eh_status static_scratch_init(int device) {
  if (device < 0 || device >= STATIC_MAX_DEVICES) {
    eh_set_error("device %d: the static staging table holds %d devices", device, STATIC_MAX_DEVICES);
    return EH_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (g_static_ptr[device]) return EH_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    eh_set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return EH_ERR_HIP;
  }
  g_static_ptr[device] = p;
  return EH_OK;
}

extern "C" eh_status eh_intersect_rows(eh_matrix *m, const uint32_t *a_rows, const uint32_t *b_rows,
                                       uint64_t npairs, uint64_t *counts) {
  EH_CHECK(isect_rows_check(m, a_rows, b_rows, npairs, counts));
  if (npairs == 0) return EH_OK;
  EH_HIP(hipSetDevice(m->ctx->device));
  DevBuf ra, rb, c;
  EH_CHECK(ra.alloc(4 * npairs + 16));
  EH_CHECK(rb.alloc(4 * npairs + 16));
  EH_CHECK(c.alloc(8 * npairs + 16));
  return isect_rows_run(m, a_rows, b_rows, npairs, counts, ra.as<uint32_t>(), rb.as<uint32_t>(),
                        c.as<unsigned long long>());
}

extern "C" eh_status eh_intersect_row_sets(eh_matrix *m, const uint32_t *row_ids, const uint64_t *row_off,
                                           uint64_t nreq, uint64_t *counts, uint32_t *elems, uint64_t elems_cap,
                                           uint64_t *elem_off) {
  if (!m || (nreq && (!row_ids || !row_off || !counts))) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (nreq == 0) {
    if (elem_off) elem_off[0] = 0;
    return EH_OK;
  }
  if (row_off[0] != 0) { eh_set_error("row_off[0] must be 0"); return EH_ERR_INVALID; }
  for (uint64_t q = 0; q < nreq; q++) {
    const uint64_t nr = row_off[q + 1] - row_off[q];
    if (row_off[q + 1] < row_off[q] || nr < 1 || nr > EH_ROW_SET_MAX) {
      eh_set_error("request %llu names %lld rows (1 to %d)", (unsigned long long)q,
                   (long long)(row_off[q + 1] - row_off[q]), EH_ROW_SET_MAX);
      return EH_ERR_INVALID;
    }
  }
  const uint64_t nids = row_off[nreq];
  for (uint64_t t = 0; t < nids; t++)
    if (row_ids[t] >= m->n) { eh_set_error("row %u outside [0, %u)", row_ids[t], m->n); return EH_ERR_INVALID; }
  eh_ctx *ctx = m->ctx;
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  DevBuf ids, roff, cnt, eoff, out;
  EH_CHECK(ids.alloc(4 * nids + 16));
  EH_CHECK(roff.alloc(8 * (nreq + 1)));
  EH_CHECK(cnt.alloc(8 * nreq));
  EH_HIP(hipMemcpyAsync(ids.p, row_ids, 4 * nids, hipMemcpyHostToDevice, st));
  EH_HIP(hipMemcpyAsync(roff.p, row_off, 8 * (nreq + 1), hipMemcpyHostToDevice, st));
  k_isect_row_sets<<<grid_for(nreq * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                          m->dunbr.as<uint32_t>(), ids.as<uint32_t>(),
                                                          roff.as<uint64_t>(), nreq, cnt.as<unsigned long long>(),
                                                          nullptr, nullptr);
  EH_HIP(hipGetLastError());
  EH_HIP(hipMemcpyAsync(counts, cnt.p, 8 * nreq, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  std::vector<uint64_t> off(nreq + 1);
  off[0] = 0;
  for (uint64_t q = 0; q < nreq; q++) off[q + 1] = off[q] + counts[q];
  if (elem_off) memcpy(elem_off, off.data(), 8 * (nreq + 1));
  if (!elems) return EH_OK;
  if (off[nreq] > elems_cap) {
    eh_set_error("the results hold %llu elements (capacity %llu)", (unsigned long long)off[nreq],
                 (unsigned long long)elems_cap);
    return EH_ERR_CAPACITY;
  }
  if (off[nreq] == 0) return EH_OK;
  EH_CHECK(eoff.alloc(8 * (nreq + 1)));
  EH_CHECK(out.alloc(4 * off[nreq] + 16));
  EH_HIP(hipMemcpyAsync(eoff.p, off.data(), 8 * (nreq + 1), hipMemcpyHostToDevice, st));
  k_isect_row_sets<<<grid_for(nreq * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                          m->dunbr.as<uint32_t>(), ids.as<uint32_t>(),
                                                          roff.as<uint64_t>(), nreq, cnt.as<unsigned long long>(),
                                                          eoff.as<uint64_t>(), out.as<uint32_t>());
  EH_HIP(hipGetLastError());
  EH_HIP(hipMemcpyAsync(elems, out.p, 4 * off[nreq], hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  return EH_OK;
}

extern "C" eh_status eh_intersect_rows_static(eh_matrix *m, const uint32_t *a_rows, const uint32_t *b_rows,
                                              uint64_t npairs, uint64_t *counts) {
  EH_CHECK(isect_rows_check(m, a_rows, b_rows, npairs, counts));
  if (npairs > EH_STATIC_MAX_PAIRS) {
    eh_set_error("eh_intersect_rows_static: %llu pairs exceed the static staging of %llu pairs "
                 "(split the call, or use eh_intersect_rows)",
                 (unsigned long long)npairs, (unsigned long long)EH_STATIC_MAX_PAIRS);
    return EH_ERR_CAPACITY;
  }
  if (npairs == 0) return EH_OK;
  const int dev = m->ctx->device;
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_static_mu);
    base = (dev >= 0 && dev < STATIC_MAX_DEVICES) ? g_static_ptr[dev] : nullptr;
  }
  if (!base) { eh_set_error("static staging of device %d not initialised", dev); return EH_ERR_HIP; }
  EH_HIP(hipSetDevice(dev));
  // every count is written by the kernel, so nothing in the array needs clearing
  uint32_t *ra = static_cast<uint32_t *>(base);
  uint32_t *rb = ra + EH_STATIC_MAX_PAIRS;
  unsigned long long *c = reinterpret_cast<unsigned long long *>(rb + EH_STATIC_MAX_PAIRS);
  return isect_rows_run(m, a_rows, b_rows, npairs, counts, ra, rb, c);
}
