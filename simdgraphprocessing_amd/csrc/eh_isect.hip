// eh_isect.hip -- the set-level operation ops::set_intersect(Set<X>*, const Set<Y>*,
// const Set<Z>*) (src/set/ops/intersection.hpp:829-1336, 2066-2118) on the device, and a
// batched form over rows of one matrix.
//
// eh_intersect takes two host sets in the reference's byte formats
//   uint   : u32[card]
//   bitset : [u64 first word][u64 words...]         (src/set/layouts/bitset.hpp:65-93)
//   pshort : u16 stream [prefix, count-1, low16...]  (src/set/layouts/pshort.hpp:46-78)
// uploads them, decodes both to sorted u32 lists, intersects, and (when asked)
// encodes the result in the reference's result layout: uint for any pair with a
// uint side, the AND over the overlapping word span for bitset x bitset, a
// pshort stream (empty partitions dropped) for pshort x pshort and pshort x bitset.
// Each call is one small chain of single-workgroup kernels; it serves wrappers and
// tests, not the counting pass.
//
// eh_intersect_rows counts |row(a[p]) & row(b[p])| for many pairs of one matrix with
// one wave per pair over the rows' sorted u32 lists already resident in HBM.
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

bool view_ok(const eh_row_view *v) {
  if (v->cardinality == 0) return true;
  if (!v->data || v->cardinality > 0xFFFFFFFFull) return false;
  if (v->type == EH_U) return true;
  if (v->type == EH_BS) return v->number_of_bytes >= 16 && (v->number_of_bytes - 8) % 8 == 0;
  if (v->type == EH_PS) return v->number_of_bytes >= 6 && v->number_of_bytes % 2 == 0 && pshort_ok(v);
  return false;
}

uint64_t view_bytes(const eh_row_view *v) {
  return v->type == EH_U ? 4ull * v->cardinality : v->number_of_bytes;
}

}  // namespace

extern "C" eh_status eh_intersect(eh_ctx *ctx, const eh_row_view *a, const eh_row_view *b, int materialize,
                                  void *out_buf, size_t out_cap, eh_row_view *out) {
  if (!ctx || !a || !b || (materialize && (!out_buf || !out))) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (!view_ok(a) || !view_ok(b)) { eh_set_error("malformed set view (type or byte count)"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(ctx->device));
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

extern "C" eh_status eh_intersect_rows(eh_matrix *m, const uint32_t *a_rows, const uint32_t *b_rows,
                                       uint64_t npairs, uint64_t *counts) {
  if (!m || (npairs && (!a_rows || !b_rows || !counts))) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (npairs == 0) return EH_OK;
  for (uint64_t p = 0; p < npairs; p++)
    if (a_rows[p] >= m->n || b_rows[p] >= m->n) {
      eh_set_error("pair %llu names a row outside [0, %u)", (unsigned long long)p, m->n);
      return EH_ERR_INVALID;
    }
  eh_ctx *ctx = m->ctx;
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  DevBuf ra, rb, c;
  EH_CHECK(ra.alloc(4 * npairs + 16));
  EH_CHECK(rb.alloc(4 * npairs + 16));
  EH_CHECK(c.alloc(8 * npairs + 16));
  EH_HIP(hipMemcpyAsync(ra.p, a_rows, 4 * npairs, hipMemcpyHostToDevice, st));
  EH_HIP(hipMemcpyAsync(rb.p, b_rows, 4 * npairs, hipMemcpyHostToDevice, st));
  k_isect_rows<<<grid_for(npairs * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                      m->dunbr.as<uint32_t>(), ra.as<uint32_t>(), rb.as<uint32_t>(),
                                                      npairs, c.as<unsigned long long>());
  EH_HIP(hipGetLastError());
  EH_HIP(hipMemcpyAsync(counts, c.p, 8 * npairs, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  return EH_OK;
}
