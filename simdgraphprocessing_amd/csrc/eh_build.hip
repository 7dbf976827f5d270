// eh_build.hip -- device layout builder (SparseMatrix::build on the GPU).
//
// Restates SparseMatrix<T,R>::build -> from_symmetric_noattribute_graph -> pack_data
// (src/SparseMatrix.hpp:249-266,268-290,406-494) with the per-row layout choice of
// hybrid::get_type (src/set/layouts/hybrid.hpp:73-100) and the layout encoders
// uinteger/pshort/bitset (uinteger.hpp:49-59, pshort.hpp:46-91, bitset.hpp:65-109).
// Passes, all on the device:
//   1. filter count   card[i] = kept entries of row i (all / nbr < node / mask)
//   2. pack           rows copied into one 16-byte-aligned uint32 arena
//   3. select         per-row type from (card, first, last, prefix partitions)
//   4. encode         bitset words and pshort streams for the rows that use them
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

// one wave per row: copy the kept entries, pad with the sentinel
__global__ void __launch_bounds__(256) k_pack(const uint64_t *off, const uint32_t *nbr, uint32_t n, int filter,
                                              const uint8_t *mask, const uint32_t *card, const uint64_t *poff,
                                              uint32_t *unbr) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    const uint64_t b = off[i], e = off[i + 1];
    const uint32_t c = card[i];
    uint32_t *dst = unbr + poff[i];
    if (filter != EH_FILTER_MASK) {
      for (uint32_t k = lane; k < c; k += 64) dst[k] = nbr[b + k];
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
    const uint32_t padded = (c + 3u) & ~3u;
    for (uint32_t k = c + lane; k < padded; k += 64) dst[k] = EH_PAD_VALUE;
  }
}

struct TypeArgs {
  const uint32_t *card;
  const uint64_t *poff;
  const uint32_t *unbr;
  const uint8_t *forced;  // learned per-row types or null
  uint32_t n;
  int policy;
  uint8_t *type;
  uint32_t *parts;
  uint32_t *bs_n;   // words per BS row
  uint32_t *ps_n;   // shorts per PS row, padded to 4
  uint64_t *refb;   // reference image bytes per row
  unsigned long long *type_rows;  // [3] non-empty rows per type
};

__global__ void k_select(TypeArgs a) {
  // wave-uniform grid stride, so the per-type row counts are added once per wave
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < a.n; base += stride) {
    const uint64_t i = base + (threadIdx.x & 63u);
    const bool valid = i < a.n;
    uint32_t counted = 3u;  // type index this lane adds to type_rows (3 = none)
    if (valid) {
      const uint32_t c = a.card[i];
      const uint32_t *r = a.unbr + a.poff[i];
      uint32_t lo = c ? r[0] : 0, hi = c ? r[c - 1] : 0;
      const uint32_t nw = c ? (hi >> 6) - (lo >> 6) + 1u : 0u;
      uint32_t parts = 0;
      const bool need_parts = a.policy == EH_LAYOUT_PSHORT || a.policy == EH_LAYOUT_HYBRID3 ||
                              a.policy == EH_LAYOUT_COST || a.policy == EH_LAYOUT_LEARNED;
      if (c && need_parts) {
        parts = 1;
        for (uint32_t k = 1; k < c; k++) parts += (r[k] >> 16) != (r[k - 1] >> 16);
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
          // fewest B-side bytes: uint 4c, pshort 2c + 4 parts, bitset 8 words
          const uint64_t bu = 4ull * c, bp = 2ull * c + 4ull * parts, bb = 8ull * nw;
          t = EH_U;
          uint64_t best = bu;
          if (c > 1 && bp < best) { t = EH_PS; best = bp; }
          if (c > 1 && bb < best) { t = EH_BS; best = bb; }
          break;
        }
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
      // reference row image: optional u64 tag, then the flattened sub-layout
      const bool tagged = a.policy >= EH_LAYOUT_HYBRID;
      uint64_t bytes = tagged ? 8ull : 0ull;
      if (is_uint_type(t)) bytes += 4ull * c;
      else if (t == EH_BS) bytes += c ? 4ull + 8ull + 8ull * nw : 0ull;
      else bytes += c ? 4ull + 2ull * shorts : 0ull;
      a.refb[i] = bytes;
      if (c) counted = t == EH_UG ? EH_U : t;
    }
    for (uint32_t q = 0; q < 3; q++) {
      const uint64_t b = __ballot(counted == q);
      if ((threadIdx.x & 63u) == 0 && b) atomicAdd(&a.type_rows[q], (unsigned long long)__popcll(b));
    }
  }
}

// one wave per bitset row: OR the bits into zeroed words
__global__ void __launch_bounds__(256) k_fill_bitset(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr,
                                                     const uint8_t *type, const uint64_t *bsoff, uint32_t n,
                                                     unsigned long long *bsw) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    if (type[i] != EH_BS) continue;
    const uint32_t c = card[i];
    if (!c) continue;
    const uint32_t *r = unbr + poff[i];
    const uint32_t w0 = r[0] >> 6;
    unsigned long long *W = bsw + bsoff[i];
    for (uint32_t k = lane; k < c; k += 64) {
      const uint32_t x = r[k];
      atomicOr(&W[(x >> 6) - w0], 1ull << (x & 63u));
    }
  }
}

// one thread per pshort row: [prefix, count-1, low16...] per 16-bit prefix (pshort.hpp:46-78)
__global__ void k_fill_pshort(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr,
                              const uint8_t *type, const uint64_t *psoff, const uint32_t *ps_n, uint32_t n,
                              uint16_t *psd) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (type[i] != EH_PS) continue;
    const uint32_t c = card[i];
    if (!c) continue;
    const uint32_t *r = unbr + poff[i];
    uint16_t *S = psd + psoff[i];
    uint32_t w = 0, hdr = 0, run = 0, cur = 0xFFFFFFFFu;
    for (uint32_t k = 0; k < c; k++) {
      const uint32_t hi = r[k] >> 16;
      if (hi != cur) {
        if (cur != 0xFFFFFFFFu) S[hdr + 1] = (uint16_t)(run - 1u);
        hdr = w;
        S[w++] = (uint16_t)hi;
        S[w++] = 0;
        cur = hi;
        run = 0;
      }
      S[w++] = (uint16_t)(r[k] & 0xFFFFu);
      run++;
    }
    S[hdr + 1] = (uint16_t)(run - 1u);
    for (uint32_t k = w; k < ps_n[i]; k++) S[k] = 0;
  }
}

__global__ void k_desc(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr, const uint8_t *type,
                       const uint32_t *parts, const uint32_t *bs_n, const uint32_t *ps_n, const uint64_t *bsoff,
                       const uint64_t *psoff, uint32_t n, RowDesc *desc) {
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
      d.tinfo = EH_BS | (bs_n[i] << 2);
      d.w0 = unbr[poff[i]] >> 6;
      d.off = (uint32_t)bsoff[i];
    } else if (c && t == EH_PS) {
      const uint32_t shorts = c + 2u * parts[i];
      d.tinfo = EH_PS | (shorts << 2);
      d.w0 = parts[i];
      d.off = (uint32_t)(psoff[i] >> 2);
    }
    desc[i] = d;
    (void)ps_n;
  }
}

// one wave per row: W_i = sum over kept j of card(j)
__global__ void __launch_bounds__(256) k_wedges(const uint32_t *card, const uint64_t *poff, const uint32_t *unbr,
                                                uint32_t n, uint64_t *w) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    const uint32_t c = card[i];
    const uint32_t *r = unbr + poff[i];
    uint64_t s = 0;
    for (uint32_t k = lane; k < c; k += 64) s += card[r[k]];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) w[i] = s;
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
                           const uint32_t *d_graph_of, eh_matrix **out) {
  (void)entries;
  if (policy < EH_LAYOUT_UINT || policy > EH_LAYOUT_COST) { eh_set_error("unknown layout policy %d", (int)policy); return EH_ERR_INVALID; }
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
  DevBuf pc, poff, type, parts, bs_n, ps_n, bsoff, psoff, refb, stats, wed;
  if ((s = m->dcard.alloc(4ull * n + 4)) != EH_OK) return fail(s);
  if ((s = m->daoff.alloc(4ull * n + 4)) != EH_OK) return fail(s);
  if ((s = pc.alloc(4ull * n + 8)) != EH_OK) return fail(s);
  if ((s = poff.alloc(8ull * n + 16)) != EH_OK) return fail(s);
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
  k_pack<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(d_off, d_nbr, n, filter, d_mask, m->dcard.as<uint32_t>(),
                                                     poff.as<uint64_t>(), m->dunbr.as<uint32_t>());
  if ((s = prim_reduce_max_u32(ctx, m->dcard.as<uint32_t>(), n, &m->max_row)) != EH_OK) return fail(s);
  // learned types are computed from the packed rows before selection
  DevBuf forced;
  if (policy == EH_LAYOUT_LEARNED) {
    if ((s = forced.alloc((size_t)n + 16)) != EH_OK) return fail(s);
    m->card = 0;
    if ((s = matrix_select_learned(m, forced.as<uint8_t>())) != EH_OK) return fail(s);
  }
  if ((s = type.alloc((size_t)n + 16)) != EH_OK) return fail(s);
  if ((s = parts.alloc(4ull * n + 4)) != EH_OK) return fail(s);
  if ((s = bs_n.alloc(4ull * n + 8)) != EH_OK) return fail(s);
  if ((s = ps_n.alloc(4ull * n + 8)) != EH_OK) return fail(s);
  if ((s = refb.alloc(8ull * n + 8)) != EH_OK) return fail(s);
  if ((s = stats.alloc(64)) != EH_OK) return fail(s);
  if (hipMemsetAsync(stats.p, 0, 64, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemsetAsync(bs_n.p, 0, 4ull * n + 8, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemsetAsync(ps_n.p, 0, 4ull * n + 8, st) != hipSuccess) return fail(EH_ERR_HIP);
  TypeArgs ta;
  ta.card = m->dcard.as<uint32_t>();
  ta.poff = poff.as<uint64_t>();
  ta.unbr = m->dunbr.as<uint32_t>();
  ta.forced = forced.p ? forced.as<uint8_t>() : nullptr;
  ta.n = n;
  ta.policy = policy;
  ta.type = type.as<uint8_t>();
  ta.parts = parts.as<uint32_t>();
  ta.bs_n = bs_n.as<uint32_t>();
  ta.ps_n = ps_n.as<uint32_t>();
  ta.refb = refb.as<uint64_t>();
  ta.type_rows = stats.as<unsigned long long>();
  k_select<<<grid_for(n), 256, 0, st>>>(ta);
  if ((s = bsoff.alloc(8ull * n + 16)) != EH_OK) return fail(s);
  if ((s = psoff.alloc(8ull * n + 16)) != EH_OK) return fail(s);
  if ((s = prim_exscan_u32_to_u64(ctx, bs_n.as<uint32_t>(), bsoff.as<uint64_t>(), (uint64_t)n + 1)) != EH_OK) return fail(s);
  if ((s = prim_exscan_u32_to_u64(ctx, ps_n.as<uint32_t>(), psoff.as<uint64_t>(), (uint64_t)n + 1)) != EH_OK) return fail(s);
  uint64_t hb[2] = {0, 0};
  if (hipMemcpyAsync(&hb[0], bsoff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipMemcpyAsync(&hb[1], psoff.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  unsigned long long trows[3];
  if (hipMemcpyAsync(trows, stats.p, 24, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hipStreamSynchronize(st) != hipSuccess) return fail(EH_ERR_HIP);
  if (hb[0] >= (1ull << 32) || (hb[1] >> 2) >= (1ull << 32)) { eh_set_error("layout payload too large for 32-bit offsets"); return fail(EH_ERR_UNSUPPORTED); }
  m->bs_words = hb[0];
  m->ps_shorts = hb[1];
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
                                               type.as<uint8_t>(), psoff.as<uint64_t>(), ps_n.as<uint32_t>(), n,
                                               m->dpsd.as<uint16_t>());
  if ((s = m->ddesc.alloc(16ull * n + 16)) != EH_OK) return fail(s);
  k_desc<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(), m->dunbr.as<uint32_t>(),
                                      type.as<uint8_t>(), parts.as<uint32_t>(), bs_n.as<uint32_t>(), ps_n.as<uint32_t>(),
                                      bsoff.as<uint64_t>(), psoff.as<uint64_t>(), n, m->ddesc.as<RowDesc>());
  if (hipEventRecord(ctx->ev1, st) != hipSuccess) return fail(EH_ERR_HIP);
  if ((s = wed.alloc(8ull * n + 8)) != EH_OK) return fail(s);
  k_wedges<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), poff.as<uint64_t>(),
                                                       m->dunbr.as<uint32_t>(), n, wed.as<uint64_t>());
  if (hipGetLastError() != hipSuccess) { eh_set_error("build kernel launch failed"); return fail(EH_ERR_HIP); }
  if ((s = prim_reduce_sum_u64(ctx, wed.as<uint64_t>(), n, &m->wedges)) != EH_OK) return fail(s);
  if ((s = prim_reduce_sum_u64(ctx, refb.as<uint64_t>(), n, &m->ref_row_bytes)) != EH_OK) return fail(s);
  // kept entries = sum of card
  {
    DevBuf c64;
    if ((s = c64.alloc(8ull * n + 8)) != EH_OK) return fail(s);
    k_u32_to_u64_plus<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), n, c64.as<uint64_t>());
    if ((s = prim_reduce_sum_u64(ctx, c64.as<uint64_t>(), n, &m->card)) != EH_OK) return fail(s);
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) ms = 0;
  m->build_ms = ms;
  if (d_graph_of) {
    if ((s = m->dgraph_of.alloc(4ull * n + 4)) != EH_OK) return fail(s);
    if (hipMemcpyAsync(m->dgraph_of.p, d_graph_of, 4ull * n, hipMemcpyDeviceToDevice, st) != hipSuccess) return fail(EH_ERR_HIP);
  }
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
                          dmask.as<uint8_t>(), policy, nullptr, out);
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
  out->payload_bytes = 8 * m->bs_words + 2 * m->ps_shorts + 4 * m->unbr_words;
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
  uint64_t bytes = 0;
  if (d.card == 0) bytes = 0;
  else if (is_uint_type(t)) bytes = 4ull * d.card;
  else if (t == EH_BS) bytes = 8ull + 8ull * desc_aux(d);
  else bytes = 2ull * desc_aux(d);
  view->data = buf;
  view->cardinality = d.card;
  view->number_of_bytes = bytes;
  view->type = (int32_t)(is_uint_type(t) ? EH_U : t);
  view->reserved = 0;
  if (bytes > cap) { eh_set_error("row needs %llu bytes", (unsigned long long)bytes); return EH_ERR_CAPACITY; }
  if (!bytes) return EH_OK;
  if (is_uint_type(t)) {
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
  for (uint32_t i = 0; i < m->n; i++) types[i] = (uint8_t)desc_type(d[i]);
  return EH_OK;
}

extern "C" void eh_matrix_free(eh_matrix *m) { delete m; }
