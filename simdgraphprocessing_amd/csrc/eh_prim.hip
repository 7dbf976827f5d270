// eh_prim.hip -- device-wide primitives (radix sort, unique, scan, reduce) used by
// the graph canonicaliser and the layout builder.  Thin wrappers over rocPRIM's
// device algorithms so the other translation units do not instantiate them.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "eh_internal.hpp"

namespace {

eh_status ensure_scratch(eh_ctx *ctx, size_t bytes) {
  if (ctx->scratch.bytes >= bytes && ctx->scratch.p) return EH_OK;
  return ctx->scratch.alloc(bytes + (bytes >> 3) + 256);
}

}  // namespace

eh_status prim_sort_u64(eh_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int end_bit) {
  if (n < 2) return EH_OK;
  size_t bytes = 0;
  EH_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys, tmp, (size_t)n, 0, (unsigned)end_bit,
                                  ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::radix_sort_keys(ctx->scratch.p, bytes, keys, tmp, (size_t)n, 0, (unsigned)end_bit,
                                  ctx->stream));
  EH_HIP(hipMemcpyAsync(keys, tmp, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
  return EH_OK;
}

// keys ordered by bits [begin_bit, end_bit) only, stably (an LSD radix sort keeps the input
// order among equal digits); end_bit must stay below 64 when begin_bit > 0 (see prim_sort_u64)
eh_status prim_sort_u64_range(eh_ctx *ctx, uint64_t *keys, uint64_t *tmp, uint64_t n, int begin_bit, int end_bit) {
  if (n < 2) return EH_OK;
  if (begin_bit < 0 || end_bit > 64 || begin_bit >= end_bit || (begin_bit > 0 && end_bit == 64)) {
    eh_set_error("prim_sort_u64_range: bit range [%d, %d) not supported", begin_bit, end_bit);
    return EH_ERR_INVALID;
  }
  size_t bytes = 0;
  EH_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys, tmp, (size_t)n, (unsigned)begin_bit, (unsigned)end_bit,
                                  ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::radix_sort_keys(ctx->scratch.p, bytes, keys, tmp, (size_t)n, (unsigned)begin_bit,
                                  (unsigned)end_bit, ctx->stream));
  EH_HIP(hipMemcpyAsync(keys, tmp, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
  return EH_OK;
}

eh_status prim_sort_pairs_u32(eh_ctx *ctx, uint32_t *keys, uint32_t *keys_tmp, uint32_t *vals,
                              uint32_t *vals_tmp, uint64_t n) {
  if (n < 2) return EH_OK;
  size_t bytes = 0;
  EH_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys_tmp, vals, vals_tmp, (size_t)n, 0, 32,
                                   ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::radix_sort_pairs(ctx->scratch.p, bytes, keys, keys_tmp, vals, vals_tmp, (size_t)n,
                                   0, 32, ctx->stream));
  EH_HIP(hipMemcpyAsync(keys, keys_tmp, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  EH_HIP(hipMemcpyAsync(vals, vals_tmp, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  return EH_OK;
}

// stable: values of equal keys keep their input order
eh_status prim_sort_pairs_u64_u32(eh_ctx *ctx, uint64_t *keys, uint64_t *keys_tmp, uint32_t *vals,
                                  uint32_t *vals_tmp, uint64_t n) {
  if (n < 2) return EH_OK;
  size_t bytes = 0;
  EH_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys_tmp, vals, vals_tmp, (size_t)n, 0, 64,
                                   ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::radix_sort_pairs(ctx->scratch.p, bytes, keys, keys_tmp, vals, vals_tmp, (size_t)n,
                                   0, 64, ctx->stream));
  EH_HIP(hipMemcpyAsync(keys, keys_tmp, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
  EH_HIP(hipMemcpyAsync(vals, vals_tmp, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  return EH_OK;
}

eh_status prim_sort_u32(eh_ctx *ctx, uint32_t *keys, uint32_t *tmp, uint64_t n) {
  if (n < 2) return EH_OK;
  size_t bytes = 0;
  EH_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys, tmp, (size_t)n, 0, 32, ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::radix_sort_keys(ctx->scratch.p, bytes, keys, tmp, (size_t)n, 0, 32, ctx->stream));
  EH_HIP(hipMemcpyAsync(keys, tmp, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  return EH_OK;
}

eh_status prim_unique_u64(eh_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n,
                          uint64_t *n_out) {
  if (n == 0) { *n_out = 0; return EH_OK; }
  EH_CHECK(ctx->counter.p ? EH_OK : ctx->counter.alloc(256));
  size_t *d_cnt = ctx->counter.as<size_t>();
  size_t bytes = 0;
  EH_HIP(rocprim::unique(nullptr, bytes, in, out, d_cnt, (size_t)n, rocprim::equal_to<uint64_t>(),
                         ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::unique(ctx->scratch.p, bytes, in, out, d_cnt, (size_t)n,
                         rocprim::equal_to<uint64_t>(), ctx->stream));
  size_t h = 0;
  EH_HIP(hipMemcpyAsync(&h, d_cnt, sizeof(size_t), hipMemcpyDeviceToHost, ctx->stream));
  EH_HIP(hipStreamSynchronize(ctx->stream));
  *n_out = h;
  return EH_OK;
}

eh_status prim_exscan_u64(eh_ctx *ctx, const uint64_t *in, uint64_t *out, uint64_t n) {
  if (n == 0) return EH_OK;
  size_t bytes = 0;
  EH_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, (uint64_t)0, (size_t)n,
                                 rocprim::plus<uint64_t>(), ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::exclusive_scan(ctx->scratch.p, bytes, in, out, (uint64_t)0, (size_t)n,
                                 rocprim::plus<uint64_t>(), ctx->stream));
  return EH_OK;
}

eh_status prim_exscan_u32_to_u64(eh_ctx *ctx, const uint32_t *in, uint64_t *out, uint64_t n) {
  if (n == 0) return EH_OK;
  auto it = rocprim::make_transform_iterator(in, [](uint32_t x) -> uint64_t { return x; });
  size_t bytes = 0;
  EH_HIP(rocprim::exclusive_scan(nullptr, bytes, it, out, (uint64_t)0, (size_t)n,
                                 rocprim::plus<uint64_t>(), ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::exclusive_scan(ctx->scratch.p, bytes, it, out, (uint64_t)0, (size_t)n,
                                 rocprim::plus<uint64_t>(), ctx->stream));
  return EH_OK;
}

eh_status prim_reduce_max_u32(eh_ctx *ctx, const uint32_t *in, uint64_t n, uint32_t *out) {
  EH_CHECK(ctx->counter.p ? EH_OK : ctx->counter.alloc(256));
  uint32_t *d = reinterpret_cast<uint32_t *>(ctx->counter.as<uint8_t>() + 64);
  size_t bytes = 0;
  EH_HIP(rocprim::reduce(nullptr, bytes, in, d, (uint32_t)0, (size_t)n, rocprim::maximum<uint32_t>(),
                         ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::reduce(ctx->scratch.p, bytes, in, d, (uint32_t)0, (size_t)n,
                         rocprim::maximum<uint32_t>(), ctx->stream));
  EH_HIP(hipMemcpyAsync(out, d, 4, hipMemcpyDeviceToHost, ctx->stream));
  EH_HIP(hipStreamSynchronize(ctx->stream));
  return EH_OK;
}

eh_status prim_reduce_sum_u64(eh_ctx *ctx, const uint64_t *in, uint64_t n, uint64_t *out) {
  EH_CHECK(ctx->counter.p ? EH_OK : ctx->counter.alloc(256));
  uint64_t *d = reinterpret_cast<uint64_t *>(ctx->counter.as<uint8_t>() + 128);
  size_t bytes = 0;
  EH_HIP(rocprim::reduce(nullptr, bytes, in, d, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(),
                         ctx->stream));
  EH_CHECK(ensure_scratch(ctx, bytes));
  EH_HIP(rocprim::reduce(ctx->scratch.p, bytes, in, d, (uint64_t)0, (size_t)n,
                         rocprim::plus<uint64_t>(), ctx->stream));
  EH_HIP(hipMemcpyAsync(out, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  EH_HIP(hipStreamSynchronize(ctx->stream));
  return EH_OK;
}
