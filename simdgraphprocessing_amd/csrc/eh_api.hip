// eh_api.hip -- context management, error reporting and the status strings of
// the C ABI.  The reference has no device or status codes: it prints and exits
// (src/MutableGraph.hpp:373-389, src/emptyheaded.hpp:77-80); here every entry
// point returns an eh_status and leaves detail in eh_last_error().
#include <cstdarg>
#include <cstring>

#include "eh_internal.hpp"

static thread_local char g_err[512] = "";

void eh_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

double ctx_elapsed_ms(eh_ctx *ctx) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) return 0;
  return ms;
}

extern "C" int eh_abi_version(void) { return EH_ABI_VERSION; }

extern "C" const char *eh_last_error(void) { return g_err; }

extern "C" const char *eh_status_str(eh_status s) {
  switch (s) {
    case EH_OK: return "ok";
    case EH_ERR_INVALID: return "invalid argument";
    case EH_ERR_OOM: return "out of memory";
    case EH_ERR_HIP: return "HIP runtime error";
    case EH_ERR_IO: return "I/O error";
    case EH_ERR_UNSUPPORTED: return "unsupported";
    case EH_ERR_CAPACITY: return "buffer too small";
  }
  return "unknown status";
}

extern "C" eh_status eh_ctx_create(int device, eh_ctx **out) {
  if (!out) { eh_set_error("eh_ctx_create: null out"); return EH_ERR_INVALID; }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    eh_set_error("no HIP device available (%s)", hipGetErrorString(e));
    return EH_ERR_HIP;
  }
  if (device < 0 || device >= count) { eh_set_error("device %d out of range (%d devices)", device, count); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(device));
  eh_ctx *c = new eh_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->num_cus = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    c->name += " / ";
    c->name += prop.name;
  }
  // a blocking stream: synchronous hipMemcpy calls on the null stream order
  // after the work queued here, which the host-side readbacks rely on
  if (hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    eh_set_error("stream/event creation failed");
    delete c;
    return EH_ERR_HIP;
  }
  if (c->counter.alloc(256) != EH_OK) { eh_ctx_destroy(c); return EH_ERR_OOM; }
  // the device address of the static staging of eh_intersect_rows_static, once per device
  const eh_status ss = static_scratch_init(device);
  if (ss != EH_OK) { eh_ctx_destroy(c); return ss; }
  *out = c;
  return EH_OK;
}

eh_status ctx_part_events(eh_ctx *ctx) {
  for (PartEv &p : ctx->parts) {
    if (!p.b) EH_HIP(hipEventCreate(&p.b));
    if (!p.e) EH_HIP(hipEventCreate(&p.e));
  }
  return EH_OK;
}

extern "C" void eh_ctx_destroy(eh_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (PartEv &p : ctx->parts) {
    if (p.b) (void)hipEventDestroy(p.b);
    if (p.e) (void)hipEventDestroy(p.e);
  }
  ctx->scratch.release();
  ctx->counter.release();
  ctx->cache.clear();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" eh_status eh_ctx_device_name(eh_ctx *ctx, char *buf, size_t cap) {
  if (!ctx || !buf || !cap) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  snprintf(buf, cap, "%s (%d CUs)", ctx->name.c_str(), ctx->num_cus);
  return EH_OK;
}
