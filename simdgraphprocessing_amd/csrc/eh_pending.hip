// eh_pending.hip -- entry points whose device implementation is not built yet.
#include "eh_internal.hpp"

#define EH_PENDING(name) \
  do { eh_set_error("%s is not implemented in this build", name); return EH_ERR_UNSUPPORTED; } while (0)

extern "C" {
eh_status eh_intersect(eh_ctx *, const eh_row_view *, const eh_row_view *, int, void *, size_t, eh_row_view *) { EH_PENDING("eh_intersect"); }
eh_status eh_intersect_rows(eh_matrix *, const uint32_t *, const uint32_t *, uint64_t, uint64_t *) { EH_PENDING("eh_intersect_rows"); }
}
