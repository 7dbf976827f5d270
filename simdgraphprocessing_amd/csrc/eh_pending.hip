// eh_pending.hip -- entry points whose device implementation is not built yet.
#include "eh_internal.hpp"

#define EH_PENDING(name) \
  do { eh_set_error("%s is not implemented in this build", name); return EH_ERR_UNSUPPORTED; } while (0)

eh_status matrix_select_learned(eh_matrix *, uint8_t *) { EH_PENDING("EH_LAYOUT_LEARNED"); }

extern "C" {
eh_status eh_intersect(eh_ctx *, const eh_row_view *, const eh_row_view *, int, void *, size_t, eh_row_view *) { EH_PENDING("eh_intersect"); }
eh_status eh_intersect_rows(eh_matrix *, const uint32_t *, const uint32_t *, uint64_t, uint64_t *) { EH_PENDING("eh_intersect_rows"); }
eh_status eh_selector_set_params(eh_ctx *, const float *) { EH_PENDING("eh_selector_set_params"); }
eh_status eh_selector_get_params(eh_ctx *, float *) { EH_PENDING("eh_selector_get_params"); }
eh_status eh_matrix_features(eh_matrix *, float *) { EH_PENDING("eh_matrix_features"); }
eh_status eh_matrix_profile_layouts(eh_matrix *, float *, float *) { EH_PENDING("eh_matrix_profile_layouts"); }
eh_status eh_mlp_forward(eh_ctx *, const float *, const float *, int, float *) { EH_PENDING("eh_mlp_forward"); }
eh_status eh_mlp_loss_grad(eh_ctx *, const float *, const float *, const float *, const float *, int, float *, float *) { EH_PENDING("eh_mlp_loss_grad"); }
eh_status eh_selector_train(eh_ctx *, const float *, const float *, const float *, int, int, int, float, uint64_t, float *) { EH_PENDING("eh_selector_train"); }
}
