// pp_batch.hip -- batched persistent PCG over many independent systems.
#include "pp_internal.h"

extern "C" {
int pp_batch_create(pp_ctx *, uint32_t, const uint32_t *, const uint32_t *, const uint32_t *, const double *,
                    pp_batch **) {
    return PP_ERR_UNSUPPORTED;
}
int pp_batch_set_values(pp_ctx *, pp_batch *, const double *, int) { return PP_ERR_UNSUPPORTED; }
int pp_batch_solve(pp_ctx *, pp_batch *, const pp_solve_opts *, const double *, double *, int32_t *,
                   pp_solve_report *) {
    return PP_ERR_UNSUPPORTED;
}
int pp_batch_solve_t(pp_ctx *, pp_batch *, const pp_solve_opts *, const double *, double *, int32_t *,
                     pp_solve_report *) {
    return PP_ERR_UNSUPPORTED;
}
int pp_batch_tune(pp_ctx *, pp_batch *, int *) { return PP_ERR_UNSUPPORTED; }
void pp_batch_destroy(pp_batch *) {}
}
