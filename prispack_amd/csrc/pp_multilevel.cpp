// pp_multilevel.cpp -- AMSED / GMSLR multilevel low-rank Schur preconditioners.
#include "pp_internal.h"

namespace pp {
void ml_setup(pp_ctx *, pp_precond *) {
    throw Error(PP_ERR_UNSUPPORTED, "multilevel preconditioners are not built yet");
}
void ml_apply(pp_ctx *, pp_precond *, int, const void *, void *) {
    throw Error(PP_ERR_UNSUPPORTED, "multilevel preconditioners are not built yet");
}
void ml_free(pp_precond *M) { M->multilevel = nullptr; }
}  // namespace pp
