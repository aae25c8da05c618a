// Entry points not yet implemented in this build.
#include "gp_internal.hpp"

extern "C" {
int gp_filter(gp_ctx* ctx, int, int, const gp_pairs*, gp_pairs*, gp_stats*) {
    if (ctx) ctx->err = "gp_filter: not implemented";
    return GP_ERR_UNSUPPORTED;
}
}
