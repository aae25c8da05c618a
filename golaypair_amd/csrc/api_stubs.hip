// Entry points not yet implemented in this build.
#include "gp_internal.hpp"

extern "C" {
int gp_uncompress_side(gp_ctx* ctx, int8_t*, int, int, int, int, int, gp_cands*, gp_stats*) {
    if (ctx) ctx->err = "gp_uncompress_side: not implemented";
    return GP_ERR_UNSUPPORTED;
}
int gp_uncompress(gp_ctx* ctx, int, int, int, int, const gp_pairs*, gp_pairs*, gp_cands*, gp_cands*, gp_stats*) {
    if (ctx) ctx->err = "gp_uncompress: not implemented";
    return GP_ERR_UNSUPPORTED;
}
int gp_filter(gp_ctx* ctx, int, int, const gp_pairs*, gp_pairs*, gp_stats*) {
    if (ctx) ctx->err = "gp_filter: not implemented";
    return GP_ERR_UNSUPPORTED;
}
}
