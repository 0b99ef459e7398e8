// Serving loop (continuous batching of independent solves).
#include "thp_internal.h"

extern "C" {
thp_status thp_serve_start(thp_ctx* ctx, const thp_serve_opts*) { return ctx ? THP_ERR_STATE : THP_ERR_ARG; }
thp_status thp_serve_add_system(thp_ctx* ctx, const thp_system_desc*, int*) { return ctx ? THP_ERR_STATE : THP_ERR_ARG; }
thp_status thp_serve_submit(thp_ctx* ctx, int, const double*, const double*, const double*, uint64_t*) { return ctx ? THP_ERR_STATE : THP_ERR_ARG; }
thp_status thp_serve_poll(thp_ctx* ctx, uint64_t, thp_result*, double) { return ctx ? THP_ERR_STATE : THP_ERR_ARG; }
thp_status thp_serve_get_dipoles(thp_ctx* ctx, int, double*) { return ctx ? THP_ERR_STATE : THP_ERR_ARG; }
thp_status thp_serve_stop(thp_ctx* ctx) { return ctx ? THP_OK : THP_ERR_ARG; }
}
