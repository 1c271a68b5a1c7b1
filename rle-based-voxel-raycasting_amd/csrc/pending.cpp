// Entry points whose implementation is being brought up.
#include "internal.h"

using namespace rlevox;

extern "C" {
rlevox_status rlevox_autotune(rlevox_ctx*, const rlevox_view*, int32_t, const rlevox_render_params*, rlevox_tuning*, double*) {
  return fail(RLEVOX_ERR_INVALID_ARG, "autotune: not available in this build");
}
rlevox_status rlevox_server_create(rlevox_ctx*, const rlevox_server_config*, rlevox_server**) {
  return fail(RLEVOX_ERR_INVALID_ARG, "server: not available in this build");
}
rlevox_status rlevox_server_submit(rlevox_server*, const rlevox_scene*, const rlevox_pose*, int32_t, int32_t, uint64_t*) {
  return fail(RLEVOX_ERR_INVALID_ARG, "server: not available in this build");
}
rlevox_status rlevox_server_wait(rlevox_server*, uint64_t, int32_t, uint16_t*, int32_t*, double*) {
  return fail(RLEVOX_ERR_INVALID_ARG, "server: not available in this build");
}
rlevox_status rlevox_server_cancel(rlevox_server*, uint64_t) {
  return fail(RLEVOX_ERR_INVALID_ARG, "server: not available in this build");
}
rlevox_status rlevox_server_stats_get(rlevox_server*, rlevox_server_stats*) {
  return fail(RLEVOX_ERR_INVALID_ARG, "server: not available in this build");
}
void rlevox_server_destroy(rlevox_server*) {}
}
