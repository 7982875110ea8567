// Entry points whose implementation is not in this build yet: they report
// OMB_EUNSUPPORTED so callers fail loudly instead of computing nothing.
#include "common.hpp"

using namespace omb;

static omb_status unsupported(const char* what) {
    set_error(std::string(what) + ": not implemented in this build");
    return OMB_EUNSUPPORTED;
}

extern "C" omb_status omb_group_plan_create(const omb_gemm_desc*, int32_t, const omb_gemv_desc*,
                                            int32_t, int32_t, omb_stream_t, omb_group_plan** p) {
    if (p) *p = nullptr;
    return unsupported("omb_group_plan_create");
}
extern "C" omb_status omb_group_plan_launch(omb_group_plan*, omb_stream_t) {
    return unsupported("omb_group_plan_launch");
}
extern "C" omb_status omb_group_plan_launch_direct(omb_group_plan*, omb_stream_t) {
    return unsupported("omb_group_plan_launch_direct");
}
extern "C" omb_status omb_group_plan_describe(const omb_group_plan*, char*, size_t) {
    return unsupported("omb_group_plan_describe");
}
extern "C" omb_status omb_group_plan_destroy(omb_group_plan*) { return OMB_OK; }
extern "C" omb_status omb_autotune_group(const omb_gemm_desc*, int32_t, int32_t, const char*,
                                         omb_stream_t) {
    return unsupported("omb_autotune_group");
}
extern "C" omb_status omb_autotune_load(const char*) { return unsupported("omb_autotune_load"); }
extern "C" omb_status omb_serve_run(omb_request*, int32_t, const omb_serve_config*,
                                    omb_request_result*, omb_serve_stats*) {
    return unsupported("omb_serve_run");
}
