// Entry points whose implementation is not in this build yet: they report
// OMB_EUNSUPPORTED so callers fail loudly instead of computing nothing.
#include "common.hpp"

using namespace omb;

static omb_status unsupported(const char* what) {
    set_error(std::string(what) + ": not implemented in this build");
    return OMB_EUNSUPPORTED;
}

extern "C" omb_status omb_serve_run(omb_request*, int32_t, const omb_serve_config*,
                                    omb_request_result*, omb_serve_stats*) {
    return unsupported("omb_serve_run");
}
