// pending.cpp -- entry points whose engines are not built yet.
#include "scmc_internal.h"

using namespace scmc;

extern "C" {
int scmc_server_start(const scmc_server_config *) { return set_error(SCMC_E_UNSUPPORTED, "server not built"); }
int scmc_server_stop(void) { return SCMC_OK; }
int scmc_submit_mccov(const scmc_problem *, int64_t, int32_t, int64_t *)
{
    return set_error(SCMC_E_UNSUPPORTED, "server not built");
}
int scmc_poll(int64_t, double *, double *, double *, double *, scmc_stats *)
{
    return set_error(SCMC_E_UNSUPPORTED, "server not built");
}
int scmc_wait(int64_t, double *, double *, double *, double *, scmc_stats *)
{
    return set_error(SCMC_E_UNSUPPORTED, "server not built");
}
int scmc_cancel(int64_t) { return set_error(SCMC_E_UNSUPPORTED, "server not built"); }
int64_t scmc_request_footprint(int32_t, int32_t, int32_t) { return -1; }
}
