// Serving loop (continuous batching) of the NDDO engine.
#include "engine_internal.h"

namespace nddo {
struct ServedJob {};
void Engine::start_server() {}
void Engine::stop_server() {}
}  // namespace nddo

extern "C" {
int nddo_submit(nddo_engine*, const nddo_job_desc*, nddo_job_id*) { nddo::set_error("serving not built"); return NDDO_BAD_ARG; }
int nddo_poll(nddo_engine*, nddo_job_id) { return NDDO_BAD_ARG; }
int nddo_wait(nddo_engine*, nddo_job_id, double) { return NDDO_BAD_ARG; }
int nddo_get_result(nddo_engine*, nddo_job_id, nddo_result*) { return NDDO_BAD_ARG; }
int nddo_release(nddo_engine*, nddo_job_id) { return NDDO_BAD_ARG; }
}
