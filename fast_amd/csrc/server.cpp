// Continuous-batching request loop (placeholder until the resident-batch server lands).
#include "fast_tau.h"
#include "tau_host.h"

extern "C" {
int fast_server_create(fast_ctx*, const fast_server_config*, fast_server** out) {
    if (out) *out = nullptr;
    return FAST_E_BADARG;
}
int fast_server_submit(fast_server*, const uint8_t*, int, int, int, const fast_tau_params*, uint64_t*) { return FAST_E_BADARG; }
int fast_server_wait(fast_server*, int, uint64_t*, fast_tau_result*, double*) { return FAST_E_BADARG; }
int fast_server_pending(fast_server*) { return 0; }
void fast_server_destroy(fast_server*) {}
}
