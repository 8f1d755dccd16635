// server.hip -- serving entry points (device-resident request loop).
// Placeholder bodies until the continuous-batching loop lands; they fail loudly.
#include "../../include/sstmc.h"

extern "C" {
int sstmc_server_create(uint32_t, uint64_t, uint32_t, sstmc_server** out) {
    if (out) *out = nullptr;
    return SSTMC_NOT_SUPPORTED;
}
int sstmc_submit(sstmc_server*, const sstmc_config*, uint64_t*) { return SSTMC_NOT_SUPPORTED; }
int sstmc_poll(sstmc_server*, uint64_t) { return SSTMC_NOT_SUPPORTED; }
int sstmc_wait(sstmc_server*, uint64_t) { return SSTMC_NOT_SUPPORTED; }
int sstmc_fetch(sstmc_server*, uint64_t, double*, double*) { return SSTMC_NOT_SUPPORTED; }
int sstmc_server_stats_get(sstmc_server*, sstmc_server_stats*) { return SSTMC_NOT_SUPPORTED; }
void sstmc_server_destroy(sstmc_server*) {}
}
