// GPU source finder (placeholder until the kernels land).
#include "sourcefind.h"

namespace stp_hip {

void source_find_device(stp_ctx*, const void*, int, int, int, const stp_sourcefind_pars&, int32_t*, stp_island*, size_t,
    size_t*, stp_sourcefind_stats*)
{
    throw Error(STP_ERR_UNSUPPORTED, "source finder not built yet");
}

} // namespace stp_hip
