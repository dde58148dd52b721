// Host entry of the GPU source finder (sourcefind.hip).
#pragma once

#include "common.h"

namespace stp_hip {

// SourceFindImage (src/stp/sourcefind/sourcefind.cpp:156-292) on a device image
// (double for STP_PRECISION_FP64, float for STP_PRECISION_FP32C).  The label map
// (device, may be NULL), islands and stats are written as the reference exposes them.
void source_find_device(stp_ctx* ctx, const void* d_image, int precision, int rows, int cols,
    const stp_sourcefind_pars& p, int32_t* d_label_map, stp_island* islands, size_t cap, size_t* n_islands,
    stp_sourcefind_stats* stats);

// IslandParams::leastsq_fit_gaussian_2d (sourcefind.cpp:719-866) for n islands of a
// source find on its device image and label map (fitting.hip).
void fit_islands_device(stp_ctx* ctx, const void* d_image, int32_t precision, int rows, int cols, const int32_t* d_labels,
    const stp_island* islands, size_t n, stp_fit_result* results);

} // namespace stp_hip
