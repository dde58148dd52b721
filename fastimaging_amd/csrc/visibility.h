// Model and residual visibilities of a point-source sky model (visibility.hip).
#pragma once

#include "common.h"

namespace stp_hip {

// Model visibilities of `n_src` sky-model rows (l, m, flux; host, n_src x 3
// column-major) at the device baselines d_uvw (n_vis x 3 column-major, wavelengths)
// into d_out (n_vis complex), or, with `subtract`, d_out -= model (the residual).
// Enqueued on the context stream.
void model_visibilities(stp_ctx* ctx, size_t n_src, const double* skymodel, size_t n_vis, const double* d_uvw, double* d_out,
    bool subtract);

} // namespace stp_hip
