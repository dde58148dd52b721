// Host-side convolution-kernel tables for the gridder (computed once per call
// on the CPU in double precision, then copied to the device).
//   functors:            src/stp/convolution/conv_func.cpp:10-202
//   make_1D_kernel:      src/stp/convolution/conv_func.h:319-337
//   populate_kernel_cache (separable form): src/stp/gridder/gridder.h:73-105
//   ImgDomKernel:        src/stp/convolution/conv_func.h:351-412
#pragma once

#include <cstddef>
#include <vector>

#include "../../include/stp_capi.h"

namespace stp_hip {

struct KernelFn {
    int kind;     // stp_kernel_function
    double param; // TopHat/Triangle: half_base_width; others: truncation radius
    double width_g, width_s;
};

KernelFn kernel_fn(int kind, double param);
// From a C-ABI kernel spec (widths as given).
KernelFn kernel_fn(const struct stp_kernel_spec& spec);
void eval_kernel(const KernelFn& k, const double* x, size_t n, double* out);
bool eval_gcf(const KernelFn& k, const double* x, size_t n, double* out);

// 1-D kernels at the oversampled offsets (a - os/2)/os, a = 0..os (os even) or
// a = 0 (os = 1), each normalised to unit sum: table[a * (2s+1) + t].
// The reference's 2-D cache entry (a, b) is table[a][row] * table[b][col].
std::vector<double> kernel_1d_table(const KernelFn& k, int support, int oversampling);

// Image-domain kernel of length kernel_size (ImgDomKernel); DFT path evaluated
// directly from the non-zero kernel samples.
std::vector<double> img_dom_kernel(const KernelFn& k, size_t kernel_size, bool normalize, bool analytic_gcf);

// Hankel W-kernels (src/stp/gridder/aw_projection.cpp:477-596).  The n x n discrete
// Hankel transform, row-major by output: t[j * n + i] = DHT(j, i), so that threads
// over outputs i read consecutive words for each input j.
std::vector<double> dht_table(size_t n);
// exp(-2 pi i m / n), m = 0..n-1, interleaved (cos, sin): forward DFT twiddles.
std::vector<double> dft_twiddles(size_t n);

} // namespace stp_hip
