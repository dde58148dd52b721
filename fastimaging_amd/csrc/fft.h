// Host entry points of the Stockham FFT kernels (fft.hip).
#pragma once

#include "common.h"

namespace stp_hip {

// Normalisation fused into the last c2r pass (imager.h:96-213, non-FFTSHIFT):
// out(i, j) = img(oi, oj) / *total_ptr [/ (gcf[ii] * gcf[jj])], cropped from the
// padded transform size to image_size.
struct NormSpec {
    int image_size = 0;
    const double* total_ptr = nullptr; // device scalar; NULL = no scaling
    const double* gcf = nullptr;       // device array of the padded size; NULL = none
    int u_l1 = 0;                      // four-step split of the u axis, N = L1 * (N / L1); 0 = balanced
    const double* const* total_tab = nullptr; // device array: grid g's scalar (replaces total_ptr)
};

// Outputs of the batch's grids: p[g] for up to 4 grids, or the device array tab.
template <typename OT>
struct OutPtrs {
    OT* p[4];
    OT* const* tab = nullptr;
};

// Unnormalised 2-D c2r of nbatch half-plane grids ((n_rows/2+1) x N each,
// contiguous) into nbatch outputs (at most 4 unless outs.tab is set); the grids
// are overwritten.
template <typename T, typename OT>
void fft_c2r_2d(stp_ctx* ctx, cplx<T>* grids, int nbatch, int n_rows, int N, const OutPtrs<OT>& outs, const NormSpec& ns);

// Unnormalised 2-D c2r of nbatch real half-plane grids of an N x N transform, each
// stored packed as (N/2+1) x (N/2) complex with element (v, u') = G(v, 2u') +
// i G(v, 2u' + 1) (contiguous); the grids are overwritten.  Half the bytes of
// fft_c2r_2d on the same grid stored complex.
template <typename T, typename OT>
void fft_c2r_2d_real(stp_ctx* ctx, cplx<T>* grids, int nbatch, int N, const OutPtrs<OT>& outs, const NormSpec& ns);

// 2-D c2c of nbatch N x N column-major matrices (contiguous); `data` is
// overwritten, results in `out`.
template <typename T>
void fft_c2c_2d(stp_ctx* ctx, cplx<T>* data, cplx<T>* out, int N, bool forward, int nbatch = 1);

} // namespace stp_hip
