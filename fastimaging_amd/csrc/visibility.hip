// Model visibilities of a local sky model of point sources and the residual
// visibilities of `reduce -d` on gfx950.
//
// generate_visibilities_from_local_skymodel (src/stp/visibility/visibility.cpp:31-41)
// adds, for every source row (l, m, flux) in sky-model order,
// visibilities_for_point_source (visibility.cpp:15-29):
//   n = sqrt(1 - l^2 - m^2),  offset = -(l, m, n - 1),
//   vis = (flux * n) * exp(i * (-2 pi) * (uvw . offset)).
// The reduce driver subtracts the model from the data before imaging
// (src/reduce/reduce.cpp:113-122).
//
// One thread per visibility walks the sources, which the workgroup stages in LDS;
// the per-source constants (offset, flux * n) are computed on the host in FP64,
// as the reference computes them once per source.
#include <cmath>
#include <vector>

#include "common.h"
#include "visibility.h"

namespace stp_hip {
namespace {

constexpr int VWG = 256;
constexpr int SRC_CHUNK = 1024; // sources staged in LDS at a time (32 KB)

struct SrcTerm {
    double o0, o1, o2, amp;
};

// mode 0: out = model; mode 1: out = out - model (the residual, in place).
__global__ __launch_bounds__(VWG) void model_vis_kernel(const double* uvw, long long n, const SrcTerm* src, int n_src,
    double* out, int mode)
{
#pragma clang fp contract(off)
    __shared__ SrcTerm s[SRC_CHUNK];
    const long long i = (long long)blockIdx.x * VWG + threadIdx.x;
    const bool on = i < n;
    const double u = on ? uvw[i] : 0.0, v = on ? uvw[n + i] : 0.0, w = on ? uvw[2 * n + i] : 0.0;
    const double c = -2.0 * M_PI;
    double ar = 0.0, ai = 0.0;
    for (int k0 = 0; k0 < n_src; k0 += SRC_CHUNK) {
        const int cnt = min(SRC_CHUNK, n_src - k0);
        __syncthreads();
        for (int k = threadIdx.x; k < cnt; k += VWG)
            s[k] = src[k0 + k];
        __syncthreads();
        for (int k = 0; k < cnt; k++) {
            const SrcTerm t = s[k];
            const double d = (u * t.o0 + v * t.o1) + w * t.o2;
            double sn, cs;
            sincos(c * d, &sn, &cs);
            ar += t.amp * cs;
            ai += t.amp * sn;
        }
    }
    if (!on)
        return;
    if (mode == 0) {
        out[2 * i] = ar;
        out[2 * i + 1] = ai;
    } else {
        out[2 * i] = out[2 * i] - ar;
        out[2 * i + 1] = out[2 * i + 1] - ai;
    }
}

} // namespace

void model_visibilities(stp_ctx* ctx, size_t n_src, const double* skymodel, size_t n_vis, const double* d_uvw, double* d_out,
    bool subtract)
{
    STP_REQUIRE(n_src == 0 || skymodel != nullptr, "skymodel is NULL");
    STP_REQUIRE(n_vis == 0 || (d_uvw != nullptr && d_out != nullptr), "visibility buffers are NULL");
    STP_REQUIRE(n_src < (size_t(1) << 31), "too many sources");
    for (size_t k = 0; k < n_src; k++) {
        const double l = skymodel[k], m = skymodel[n_src + k], flux = skymodel[2 * n_src + k];
        STP_REQUIRE(std::isfinite(l) && std::isfinite(m) && std::isfinite(flux), "skymodel entries must be finite");
        STP_REQUIRE(l * l + m * m <= 1.0, "skymodel source outside the unit sphere (l^2 + m^2 > 1)");
    }
    if (n_vis == 0)
        return;
    // per-source constants (visibility.cpp:19-24) in sky-model order, staged in
    // page-locked memory once the previous call's upload from it has finished
    const size_t bytes = sizeof(SrcTerm) * std::max<size_t>(1, n_src);
    if (!ctx->skymodel_event)
        HIP_CHECK(hipEventCreateWithFlags(&ctx->skymodel_event, hipEventDisableTiming));
    HIP_CHECK(hipEventSynchronize(ctx->skymodel_event));
    SrcTerm* h = (SrcTerm*)host_stage(ctx, "skymodel_terms", bytes);
    for (size_t k = 0; k < n_src; k++) {
        const double l = skymodel[k], m = skymodel[n_src + k], flux = skymodel[2 * n_src + k];
        const double src_n = std::sqrt(1.0 - l * l - m * m);
        h[k] = { -l, -m, -(src_n - 1.0), flux * src_n };
    }
    SrcTerm* d_src = (SrcTerm*)ws_get(ctx, "skymodel_terms", bytes);
    if (n_src)
        HIP_CHECK(hipMemcpyAsync(d_src, h, sizeof(SrcTerm) * n_src, hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipEventRecord(ctx->skymodel_event, ctx->stream));
    launch(ctx, "model_vis", model_vis_kernel, dim3(div_up(n_vis, VWG)), dim3(VWG), 0, d_uvw, (long long)n_vis,
        (const SrcTerm*)d_src, int(n_src), d_out, subtract ? 1 : 0);
}

} // namespace stp_hip
