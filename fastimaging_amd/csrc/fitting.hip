// Least-squares 2-D Gaussian fit of source-finder islands on gfx950
// (IslandParams::leastsq_fit_gaussian_2d, src/stp/sourcefind/sourcefind.cpp:719-866;
// model and analytic Jacobian: src/stp/sourcefind/fitting.cpp:59-120).
//
// Three passes:
//   fit_bbox     bounding box of every island in visual coordinates
//                (sourcefind.cpp:449-666): one thread per pixel, atomic min / max;
//   fit_gather   one wave per island walks its box in the reference's residual
//                order (columns left..right, rows top..bottom, sourcefind.cpp:756-768)
//                and compacts the pixels carrying the island's label (ballots);
//   fit_lm       one wave per island minimises the sum of squared residuals:
//                Levenberg-Marquardt with Marquardt's diagonal scaling and
//                Nielsen's damping update.  Lanes stride over the island's pixels;
//                J^T J, J^T r and the cost are summed by xor butterflies (every lane
//                ends with the same bits), so all lanes run the same 6x6 Cholesky
//                solve and take the same decisions.
// The reference minimises with Ceres; the least-squares minimum this returns is
// the one every converged Ceres solver type returns.
#include <cmath>
#include <vector>

#include "common.h"

namespace {

struct FitJob {
    int32_t label;   // island label (+id / -id)
    int32_t count;   // num_samples
    uint32_t offset; // first sample of the island in the gathered list
    int32_t pad;
    double start[6]; // amplitude, x, y, x_stddev, y_stddev, theta
    double extremum_val;
};

struct Sample {
    double x, y, v;
};

struct FitOut {
    double p[6];
    int32_t iterations, status, found, pad;
};

constexpr int FWG = 256; // four waves: four islands per workgroup

__global__ __launch_bounds__(256) void fit_bbox_init_kernel(int4* bbox, int n, int rows, int cols)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n)
        bbox[k] = make_int4(rows, -1, cols, -1); // top, bottom, left, right
}

__global__ __launch_bounds__(256) void fit_bbox_kernel(const int32_t* labels, int rows, int cols, const int32_t* isl_of,
    int maxlab, int4* bbox)
{
    const long long li = (long long)blockIdx.x * 256 + threadIdx.x;
    if (li >= (long long)rows * cols)
        return;
    const int32_t l = labels[li];
    if (l == 0 || l > maxlab || l < -maxlab)
        return;
    const int k = isl_of[l + maxlab];
    if (k < 0)
        return;
    const int r = int(li % rows), c = int(li / rows);
    const int hs = cols / 2, vs = rows / 2;
    const int row = r < vs ? r + vs : r - vs;
    const int col = c < hs ? c + hs : c - hs;
    int* b = reinterpret_cast<int*>(&bbox[k]);
    atomicMin(&b[0], row);
    atomicMax(&b[1], row);
    atomicMin(&b[2], col);
    atomicMax(&b[3], col);
}

template <typename T>
__global__ __launch_bounds__(FWG) void fit_gather_kernel(const T* img, const int32_t* labels, int rows, int cols,
    const FitJob* jobs, const int4* bbox, int n, Sample* samples, FitOut* out)
{
    const int wid = (blockIdx.x * FWG + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wid >= n)
        return;
    const FitJob J = jobs[wid];
    const int4 b = bbox[wid];
    const int hs = cols / 2, vs = rows / 2;
    const int H = b.y - b.x + 1, W = b.w - b.z + 1;
    int found = 0;
    if (H > 0 && W > 0) {
        const long long total = (long long)H * W;
        for (long long e0 = 0; e0 < total; e0 += 64) {
            const long long e = e0 + lane;
            bool hit = false;
            double v = 0.0;
            int i = 0, j = 0;
            if (e < total) {
                i = b.z + int(e / H);
                j = b.x + int(e % H);
                const int ii = i < hs ? i + hs : i - hs;
                const int jj = j < vs ? j + vs : j - vs;
                const size_t li = size_t(jj) + size_t(ii) * rows;
                hit = labels[li] == J.label;
                if (hit)
                    v = double(img[li]);
            }
            const unsigned long long m = __ballot(hit);
            const int pos = found + __popcll(m & ((1ull << lane) - 1ull));
            if (hit && pos < J.count)
                samples[J.offset + pos] = Sample{ double(i), double(j), v };
            found += __popcll(m);
        }
    }
    if (lane == 0)
        out[wid].found = found;
}

// a, b, c of the rotated Gaussian and their derivatives (fitting.cpp:66-101).
struct Shape {
    double a, b, c;
    double da_dt, da_dsx, da_dsy, db_dt, db_dsx, db_dsy, dc_dt, dc_dsx, dc_dsy;
};

__device__ Shape shape_of(const double* p)
{
    const double sx = p[3], sy = p[4], t = p[5];
    const double ct = cos(t), st = sin(t);
    const double ct2 = ct * ct, st2 = st * st, s2t = sin(2.0 * t), c2t = cos(2.0 * t);
    const double sx2 = sx * sx, sy2 = sy * sy, sx3 = sx2 * sx, sy3 = sy2 * sy;
    Shape s;
    s.a = 0.5 * ((ct2 / sx2) + (st2 / sy2));
    s.b = 0.5 * ((s2t / sx2) - (s2t / sy2));
    s.c = 0.5 * ((st2 / sx2) + (ct2 / sy2));
    s.da_dt = st * ct * ((1.0 / sy2) - (1.0 / sx2));
    s.da_dsx = -ct2 / sx3;
    s.da_dsy = -st2 / sy3;
    s.db_dt = (c2t / sx2) - (c2t / sy2);
    s.db_dsx = -s2t / sx3;
    s.db_dsy = s2t / sy3;
    s.dc_dt = -s.da_dt;
    s.dc_dsx = -st2 / sx3;
    s.dc_dsy = -ct2 / sy3;
    return s;
}

// Residual model - data of one sample and (when J) its Jacobian row.
__device__ __forceinline__ double residual(const double* p, const Shape& s, const Sample& q, double* J)
{
    const double dx = q.x - p[1], dy = q.y - p[2];
    const double dx2 = dx * dx, dy2 = dy * dy, dxy = dx * dy;
    const double e = exp(-((s.a * dx2) + (s.b * dxy) + (s.c * dy2)));
    const double g = p[0] * e;
    if (J) {
        J[0] = e; // dg/dA (g / A in fitting.cpp:103, without the division)
        J[1] = g * ((2.0 * s.a * dx) + (s.b * dy));
        J[2] = g * ((s.b * dx) + (2.0 * s.c * dy));
        J[3] = g * (-(s.da_dsx * dx2 + s.db_dsx * dxy + s.dc_dsx * dy2));
        J[4] = g * (-(s.da_dsy * dx2 + s.db_dsy * dxy + s.dc_dsy * dy2));
        J[5] = g * (-(s.da_dt * dx2 + s.db_dt * dxy + s.dc_dt * dy2));
    }
    return g - q.v;
}

// Sum over the wave; every lane receives the same bits (a + b == b + a).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

__device__ double half_cost(const double* p, const Sample* q, int m, int lane)
{
    const Shape s = shape_of(p);
    double c = 0.0;
    for (int k = lane; k < m; k += 64) {
        const double r = residual(p, s, q[k], nullptr);
        c += r * r;
    }
    return 0.5 * wave_sum(c);
}

__device__ bool chol_solve6(const double* M, const double* rhs, double* d)
{
    double L[21];
    // packed lower triangle, row-major: (i, j) at i (i + 1) / 2 + j
    for (int j = 0; j < 6; j++) {
        double s = M[j * (j + 1) / 2 + j];
        for (int q = 0; q < j; q++)
            s -= L[j * (j + 1) / 2 + q] * L[j * (j + 1) / 2 + q];
        if (!(s > 0.0))
            return false;
        const double ljj = sqrt(s);
        L[j * (j + 1) / 2 + j] = ljj;
        for (int i = j + 1; i < 6; i++) {
            double t = M[i * (i + 1) / 2 + j];
            for (int q = 0; q < j; q++)
                t -= L[i * (i + 1) / 2 + q] * L[j * (j + 1) / 2 + q];
            L[i * (i + 1) / 2 + j] = t / ljj;
        }
    }
    double y[6];
    for (int i = 0; i < 6; i++) {
        double t = rhs[i];
        for (int q = 0; q < i; q++)
            t -= L[i * (i + 1) / 2 + q] * y[q];
        y[i] = t / L[i * (i + 1) / 2 + i];
    }
    for (int i = 5; i >= 0; i--) {
        double t = y[i];
        for (int q = i + 1; q < 6; q++)
            t -= L[q * (q + 1) / 2 + i] * d[q];
        d[i] = t / L[i * (i + 1) / 2 + i];
    }
    return true;
}

__global__ __launch_bounds__(FWG) void fit_lm_kernel(const FitJob* jobs, int n, const Sample* samples, FitOut* out)
{
#pragma clang fp contract(off)
    const int wid = (blockIdx.x * FWG + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wid >= n)
        return;
    const FitJob J = jobs[wid];
    FitOut& o = out[wid];
    if (J.count < 3) {
        // sourcefind.cpp:724-728: no solve for one- or two-pixel islands
        if (lane == 0) {
            o.p[0] = J.extremum_val;
            for (int i = 1; i < 6; i++)
                o.p[i] = 0.0;
            o.iterations = 0;
            o.status = 2;
        }
        return;
    }
    const Sample* q = samples + J.offset;
    const int m = J.count;
    double x[6];
    for (int i = 0; i < 6; i++)
        x[i] = J.start[i];
    double cost = half_cost(x, q, m, lane);
    double mu = STP_FIT_MU0, nu = 2.0;
    int it = 0, status = 1;
    bool done = false;
    while (it < STP_FIT_MAX_ITERS && !done) {
        // normal equations at x
        double A[21], g[6];
        for (int k = 0; k < 21; k++)
            A[k] = 0.0;
        for (int i = 0; i < 6; i++)
            g[i] = 0.0;
        {
            const Shape s = shape_of(x);
            for (int k = lane; k < m; k += 64) {
                double Jr[6];
                const double r = residual(x, s, q[k], Jr);
                int t = 0;
                for (int i = 0; i < 6; i++) {
                    g[i] += Jr[i] * r;
                    for (int j = 0; j <= i; j++, t++)
                        A[t] += Jr[i] * Jr[j];
                }
            }
        }
        for (int k = 0; k < 21; k++)
            A[k] = wave_sum(A[k]);
        for (int i = 0; i < 6; i++)
            g[i] = wave_sum(g[i]);
        bool accepted = false;
        while (it < STP_FIT_MAX_ITERS && !accepted && !done) {
            it++;
            double M[21], rhs[6], d[6];
            for (int k = 0; k < 21; k++)
                M[k] = A[k];
            for (int i = 0; i < 6; i++) {
                const int kk = i * (i + 1) / 2 + i;
                const double dg = fmin(fmax(A[kk], STP_FIT_DIAG_MIN), STP_FIT_DIAG_MAX);
                M[kk] = A[kk] + mu * dg;
                rhs[i] = -g[i];
            }
            if (!chol_solve6(M, rhs, d)) {
                mu *= nu;
                nu *= 2.0;
                done = mu > STP_FIT_MU_MAX;
                continue;
            }
            double xn[6];
            for (int i = 0; i < 6; i++)
                xn[i] = x[i] + d[i];
            const double cn = half_cost(xn, q, m, lane);
            // predicted decrease of the Gauss-Newton model: -(d.g + d.A.d / 2)
            double dg = 0.0, dad = 0.0;
            int t = 0;
            for (int i = 0; i < 6; i++) {
                dg += d[i] * g[i];
                for (int j = 0; j <= i; j++, t++)
                    dad += (i == j ? 1.0 : 2.0) * d[i] * A[t] * d[j];
            }
            const double pred = -(dg + 0.5 * dad);
            const double actual = cost - cn;
            if (actual > 0.0 && pred > 0.0) {
                const double rho = actual / pred;
                double nx = 0.0, nd = 0.0;
                for (int i = 0; i < 6; i++) {
                    nx += x[i] * x[i];
                    nd += d[i] * d[i];
                }
                for (int i = 0; i < 6; i++)
                    x[i] = xn[i];
                const bool small = actual <= STP_FIT_FTOL * cost || sqrt(nd) <= STP_FIT_XTOL * (sqrt(nx) + STP_FIT_XTOL);
                cost = cn;
                const double tt = 2.0 * rho - 1.0;
                mu *= fmax(1.0 / 3.0, 1.0 - tt * tt * tt);
                nu = 2.0;
                accepted = true;
                if (small || cost == 0.0) {
                    status = 0;
                    done = true;
                }
            } else {
                mu *= nu;
                nu *= 2.0;
                if (mu > STP_FIT_MU_MAX) {
                    status = 0; // no decrease left at machine precision
                    done = true;
                }
            }
        }
    }
    if (lane == 0) {
        for (int i = 0; i < 6; i++)
            o.p[i] = x[i];
        o.iterations = it;
        o.status = status;
    }
}

// Gaussian2dParams::convert_to_constrained_parameters (fitting.cpp:32-57).
stp_gaussian2d constrain(const double* p)
{
    stp_gaussian2d g{ p[0], p[1], p[2], p[3], p[4], p[5] };
    const double half_pi = M_PI / 2.0;
    g.semimajor = std::fabs(g.semimajor);
    g.semiminor = std::fabs(g.semiminor);
    if (g.semimajor < g.semiminor) {
        std::swap(g.semimajor, g.semiminor);
        g.theta += half_pi;
    }
    double mt = std::fmod(g.theta, M_PI);
    if (mt <= -half_pi)
        mt += M_PI;
    else if (mt > half_pi)
        mt -= M_PI;
    g.theta = mt;
    return g;
}

} // namespace

namespace stp_hip {

void fit_islands_device(stp_ctx* ctx, const void* d_image, int32_t precision, int rows, int cols, const int32_t* d_labels,
    const stp_island* islands, size_t n, stp_fit_result* results)
{
    STP_REQUIRE(rows >= 2 && cols >= 2, "image must be at least 2 x 2");
    STP_REQUIRE(d_image && d_labels, "image or label map is NULL");
    STP_REQUIRE(precision == STP_PRECISION_FP64 || precision == STP_PRECISION_FP32C ||
            precision == STP_PRECISION_FP32_TWOFLOAT, "unknown precision");
    STP_REQUIRE(n < (size_t(1) << 28), "too many islands for one call");
    if (n == 0)
        return;
    STP_REQUIRE(islands && results, "island or result array is NULL");
    int32_t maxlab = 0;
    for (size_t k = 0; k < n; k++) {
        STP_REQUIRE(islands[k].label_idx != 0 && islands[k].label_idx != INT32_MIN, "island with label 0");
        maxlab = std::max(maxlab, std::abs(islands[k].label_idx));
    }
    std::vector<int32_t> isl_of(size_t(2) * maxlab + 1, -1);
    std::vector<FitJob> jobs(n);
    size_t total = 0;
    for (size_t k = 0; k < n; k++) {
        const stp_island& I = islands[k];
        STP_REQUIRE(I.num_samples >= 0, "negative island sample count");
        STP_REQUIRE(isl_of[size_t(I.label_idx + maxlab)] < 0, "two islands with one label");
        isl_of[size_t(I.label_idx + maxlab)] = int32_t(k);
        FitJob& J = jobs[k];
        J.label = I.label_idx;
        J.count = I.num_samples;
        J.offset = uint32_t(total);
        J.pad = 0;
        const stp_gaussian2d& m = I.moments_fit;
        // sourcefind.cpp:736-743: non-positive axes start at 1
        const double st[6] = { m.amplitude, m.x_centre, m.y_centre, m.semimajor > 0.0 ? m.semimajor : 1.0,
            m.semiminor > 0.0 ? m.semiminor : 1.0, m.theta };
        for (int i = 0; i < 6; i++)
            J.start[i] = st[i];
        J.extremum_val = I.extremum_val;
        total += size_t(I.num_samples);
        STP_REQUIRE(total < (size_t(1) << 32), "too many island samples for one call");
    }
    int32_t* d_isl_of = (int32_t*)ws_get(ctx, "fit_isl_of", sizeof(int32_t) * isl_of.size());
    FitJob* d_jobs = (FitJob*)ws_get(ctx, "fit_jobs", sizeof(FitJob) * n);
    int4* d_bbox = (int4*)ws_get(ctx, "fit_bbox", sizeof(int4) * n);
    Sample* d_samples = (Sample*)ws_get(ctx, "fit_samples", sizeof(Sample) * std::max<size_t>(1, total));
    FitOut* d_out = (FitOut*)ws_get(ctx, "fit_out", sizeof(FitOut) * n);
    HIP_CHECK(hipMemcpyAsync(d_isl_of, isl_of.data(), sizeof(int32_t) * isl_of.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(FitJob) * n, hipMemcpyHostToDevice, ctx->stream));
    launch(ctx, "fit_bbox_init", fit_bbox_init_kernel, dim3(div_up(n, 256)), dim3(256), 0, d_bbox, int(n), rows, cols);
    const size_t np = size_t(rows) * cols;
    launch(ctx, "fit_bbox", fit_bbox_kernel, dim3(div_up(np, 256)), dim3(256), 0, d_labels, rows, cols,
        (const int32_t*)d_isl_of, int(maxlab), d_bbox);
    const unsigned wg = div_up(n, FWG / 64);
    if (precision == STP_PRECISION_FP64)
        launch(ctx, "fit_gather", fit_gather_kernel<double>, dim3(wg), dim3(FWG), 0, (const double*)d_image, d_labels, rows,
            cols, (const FitJob*)d_jobs, (const int4*)d_bbox, int(n), d_samples, d_out);
    else
        launch(ctx, "fit_gather", fit_gather_kernel<float>, dim3(wg), dim3(FWG), 0, (const float*)d_image, d_labels, rows,
            cols, (const FitJob*)d_jobs, (const int4*)d_bbox, int(n), d_samples, d_out);
    launch(ctx, "fit_lm", fit_lm_kernel, dim3(wg), dim3(FWG), 0, (const FitJob*)d_jobs, int(n), (const Sample*)d_samples,
        d_out);
    std::vector<FitOut> h_out(n);
    std::vector<int4> h_bbox(n);
    HIP_CHECK(hipMemcpyAsync(h_out.data(), d_out, sizeof(FitOut) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(h_bbox.data(), d_bbox, sizeof(int4) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < n; k++) {
        if (islands[k].num_samples >= 3 && h_out[k].found != islands[k].num_samples)
            throw Error(STP_ERR_INVALID_ARGUMENT, "island " + std::to_string(k) + " (label " +
                    std::to_string(islands[k].label_idx) + ") has " + std::to_string(islands[k].num_samples) +
                    " samples but the label map holds " + std::to_string(h_out[k].found));
        stp_fit_result& r = results[k];
        r.leastsq_fit = h_out[k].status == 2 ? stp_gaussian2d{ h_out[k].p[0], 0.0, 0.0, 0.0, 0.0, 0.0 } : constrain(h_out[k].p);
        r.bounding_box = stp_bounding_box{ h_bbox[k].x, h_bbox[k].y, h_bbox[k].z, h_bbox[k].w };
        r.iterations = h_out[k].iterations;
        r.status = h_out[k].status;
    }
}

} // namespace stp_hip
