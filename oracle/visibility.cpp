/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Model visibilities of a local sky model of point sources and their subtraction
 * from the data (the `reduce -d` residual path).
 */
#include <cmath>
#include <thread>
#include <vector>

#include "oracle.h"

namespace oracle {

// generate_visibilities_from_local_skymodel (src/stp/visibility/visibility.cpp:31-41):
// model_vis starts at zero and each source row (l, m, flux) of the sky model adds
// visibilities_for_point_source (visibility.cpp:15-29):
//   n = sqrt(1 - l^2 - m^2);  offset = -(l, m, n - 1);
//   vis = (flux * n) * exp(i * (-2 pi) * (uvw . offset)).
// The exponent is the complex scalar (0, -2 pi) times the real dot product, so the
// exponential is (cos a, sin a) with a = -2 pi * (u*o0 + v*o1 + w*o2).  Sources are
// added in sky-model order, for every visibility.
// skymodel: n_src x 3 column-major; uvw: n_vis x 3 column-major.
void model_visibilities(const double* skymodel, size_t n_src, const double* uvw, size_t n_vis, cx* out, int nthreads)
{
    std::vector<double> o0(n_src), o1(n_src), o2(n_src), amp(n_src);
    for (size_t k = 0; k < n_src; k++) {
        const double l = skymodel[k], m = skymodel[n_src + k], flux = skymodel[2 * n_src + k];
        const double src_n = std::sqrt(1.0 - l * l - m * m);
        o0[k] = -l;
        o1[k] = -m;
        o2[k] = -(src_n - 1.0);
        amp[k] = flux * src_n;
    }
    const double c = -2.0 * M_PI;
    auto work = [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i++) {
            const double u = uvw[i], v = uvw[n_vis + i], w = uvw[2 * n_vis + i];
            cx acc(0.0, 0.0);
            for (size_t k = 0; k < n_src; k++) {
                const double d = (u * o0[k] + v * o1[k]) + w * o2[k];
                const double ang = c * d;
                acc += cx(amp[k] * std::cos(ang), amp[k] * std::sin(ang));
            }
            out[i] = acc;
        }
    };
    if (nthreads <= 1 || n_vis < 4096) {
        work(0, n_vis);
        return;
    }
    std::vector<std::thread> th;
    const size_t chunk = (n_vis + size_t(nthreads) - 1) / size_t(nthreads);
    for (int t = 0; t < nthreads; t++) {
        const size_t a = std::min(n_vis, size_t(t) * chunk), b = std::min(n_vis, a + chunk);
        th.emplace_back(work, a, b);
    }
    for (auto& x : th)
        x.join();
}

} // namespace oracle
