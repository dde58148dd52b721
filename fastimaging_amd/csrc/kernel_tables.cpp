// Host-side convolution-kernel tables; see kernel_tables.h for the reference citations.
#include "kernel_tables.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "../../include/stp_capi.h"

namespace stp_hip {

KernelFn kernel_fn(int kind, double param)
{
    KernelFn k{ kind, param, 1.0, 1.0 };
    if (kind == STP_KERNEL_GAUSSIANSINC) {
        k.width_g = 2.52; // conv_func.h:216
        k.width_s = 1.55; // conv_func.h:217
    }
    return k;
}

KernelFn kernel_fn(const stp_kernel_spec& s)
{
    KernelFn k{ s.kernel_function, s.param, s.width_normalization_gaussian, s.width_normalization_sinc };
    return k;
}

static inline double gauss(double x, double w)
{
    const double c = -1.0 / (w * w);
    return std::exp(c * (x * x));
}

static inline double sinc(double x, double w)
{
    const double c = M_PI / w;
    double r = x * c;
    if (r == 0.0)
        r = 1.0e-20; // conv_func.cpp:49
    return std::sin(r) / r;
}

// Spheroidal function, rational approximation of F. Schwab (Indirect Imaging, 1984),
// M = 6, alpha = 1 (conv_func.cpp:112-182).
static double spheroidal(double nu)
{
    static const double p[2][5] = { { 8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1 },
        { 4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2 } };
    static const double q[2][3] = { { 1.0, 8.212018e-1, 2.078043e-1 }, { 1.0, 9.599102e-1, 2.918724e-1 } };
    nu = std::fabs(nu);
    const int part = (nu >= 0.75 && nu <= 1.0) ? 1 : 0;
    const double nuend = (nu >= 0.0 && nu < 0.75) ? 0.75 : ((nu <= 1.0) ? 1.0 : 0.0);
    const double d = std::pow(nu, 2) - std::pow(nuend, 2);
    double top = p[part][0], bot = q[part][0];
    for (int k = 1; k < 5; k++)
        top += p[part][k] * std::pow(d, k);
    for (int k = 1; k < 3; k++)
        bot += q[part][k] * std::pow(d, k);
    return bot > 0.0 ? top / bot : 0.0;
}

void eval_kernel(const KernelFn& k, const double* x, size_t n, double* out)
{
    for (size_t i = 0; i < n; i++) {
        const double r = x[i];
        double v;
        switch (k.kind) {
        case STP_KERNEL_TOPHAT:
            v = (std::fabs(r) < k.param) ? 1.0 : 0.0;
            break;
        case STP_KERNEL_TRIANGLE:
            v = std::max(1.0 - std::fabs(r) / k.param, 0.0);
            break;
        case STP_KERNEL_SINC:
            v = sinc(r, k.width_s);
            break;
        case STP_KERNEL_GAUSSIAN:
            v = gauss(r, k.width_g);
            break;
        case STP_KERNEL_GAUSSIANSINC:
            v = gauss(r, k.width_g) * sinc(r, k.width_s);
            break;
        case STP_KERNEL_PSWF: {
            const double nu = r / k.param;
            v = spheroidal(nu) * (1.0 - std::pow(nu, 2));
            break;
        }
        default:
            throw std::invalid_argument("unknown kernel function");
        }
        if (k.kind != STP_KERNEL_TOPHAT && k.kind != STP_KERNEL_TRIANGLE && k.param > 0.0 && std::fabs(r) > k.param)
            v = 0.0;
        out[i] = v;
    }
}

bool eval_gcf(const KernelFn& k, const double* x, size_t n, double* out)
{
    for (size_t i = 0; i < n; i++) {
        if (k.kind == STP_KERNEL_GAUSSIAN) {
            const double a = x[i] * M_PI / (k.width_g * 2);
            out[i] = std::exp(-1.0 * (a * a));
        } else if (k.kind == STP_KERNEL_SINC) {
            out[i] = (std::fabs(x[i]) < (1.0 * k.width_s)) ? 1.0 : 0.0;
        } else if (k.kind == STP_KERNEL_PSWF) {
            out[i] = spheroidal(x[i]);
        } else {
            return false; // the reference asserts (conv_func.cpp:18-20,37-39,106-110)
        }
    }
    return true;
}

std::vector<double> kernel_1d_table(const KernelFn& k, int support, int oversampling)
{
    const int op = oversampling / 2, cs = 2 * op + 1, ks = 2 * support + 1;
    std::vector<double> t(size_t(cs) * ks), d(ks), v(ks);
    for (int a = 0; a < cs; a++) {
        const double off = (double(a) - double(op)) / double(oversampling);
        for (int i = 0; i < ks; i++)
            d[i] = (double(i) - double(support)) / double(1) - off;
        eval_kernel(k, d.data(), ks, v.data());
        double s = 0.0;
        for (double x : v)
            s += x;
        for (int i = 0; i < ks; i++)
            t[size_t(a) * ks + i] = v[i] / s;
    }
    return t;
}

std::vector<double> img_dom_kernel(const KernelFn& k, size_t K, bool normalize, bool analytic_gcf)
{
    std::vector<double> aa(K, 0.0);
    const size_t centre = K / 2;
    if (analytic_gcf) {
        std::vector<double> r(K);
        for (size_t i = 0; i < K; i++)
            r[i] = (double(i) - double(centre)) / double(centre);
        if (!eval_gcf(k, r.data(), K, aa.data()))
            throw std::invalid_argument("analytic gcf is not defined for this kernel function");
    } else {
        std::vector<double> d(K), kv(K);
        for (size_t i = 0; i < K; i++)
            d[i] = double(i) - double(centre);
        eval_kernel(k, d.data(), K, kv.data());
        double s = 0.0;
        for (double x : kv)
            s += x;
        if (s > 0.0)
            for (auto& x : kv)
                x = x / s;
        // aa[K/2 + i] = sum_x k(x) cos(2 pi x i / K), x in [-K/2, K/2), from the
        // non-zero samples only (R2HC of the wrapped kernel, real part).
        std::vector<std::pair<long, double>> nz;
        for (size_t i = 0; i < K; i++)
            if (kv[i] != 0.0)
                nz.push_back({ long(i) - long(centre), kv[i] });
        const long double two_pi = 6.283185307179586476925286766559005768L;
        for (size_t j = 0; j < K; j++) {
            const long f = long(j) - long(centre);
            long double acc = 0.0L;
            for (auto& e : nz) {
                const long m = ((e.first * f) % long(K) + long(K)) % long(K);
                acc += (long double)e.second * cosl(two_pi * (long double)m / (long double)K);
            }
            aa[j] = double(acc);
        }
    }
    if (normalize) {
        double s = 0.0;
        for (double x : aa)
            s += x;
        if (s > 0.0)
            for (auto& x : aa)
                x = x / s;
    }
    return aa;
}

} // namespace stp_hip
