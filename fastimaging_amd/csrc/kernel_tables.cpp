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

// J1 by the rational approximation of Numerical Recipes in C (2nd ed., section 6.5,
// bessj1), as the reference's besselj1 (aw_projection.cpp:490-508) uses it.
static double bessel_j1(double x)
{
    if (x < 8.0) {
        const double x2 = x * x;
        const double num = x * (72362614232.0 + x2 * (-7895059235.0 + x2 * (242396853.1 + x2 * (-2972611.439 +
            x2 * (15704.48260 + x2 * (-30.16036606))))));
        const double den = 144725228442.0 + x2 * (2300535178.0 + x2 * (18583304.74 + x2 * (99447.43394 +
            x2 * (376.9991397 + x2))));
        return num / den;
    }
    const double z = 8.0 / x, z2 = z * z, ph = x - 2.356194491;
    const double pa = 1.0 + z2 * (0.183105e-2 + z2 * (-0.3516396496e-4 + z2 * (0.2457520174e-5 + z2 * (-0.240337019e-6))));
    const double pb = 0.04687499995 + z2 * (-0.2002690873e-3 + z2 * (0.8449199096e-5 + z2 * (-0.88228987e-6 +
        z2 * 0.105787412e-6)));
    return std::sqrt(0.636619772 / x) * (std::cos(ph) * pa - z * std::sin(ph) * pb);
}

// DHT(j, i) = P_i(r_j) - P_i(r_{j-1}) with P_i(r) = (2 pi / k_i) r J1(k_i r), the integral
// of 2 pi r J0(k_i r) from 0 to r, k_i = pi i / n, ring edges r_j = j + 1/2 (the last
// n - 1); for k_0 = 0 the ring areas pi r_j^2 - pi r_{j-1}^2 (aw_projection.cpp:530-596).
std::vector<double> dht_table(size_t n)
{
    std::vector<double> t(n * n);
    std::vector<double> r(n);
    for (size_t j = 0; j < n; j++)
        r[j] = double(j) + 0.5;
    r[n - 1] = double(n - 1);
    double prev = M_PI * r[0] * r[0];
    t[0] = prev;
    for (size_t j = 1; j < n; j++) {
        const double cur = M_PI * r[j] * r[j];
        t[j * n] = cur - prev;
        prev = cur;
    }
    const double kf = M_PI / double(n);
    for (size_t i = 1; i < n; i++) {
        const double k = kf * double(i), kn = (2.0 * M_PI) / k;
        double hi = kn * r[n - 1] * bessel_j1(k * r[n - 1]);
        for (size_t j = n - 1; j > 0; j--) {
            const double lo = kn * r[j - 1] * bessel_j1(k * r[j - 1]);
            t[j * n + i] = hi - lo;
            hi = lo;
        }
        t[i] = hi;
    }
    return t;
}

std::vector<double> dft_twiddles(size_t n)
{
    std::vector<double> t(2 * n);
    for (size_t m = 0; m < n; m++) {
        const double a = -2.0 * M_PI * double(m) / double(n);
        t[2 * m] = std::cos(a);
        t[2 * m + 1] = std::sin(a);
    }
    return t;
}

} // namespace stp_hip
