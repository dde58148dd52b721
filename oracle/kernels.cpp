/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Convolution-kernel functors and kernel sampling:
 *   src/stp/convolution/conv_func.cpp:10-202, conv_func.h:176-222,279-412.
 */
#include <cmath>
#include <stdexcept>

#include "oracle.h"

namespace oracle {

KernelFunc make_kernel_func(int kind, double param)
{
    KernelFunc k{ kind, param, 1.0, 1.0 };
    if (kind == STP_KERNEL_GAUSSIANSINC) {
        // GaussianSinc defaults (conv_func.h:216-217)
        k.width_g = 2.52;
        k.width_s = 1.55;
    }
    return k;
}

KernelFunc make_kernel_func(const stp_kernel_spec& s)
{
    return KernelFunc{ s.kernel_function, s.param, s.width_normalization_gaussian, s.width_normalization_sinc };
}

// Gaussian::operator() (conv_func.cpp:71-82) without truncation.
static inline double gaussian_value(double x, double w)
{
    const double c = -1.0 / (w * w);
    return std::exp(c * (x * x));
}

// Sinc::operator() (conv_func.cpp:44-59) without truncation; 0 -> 1e-20 guard (:49).
static inline double sinc_value(double x, double w)
{
    const double c = M_PI / w;
    double r = x * c;
    if (r == 0.0)
        r = 1.0e-20;
    return std::sin(r) / r;
}

// __grdsf (conv_func.cpp:112-182): Schwab's rational approximation of the
// spheroidal function, M = 6, alpha = 1 (F. Schwab, "Optimal gridding of
// visibility data in radio interferometry", 1984).
static double grdsf(double nu)
{
    static const double p[2][5] = { { 8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1 },
        { 4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2 } };
    static const double q[2][3] = { { 1.0000000e0, 8.212018e-1, 2.078043e-1 }, { 1.0000000e0, 9.599102e-1, 2.918724e-1 } };
    nu = std::fabs(nu);
    int part;
    double nuend;
    if (nu >= 0.0 && nu < 0.75) {
        part = 0;
        nuend = 0.75;
    } else if (nu >= 0.75 && nu <= 1.0) {
        part = 1;
        nuend = 1.0;
    } else {
        part = 0;
        nuend = 0.0;
    }
    double top = p[part][0], bot = q[part][0];
    const double delnusq = std::pow(nu, 2) - std::pow(nuend, 2);
    for (int k = 1; k < 5; k++)
        top += p[part][k] * std::pow(delnusq, k);
    for (int k = 1; k < 3; k++)
        bot += q[part][k] * std::pow(delnusq, k);
    return (bot > 0.0) ? top / bot : 0.0;
}

void eval_kernel(const KernelFunc& k, const double* x, size_t n, double* out)
{
    for (size_t i = 0; i < n; i++) {
        const double r = x[i];
        double v = 0.0;
        switch (k.kind) {
        case STP_KERNEL_TOPHAT: // conv_func.cpp:25-34
            v = (std::fabs(r) < k.param) ? 1.0 : 0.0;
            break;
        case STP_KERNEL_TRIANGLE: // conv_func.cpp:11-15
            v = std::max(1.0 - std::fabs(r) / k.param, 0.0);
            break;
        case STP_KERNEL_SINC:
            v = sinc_value(r, k.width_s);
            if (k.param > 0.0 && std::fabs(r) > k.param)
                v = 0.0;
            break;
        case STP_KERNEL_GAUSSIAN:
            v = gaussian_value(r, k.width_g);
            if (k.param > 0.0 && std::fabs(r) > k.param)
                v = 0.0;
            break;
        case STP_KERNEL_GAUSSIANSINC: // conv_func.cpp:92-103
            v = gaussian_value(r, k.width_g) * sinc_value(r, k.width_s);
            if (k.param > 0.0 && std::fabs(r) > k.param)
                v = 0.0;
            break;
        case STP_KERNEL_PSWF: { // conv_func.cpp:185-196
            const double nu = r / k.param;
            v = grdsf(nu) * (1.0 - std::pow(nu, 2));
            if (k.param > 0.0 && std::fabs(r) > k.param)
                v = 0.0;
            break;
        }
        default:
            throw std::invalid_argument("unknown kernel function");
        }
        out[i] = v;
    }
}

bool eval_gcf(const KernelFunc& k, const double* x, size_t n, double* out)
{
    for (size_t i = 0; i < n; i++) {
        switch (k.kind) {
        case STP_KERNEL_GAUSSIAN: { // conv_func.cpp:85-89
            const double a = x[i] * M_PI / (k.width_g * 2);
            out[i] = std::exp(-1.0 * (a * a));
            break;
        }
        case STP_KERNEL_SINC: // conv_func.cpp:62-67
            out[i] = (std::fabs(x[i]) < (1.0 * k.width_s)) ? 1.0 : 0.0;
            break;
        case STP_KERNEL_PSWF: // conv_func.cpp:199-202
            out[i] = grdsf(x[i]);
            break;
        default: // TopHat, Triangle, GaussianSinc assert(false) (conv_func.cpp:18-20,37-39,106-110)
            return false;
        }
    }
    return true;
}

// Sequential sum (arma::accu); the reference's exact summation order is not pinned.
static double accu(const std::vector<double>& v)
{
    double s = 0.0;
    for (double x : v)
        s += x;
    return s;
}

std::vector<double> make_1d_kernel(const KernelFunc& k, int support, double offset, int oversampling, bool pad, bool normalize)
{
    const int local_pad = pad ? 1 : 0;
    const int array_size = 2 * (support + local_pad) * oversampling + 1;
    const int centre_idx = (support + local_pad) * oversampling;
    std::vector<double> dist(array_size), out(array_size);
    for (int i = 0; i < array_size; i++)
        dist[i] = (double(i) - double(centre_idx)) / double(oversampling) - offset;
    eval_kernel(k, dist.data(), dist.size(), out.data());
    if (normalize) {
        const double s = accu(out);
        for (auto& v : out)
            v = v / s;
    }
    return out;
}

std::vector<double> make_kernel_array(const KernelFunc& k, int support, double off_x, double off_y, int oversampling, bool pad, bool normalize)
{
    const int local_pad = pad ? 1 : 0;
    const int n = 2 * (support + local_pad) * oversampling + 1;
    const int centre_idx = (support + local_pad) * oversampling;
    std::vector<double> dx(n), dy(n), kx(n), ky(n);
    for (int i = 0; i < n; i++) {
        const double d = (double(i) - double(centre_idx)) / double(oversampling);
        dx[i] = d - off_x;
        dy[i] = d - off_y;
    }
    eval_kernel(k, dx.data(), n, kx.data());
    eval_kernel(k, dy.data(), n, ky.data());
    std::vector<double> res(size_t(n) * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            res[i + size_t(j) * n] = ky[i] * kx[j];
    if (normalize) {
        const double s = accu(res);
        for (auto& v : res)
            v = v / s;
    }
    return res;
}

std::vector<double> img_dom_kernel(const KernelFunc& k, size_t kernel_size, bool normalize, bool analytic_gcf)
{
    std::vector<double> aa(kernel_size, 0.0);
    const size_t centre = kernel_size / 2;
    if (analytic_gcf) {
        std::vector<double> r(kernel_size);
        for (size_t i = 0; i < kernel_size; i++)
            r[i] = (double(i) - double(centre)) / double(centre);
        if (!eval_gcf(k, r.data(), kernel_size, aa.data()))
            throw std::invalid_argument("analytic gcf not defined for this kernel function");
    } else {
        std::vector<double> d(kernel_size), k1d(kernel_size);
        for (size_t i = 0; i < kernel_size; i++)
            d[i] = double(i) - double(centre);
        eval_kernel(k, d.data(), kernel_size, k1d.data());
        const double s = accu(k1d);
        if (s > 0.0)
            for (auto& v : k1d)
                v = v / s;
        std::vector<double> arr(kernel_size, 0.0), f(kernel_size);
        for (size_t i = 0; i < centre; i++) {
            arr[i] = k1d[centre + i];
            arr[kernel_size - centre + i] = k1d[i];
        }
        fft_r2hc_1d(arr.data(), kernel_size, f.data());
        for (size_t i = 0; i < kernel_size / 2; i++)
            aa[kernel_size / 2 + i] = f[i];
        for (size_t i = 1; i < kernel_size / 2 + 1; i++)
            aa[kernel_size / 2 - i] = f[i];
    }
    if (normalize) {
        const double s = accu(aa);
        if (s > 0.0)
            for (auto& v : aa)
                v = v / s;
    }
    return aa;
}

} // namespace oracle
