// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// Least-squares 2-D Gaussian fit of source-finder islands
// (IslandParams::leastsq_fit_gaussian_2d, src/stp/sourcefind/sourcefind.cpp:719-866):
// the island's pixels are those of its bounding box whose label equals the
// island's label (sourcefind.cpp:449-666 for the box, :756-768 for the pixel
// order), the residuals are model - data with the model and analytic Jacobian of
// src/stp/sourcefind/fitting.cpp:59-120 (fitting.h:176-215), and the start is the
// moments fit with non-positive axes replaced by 1 (sourcefind.cpp:736-743).
//
// The reference minimises with Ceres (absent here: a third-party library, not
// vendored).  Its minimiser is not restated; this file uses Levenberg-Marquardt
// (Marquardt's diagonal scaling, Nielsen's damping update) with tolerances tight
// enough that the result is the least-squares minimum itself, which every
// converged Ceres solver type also returns (the reference's own test accepts the
// same answer from LBFGS, BFGS and the trust-region solver,
// src/test/sourcefind/sourcefind_test_Fitting.cpp:107-123).  The HIP fit kernel
// runs the same iteration.
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace oracle {

namespace {

struct Px {
    double x, y, v;
};

// a, b, c of the rotated Gaussian and their derivatives (fitting.cpp:66-101).
struct Shape {
    double a, b, c;
    double da_dt, da_dsx, da_dsy, db_dt, db_dsx, db_dsy, dc_dt, dc_dsx, dc_dsy;
};

Shape shape_of(const double* p)
{
    const double sx = p[3], sy = p[4], t = p[5];
    const double ct = std::cos(t), st = std::sin(t);
    const double ct2 = ct * ct, st2 = st * st, s2t = std::sin(2.0 * t), c2t = std::cos(2.0 * t);
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

// Residual of one pixel and (if J) its Jacobian row.
double residual(const double* p, const Shape& s, const Px& q, double* J)
{
    const double dx = q.x - p[1], dy = q.y - p[2];
    const double dx2 = dx * dx, dy2 = dy * dy, dxy = dx * dy;
    const double e = std::exp(-((s.a * dx2) + (s.b * dxy) + (s.c * dy2)));
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

double half_cost(const double* p, const std::vector<Px>& px)
{
    const Shape s = shape_of(p);
    double c = 0.0;
    for (const Px& q : px) {
        const double r = residual(p, s, q, nullptr);
        c += r * r;
    }
    return 0.5 * c;
}

// Cholesky solve of the 6x6 system M d = rhs (M packed lower, row-major 21).
bool chol_solve6(const double* M, const double* rhs, double* d)
{
    double L[6][6] = {};
    int k = 0;
    double A[6][6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j <= i; j++, k++)
            A[i][j] = A[j][i] = M[k];
    for (int j = 0; j < 6; j++) {
        double s = A[j][j];
        for (int q = 0; q < j; q++)
            s -= L[j][q] * L[j][q];
        if (!(s > 0.0))
            return false;
        L[j][j] = std::sqrt(s);
        for (int i = j + 1; i < 6; i++) {
            double t = A[i][j];
            for (int q = 0; q < j; q++)
                t -= L[i][q] * L[j][q];
            L[i][j] = t / L[j][j];
        }
    }
    double y[6];
    for (int i = 0; i < 6; i++) {
        double t = rhs[i];
        for (int q = 0; q < i; q++)
            t -= L[i][q] * y[q];
        y[i] = t / L[i][i];
    }
    for (int i = 5; i >= 0; i--) {
        double t = y[i];
        for (int q = i + 1; q < 6; q++)
            t -= L[q][i] * d[q];
        d[i] = t / L[i][i];
    }
    return true;
}

} // namespace

void fit_constrain(stp_gaussian2d& g)
{
    // Gaussian2dParams::convert_to_constrained_parameters (fitting.cpp:32-57)
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
}

void gaussian_jacobian(const double* p, double x, double y, double v, double* r, double* J)
{
    const Shape s = shape_of(p);
    *r = residual(p, s, Px{ x, y, v }, J);
}

FitResult lm_fit(const double* xs, const double* ys, const double* vs, size_t m, const double* start)
{
    std::vector<Px> px(m);
    for (size_t k = 0; k < m; k++)
        px[k] = Px{ xs[k], ys[k], vs[k] };
    FitResult fr;
    double x[6];
    std::memcpy(x, start, sizeof(x));
    double cost = half_cost(x, px);
    double mu = LM_MU0, nu = 2.0;
    int it = 0;
    fr.status = 1;
    while (it < LM_MAX_ITERS) {
        // normal equations at x
        double A[21] = {}, g[6] = {};
        const Shape s = shape_of(x);
        for (const Px& q : px) {
            double J[6];
            const double r = residual(x, s, q, J);
            int k = 0;
            for (int i = 0; i < 6; i++) {
                g[i] += J[i] * r;
                for (int j = 0; j <= i; j++, k++)
                    A[k] += J[i] * J[j];
            }
        }
        bool accepted = false, done = false;
        while (it < LM_MAX_ITERS && !accepted && !done) {
            it++;
            double M[21], rhs[6], d[6];
            std::memcpy(M, A, sizeof(M));
            for (int i = 0; i < 6; i++) {
                const int kk = i * (i + 1) / 2 + i;
                const double dg = std::fmin(std::fmax(A[kk], LM_DIAG_MIN), LM_DIAG_MAX);
                M[kk] = A[kk] + mu * dg;
                rhs[i] = -g[i];
            }
            if (!chol_solve6(M, rhs, d)) {
                mu *= nu;
                nu *= 2.0;
                done = mu > LM_MU_MAX;
                continue;
            }
            double xn[6];
            for (int i = 0; i < 6; i++)
                xn[i] = x[i] + d[i];
            const double cn = half_cost(xn, px);
            // predicted decrease of the Gauss-Newton model: -(d.g + d.A.d / 2)
            double dg = 0.0, dad = 0.0;
            int k = 0;
            for (int i = 0; i < 6; i++) {
                dg += d[i] * g[i];
                for (int j = 0; j <= i; j++, k++)
                    dad += (i == j ? 1.0 : 2.0) * d[i] * A[k] * d[j];
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
                std::memcpy(x, xn, sizeof(x));
                const bool small = actual <= LM_FTOL * cost || std::sqrt(nd) <= LM_XTOL * (std::sqrt(nx) + LM_XTOL);
                cost = cn;
                const double t = 2.0 * rho - 1.0;
                mu *= std::fmax(1.0 / 3.0, 1.0 - t * t * t);
                nu = 2.0;
                accepted = true;
                if (small || cost == 0.0) {
                    fr.status = 0;
                    done = true;
                }
            } else {
                mu *= nu;
                nu *= 2.0;
                if (mu > LM_MU_MAX) {
                    fr.status = 0; // no decrease possible at machine precision
                    done = true;
                }
            }
        }
        if (done)
            break;
    }
    fr.iterations = it;
    fr.fit = stp_gaussian2d{ x[0], x[1], x[2], x[3], x[4], x[5] };
    fr.half_cost = cost;
    fit_constrain(fr.fit);
    return fr;
}

std::vector<stp_fit_result> fit_islands(const double* data, int rows, int cols, const int32_t* labels,
    const stp_island* islands, size_t n)
{
    std::vector<stp_fit_result> out(n);
    const int h_shift = cols / 2, v_shift = rows / 2;
    // label -> island (islands cover a subset of the labels)
    int32_t maxlab = 0;
    for (size_t k = 0; k < n; k++)
        maxlab = std::max(maxlab, std::abs(islands[k].label_idx));
    std::vector<int32_t> isl_of(size_t(2) * maxlab + 1, -1);
    for (size_t k = 0; k < n; k++)
        isl_of[size_t(islands[k].label_idx + maxlab)] = int32_t(k);
    // bounding boxes in visual coordinates (sourcefind.cpp:449-666)
    for (size_t k = 0; k < n; k++)
        out[k].bounding_box = stp_bounding_box{ rows, -1, cols, -1 };
    for (int c = 0; c < cols; c++) {
        for (int r = 0; r < rows; r++) {
            const int32_t l = labels[size_t(r) + size_t(c) * rows];
            if (l == 0 || std::abs(l) > maxlab || isl_of[size_t(l + maxlab)] < 0)
                continue;
            stp_bounding_box& b = out[size_t(isl_of[size_t(l + maxlab)])].bounding_box;
            const int row = r < v_shift ? r + v_shift : r - v_shift;
            const int col = c < h_shift ? c + h_shift : c - h_shift;
            b.left = std::min(b.left, col);
            b.right = std::max(b.right, col);
            b.top = std::min(b.top, row);
            b.bottom = std::max(b.bottom, row);
        }
    }
    for (size_t k = 0; k < n; k++) {
        const stp_island& isl = islands[k];
        stp_fit_result& res = out[k];
        if (isl.num_samples < 3) {
            // sourcefind.cpp:724-728
            res.leastsq_fit = stp_gaussian2d{ isl.extremum_val, 0.0, 0.0, 0.0, 0.0, 0.0 };
            res.iterations = 0;
            res.status = 2;
            continue;
        }
        // residuals in the reference's order: columns left..right, rows top..bottom
        std::vector<double> xs, ys, vs;
        const stp_bounding_box& b = res.bounding_box;
        for (int i = b.left; i <= b.right; i++) {
            for (int j = b.top; j <= b.bottom; j++) {
                const int ii = i < h_shift ? i + h_shift : i - h_shift;
                const int jj = j < v_shift ? j + v_shift : j - v_shift;
                const size_t li = size_t(jj) + size_t(ii) * rows;
                if (labels[li] != isl.label_idx)
                    continue;
                xs.push_back(double(i));
                ys.push_back(double(j));
                vs.push_back(data[li]);
            }
        }
        const stp_gaussian2d& m = isl.moments_fit;
        const double start[6] = { m.amplitude, m.x_centre, m.y_centre, m.semimajor > 0.0 ? m.semimajor : 1.0,
            m.semiminor > 0.0 ? m.semiminor : 1.0, m.theta };
        const FitResult fr = lm_fit(xs.data(), ys.data(), vs.data(), xs.size(), start);
        res.leastsq_fit = fr.fit;
        res.iterations = fr.iterations;
        res.status = fr.status;
    }
    return out;
}

} // namespace oracle
