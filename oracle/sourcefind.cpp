/*
 * ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Source finder: src/stp/sourcefind/sourcefind.cpp:156-717 (FFTSHIFT=OFF),
 * connected-component labelling: src/stp/common/ccl.h:19-773,
 * constrained Gaussian parameters: src/stp/sourcefind/fitting.cpp:32-57.
 */
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "oracle.h"

namespace oracle {

namespace {

struct UF {
    std::vector<uint32_t> p;
    explicit UF(size_t n) : p(n)
    {
        std::iota(p.begin(), p.end(), 0u);
    }
    uint32_t find(uint32_t x)
    {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    void unite(uint32_t a, uint32_t b)
    {
        a = find(a);
        b = find(b);
        if (a == b)
            return;
        if (a < b)
            p[b] = a;
        else
            p[a] = b;
    }
};

} // namespace

/*
 * Connected components of the thresholded image with the reference's ids.
 *
 * Pixel classes (ccl.h:187,246): positive if value > thr_pos, otherwise
 * negative if find_negative and value < thr_neg.  Adjacency is the 8- (or 4-)
 * neighbourhood of the *visual* (fft-shifted) image without wrap-around: the
 * un-shifted imager output is read with visual row vr = (r + rows/2) % rows and
 * visual column c = (mc + cols/2) % cols (ccl.h:113-114,169,205-207).
 *
 * Final ids (ccl.h:453-480): the reference numbers components in the order of
 * their smallest provisional label, and provisional labels grow along its scan
 * (visual column c = 0..cols-1, memory row r = 0..rows-1 inside a column), so
 * the id of a component is the rank of its first pixel in that scan.  Positive
 * and negative components are ranked separately.
 *
 * The reference's 8-connected slab scan does not test the link between visual
 * (row rows/2, col c-1) and (row rows/2-1, col c) - memory rows 0 and rows-1 -
 * when column c is not the first column of a TBB slab (ccl.h:205-207 excludes
 * it from the scan; only the slab-border merge at :385-386 tests it).  That
 * makes the reference's result depend on the slab partition; this restatement
 * always links such pixels (ccl_label_sauf below reproduces the slab scan).
 */
void ccl_label(const double* img, int rows, int cols, double thr_pos, double thr_neg,
    bool find_negative, bool conn4, int32_t* labels, uint32_t* n_pos, uint32_t* n_neg)
{
    if ((rows % 2) || (cols % 2))
        throw std::invalid_argument("ccl: rows and cols must be even (ccl.h:116-117)");
    const int rshift = rows / 2, cshift = cols / 2;
    const size_t npx = size_t(rows) * cols;
    // class per visual pixel; key = c * rows + memory row  (scan order)
    std::vector<int8_t> cls(npx, 0);
    auto mem_index = [&](int c, int r) -> size_t { return size_t(r) + size_t((c + cshift) % cols) * rows; };
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++) {
            const double v = img[mem_index(c, r)];
            int8_t k = 0;
            if (v > thr_pos)
                k = 1;
            else if (find_negative && v < thr_neg)
                k = -1;
            cls[size_t(c) * rows + r] = k;
        }
    UF uf(npx);
    auto key_of_visual = [&](int c, int vr) -> uint32_t {
        const int r = (vr + rows - rshift) % rows; // memory row of visual row vr
        return uint32_t(size_t(c) * rows + r);
    };
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++) {
            const uint32_t k = uint32_t(size_t(c) * rows + r);
            const int8_t cl = cls[k];
            if (!cl)
                continue;
            const int vr = (r + rshift) % rows;
            auto link = [&](int nc, int nvr) {
                if (nc < 0 || nc >= cols || nvr < 0 || nvr >= rows)
                    return;
                const uint32_t nk = key_of_visual(nc, nvr);
                if (cls[nk] == cl)
                    uf.unite(k, nk);
            };
            link(c - 1, vr);
            link(c, vr - 1);
            if (!conn4) {
                link(c - 1, vr - 1);
                link(c - 1, vr + 1);
            }
        }
    // rank roots in key order
    std::vector<int32_t> id(npx, 0);
    int32_t np = 0, nn = 0;
    for (size_t k = 0; k < npx; k++) {
        if (!cls[k])
            continue;
        if (uf.find(uint32_t(k)) == k)
            id[k] = (cls[k] > 0) ? ++np : -(++nn);
    }
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++) {
            const size_t k = size_t(c) * rows + r;
            labels[mem_index(c, r)] = cls[k] ? id[uf.find(uint32_t(k))] : 0;
        }
    *n_pos = uint32_t(np);
    *n_neg = uint32_t(nn);
}

// --- literal restatement of the reference SAUF slab scan (ccl.h:107-484, :505-773) ---
static uint32_t sauf_find_root(const std::vector<uint32_t>& P, uint32_t i)
{
    uint32_t root = i;
    while (P[root] < root)
        root = P[root];
    return root;
}
static void sauf_set_root(std::vector<uint32_t>& P, uint32_t i, uint32_t root)
{
    while (P[i] < i) {
        const uint32_t j = P[i];
        P[i] = root;
        i = j;
    }
    P[i] = root;
}
static uint32_t sauf_union(std::vector<uint32_t>& P, uint32_t i, uint32_t j)
{
    uint32_t root = sauf_find_root(P, i);
    if (i != j) {
        const uint32_t rj = sauf_find_root(P, j);
        if (root > rj)
            root = rj;
        sauf_set_root(P, j, root);
    }
    sauf_set_root(P, i, root);
    return root;
}

void ccl_label_sauf(const double* img, int rows_i, int cols_i, double thr_pos, double thr_neg,
    bool find_negative, bool conn4, int slab_cols, int32_t* L, uint32_t* n_pos, uint32_t* n_neg)
{
    const uint32_t rows = uint32_t(rows_i), cols = uint32_t(cols_i);
    const uint32_t cshift = cols / 2, rshift = rows / 2;
    std::fill(L, L + size_t(rows) * cols, 0);
    const size_t Plen = size_t(cols) * rows / 2;
    std::vector<uint32_t> Pp(Plen, 0), Pn(Plen, 0);
    struct Slab {
        uint32_t col_start, lunique_start, lunique_p, lunique_n;
    };
    std::vector<Slab> slabs;
    auto mcol = [&](uint32_t c) { return c >= cshift ? c - cshift : c + cshift; };
    for (uint32_t cs = 0; cs < cols; cs += uint32_t(slab_cols)) {
        const uint32_t ce = std::min(cols, cs + uint32_t(slab_cols));
        const uint32_t lunique_start = cs * (rows / 2);
        Pp[lunique_start] = 0;
        Pn[lunique_start] = 0;
        uint32_t lp = lunique_start + 1, ln = lunique_start + 1;
        for (uint32_t ci = cs; ci < ce; ci++) {
            const uint32_t m = mcol(ci), mprev = (m == 0) ? cols - 1 : m - 1;
            int32_t* Lc = L + size_t(m) * rows;
            int32_t* Lp = L + size_t(mprev) * rows;
            const double* Ic = img + size_t(m) * rows;
            const bool left_valid = ci != cs;
            for (uint32_t r = 0; r < rows; r++) {
                const double v = Ic[r];
                for (int sgn = 1; sgn >= -1; sgn -= 2) {
                    if (sgn == 1 && !(v > thr_pos))
                        continue;
                    if (sgn == -1 && (!find_negative || v > thr_pos || !(v < thr_neg)))
                        continue;
                    std::vector<uint32_t>& P = (sgn == 1) ? Pp : Pn;
                    auto same = [&](int32_t x) { return sgn == 1 ? x > 0 : x < 0; };
                    const uint32_t rprev = (r == 0) ? r : r - 1;
                    const int32_t top = Lc[rprev], left = Lp[r];
                    const bool T_left = left_valid && same(left);
                    const bool T_top = (r == rshift || r == 0) ? false : same(top);
                    int32_t* cur = Lc + r;
                    if (conn4) {
                        if (T_left) {
                            *cur = T_top ? sgn * int32_t(sauf_union(P, uint32_t(sgn * top), uint32_t(sgn * left))) : left;
                        } else if (T_top) {
                            *cur = top;
                        } else {
                            *cur = sgn * int32_t(sgn == 1 ? lp : ln);
                            if (sgn == 1) {
                                P[lp] = lp;
                                lp++;
                            } else {
                                P[ln] = ln;
                                ln++;
                            }
                        }
                    } else {
                        const int32_t tl = Lp[rprev];
                        const uint32_t rnext = (r == rows - 1) ? r : r + 1;
                        const int32_t bl = Lp[rnext];
                        const bool T_tl = (r == rshift || r == 0) ? false : (left_valid && same(tl));
                        const bool T_bl = (r == rshift - 1 || r == rows - 1) ? false : (left_valid && same(bl));
                        if (T_left) {
                            *cur = left;
                        } else if (T_bl) {
                            if (T_tl)
                                *cur = sgn * int32_t(sauf_union(P, uint32_t(sgn * tl), uint32_t(sgn * bl)));
                            else if (T_top)
                                *cur = sgn * int32_t(sauf_union(P, uint32_t(sgn * top), uint32_t(sgn * bl)));
                            else
                                *cur = bl;
                        } else if (T_tl) {
                            *cur = tl;
                        } else if (T_top) {
                            *cur = top;
                        } else {
                            *cur = sgn * int32_t(sgn == 1 ? lp : ln);
                            if (sgn == 1) {
                                P[lp] = lp;
                                lp++;
                            } else {
                                P[ln] = ln;
                                ln++;
                            }
                        }
                    }
                }
            }
            // row border merging (ccl.h:308-342, :660-683)
            int32_t* cur = Lc;
            const int32_t top = Lc[rows - 1], tl = Lp[rows - 1];
            if (*cur > 0) {
                if (top > 0)
                    *cur = int32_t(sauf_union(Pp, uint32_t(*cur), uint32_t(top)));
                if (!conn4 && tl > 0 && top != tl)
                    *cur = int32_t(sauf_union(Pp, uint32_t(*cur), uint32_t(tl)));
            } else if (find_negative && *cur < 0) {
                if (top < 0)
                    *cur = -int32_t(sauf_union(Pn, uint32_t(-*cur), uint32_t(-top)));
                if (!conn4 && tl < 0 && top != tl)
                    *cur = -int32_t(sauf_union(Pn, uint32_t(-*cur), uint32_t(-tl)));
            }
        }
        slabs.push_back({ cs, lunique_start, lp, ln });
    }
    // column border merging (ccl.h:355-448, :696-737)
    for (const Slab& s : slabs) {
        const uint32_t ci = s.col_start, m = mcol(ci), mprev = (m == 0) ? cols - 1 : m - 1;
        int32_t* Lc = L + size_t(m) * rows;
        int32_t* Lp = L + size_t(mprev) * rows;
        const bool left_valid = ci != 0;
        for (uint32_t r = 0; r < rows; r++) {
            const int32_t cur = Lc[r];
            if (cur == 0 || (cur < 0 && !find_negative))
                continue;
            const int sgn = cur > 0 ? 1 : -1;
            std::vector<uint32_t>& P = (sgn == 1) ? Pp : Pn;
            auto same = [&](int32_t x) { return sgn == 1 ? x > 0 : x < 0; };
            const int32_t left = Lp[r];
            if (conn4) {
                if (left_valid && same(left))
                    Lc[r] = sgn * int32_t(sauf_union(P, uint32_t(sgn * cur), uint32_t(sgn * left)));
                continue;
            }
            const uint32_t rprev = (r == 0) ? rows - 1 : (r == rshift ? r : r - 1);
            const uint32_t rnext = (r == rows - 1) ? 0 : (r == rshift - 1 ? r : r + 1);
            const int32_t tl = Lp[rprev], bl = Lp[rnext];
            const bool T_left = left_valid && same(left);
            const bool T_tl = (r == rshift) ? false : (left_valid && same(tl));
            const bool T_bl = (r == rshift - 1) ? false : (left_valid && same(bl));
            if (T_tl)
                Lc[r] = sgn * int32_t(sauf_union(P, uint32_t(sgn * cur), uint32_t(sgn * tl)));
            if (T_left)
                Lc[r] = sgn * int32_t(sauf_union(P, uint32_t(sgn * cur), uint32_t(sgn * left)));
            if (T_bl)
                Lc[r] = sgn * int32_t(sauf_union(P, uint32_t(sgn * cur), uint32_t(sgn * bl)));
        }
    }
    // analysis (ccl.h:452-480)
    uint32_t k = 1;
    for (const Slab& s : slabs)
        for (uint32_t i = s.lunique_start + 1; i < s.lunique_p; i++)
            Pp[i] = (Pp[i] < i) ? Pp[Pp[i]] : k++;
    *n_pos = k - 1;
    k = 1;
    if (find_negative)
        for (const Slab& s : slabs)
            for (uint32_t i = s.lunique_start + 1; i < s.lunique_n; i++)
                Pn[i] = (Pn[i] < i) ? Pn[Pn[i]] : k++;
    *n_neg = k - 1;
    // final labelling (sourcefind.cpp:346-377)
    for (size_t i = 0; i < size_t(rows) * cols; i++) {
        if (L[i] > 0)
            L[i] = int32_t(Pp[L[i]]);
        else if (L[i] < 0)
            L[i] = -int32_t(Pn[-L[i]]);
    }
}

// IslandParams::estimate_moments_fit (sourcefind.cpp:691-717) and
// Gaussian2dParams::convert_to_constrained_parameters (fitting.cpp:32-57)
void estimate_moments_fit(stp_island& isl, double x_bar, double y_bar, double xx_bar, double yy_bar, double xy_bar,
    double rms_est, double analysis_n_sigma)
{
    const double working1 = (xx_bar + yy_bar) / 2.0;
    const double working2 = std::sqrt(((xx_bar - yy_bar) * (xx_bar - yy_bar) / 4) + xy_bar * xy_bar);
    const double trunc_semimajor_sq = working1 + working2;
    const double trunc_semiminor_sq = working1 - working2;
    const double pixel_threshold = analysis_n_sigma * rms_est;
    const double cutoff_ratio = isl.sign * isl.extremum_val / pixel_threshold;
    const double axes_scale_factor = 1.0 - std::log(cutoff_ratio) / (cutoff_ratio - 1.0);
    const double semimajor_est = std::sqrt(trunc_semimajor_sq / axes_scale_factor);
    const double semiminor_est = std::sqrt(trunc_semiminor_sq / axes_scale_factor);
    double theta = 0.5 * std::atan(2.0 * xy_bar / (xx_bar - yy_bar));
    if (theta * xy_bar < 0.0)
        theta += M_PI / 2.0;
    stp_gaussian2d g{ isl.extremum_val, x_bar, y_bar, semimajor_est, semiminor_est, theta };
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
    isl.moments_fit = g;
}

// SourceFindImage::SourceFindImage + _label_detection_islands (sourcefind.cpp:156-671)
SourceFindResult source_find(const double* data, int rows, int cols, const stp_sourcefind_pars& p)
{
    if (p.gaussian_fitting)
        throw std::invalid_argument("gaussian fitting is not restated");
    const size_t n = size_t(rows) * cols;
    SourceFindResult res;
    Stats st;
    switch (p.median_method) {
    case STP_MEDIAN_ZEROMEDIAN:
        st.median = 0.0;
        st.median_valid = true;
        break;
    case STP_MEDIAN_BINAPPROX:
        st = mat_median_binapprox(data, n);
        break;
    case STP_MEDIAN_BINMEDIAN:
        st = mat_binmedian(data, n);
        break;
    case STP_MEDIAN_NTHELEMENT:
        st.median = mat_median_exact(data, n);
        st.median_valid = true;
        break;
    default:
        throw std::invalid_argument("unknown median method");
    }
    const double bg = st.median;
    double rms;
    if (std::fabs(p.rms_est) > 0.0) {
        rms = p.rms_est;
    } else {
        if (!st.mean_valid || !st.sigma_valid) {
            const Stats t = mat_mean_and_stddev(data, n);
            st.mean = t.mean;
            st.sigma = t.sigma;
            st.mean_valid = st.sigma_valid = true;
        }
        rms = estimate_rms(data, n, 3, unsigned(p.sigma_clip_iters), st);
    }
    res.stats.bg_level = bg;
    res.stats.rms_est = rms;
    res.stats.mean = st.mean_valid ? st.mean : 0.0;
    res.stats.sigma = st.sigma_valid ? st.sigma : 0.0;

    const double a_pos = bg + p.analysis_n_sigma * rms;
    const double a_neg = bg - p.analysis_n_sigma * rms;
    const double d_pos = bg + p.detection_n_sigma * rms;
    const double d_neg = bg - p.detection_n_sigma * rms;
    res.label_map.assign(n, 0);
    uint32_t np = 0, nn = 0;
    ccl_label(data, rows, cols, a_pos, a_neg, p.find_negative_sources != 0, p.ccl_4connectivity != 0,
        res.label_map.data(), &np, &nn);
    res.stats.num_labels_pos = np;
    res.stats.num_labels_neg = nn;
    int32_t* L = res.label_map.data();

    // extrema: strict compare from the detection threshold; ties -> smallest memory index
    std::vector<double> ext_pos(np, d_pos), ext_neg(nn, d_neg);
    std::vector<size_t> idx_pos(np, 0), idx_neg(nn, 0);
    for (size_t i = 0; i < n; i++) {
        const int32_t l = L[i];
        if (l > 0) {
            if (data[i] > ext_pos[l - 1]) {
                ext_pos[l - 1] = data[i];
                idx_pos[l - 1] = i;
            }
        } else if (l < 0) {
            if (data[i] < ext_neg[-l - 1]) {
                ext_neg[-l - 1] = data[i];
                idx_neg[-l - 1] = i;
            }
        }
    }
    std::vector<int32_t> id_pos(np), id_neg(nn);
    uint32_t valid = np + nn;
    for (uint32_t i = 0; i < np; i++) {
        id_pos[i] = (ext_pos[i] > d_pos) ? int32_t(i + 1) : 0;
        if (!id_pos[i])
            valid--;
    }
    for (uint32_t i = 0; i < nn; i++) {
        id_neg[i] = (ext_neg[i] < d_neg) ? -int32_t(i + 1) : 0;
        if (!id_neg[i])
            valid--;
    }
    res.stats.num_valid_labels = valid;
    // moments and counts (sourcefind.cpp:458-569), visual coordinates
    std::vector<double> mp(size_t(6) * np, 0.0), mn(size_t(6) * nn, 0.0);
    std::vector<int32_t> cnt_pos(np, 0), cnt_neg(nn, 0);
    const int h_shift = cols / 2, v_shift = rows / 2;
    for (int i = 0; i < cols; i++) {
        for (int j = 0; j < rows; j++) {
            const size_t li = size_t(j) + size_t(i) * rows;
            const int32_t label = L[li];
            if (label == 0)
                continue;
            int idx;
            if (label > 0) {
                idx = label - 1;
                if (id_pos[idx] == 0) {
                    if (p.generate_labelmap)
                        L[li] = 0;
                    continue;
                }
            } else {
                idx = -label - 1;
                if (id_neg[idx] == 0) {
                    if (p.generate_labelmap)
                        L[li] = 0;
                    continue;
                }
            }
            const double val = data[li];
            const double y_idx = j < v_shift ? double(j + v_shift) : double(j - v_shift);
            const double x_idx = i < h_shift ? double(i + h_shift) : double(i - h_shift);
            const double x_bar = x_idx * val, y_bar = y_idx * val;
            const double xx = x_bar * x_idx, yy = y_bar * y_idx, xy = x_bar * y_idx;
            double* m = (label > 0) ? &mp[size_t(6) * idx] : &mn[size_t(6) * idx];
            if (label > 0)
                cnt_pos[idx]++;
            else
                cnt_neg[idx]++;
            m[0] += x_bar;
            m[1] += y_bar;
            m[2] += xx;
            m[3] += yy;
            m[4] += xy;
            m[5] += val;
        }
    }
    auto make_islands = [&](uint32_t count, const std::vector<int32_t>& ids, const std::vector<double>& ext,
                            const std::vector<size_t>& eidx, const std::vector<int32_t>& cnt, const std::vector<double>& m) {
        for (uint32_t l = 0; l < count; l++) {
            if (!ids[l])
                continue;
            const size_t li = eidx[l];
            const int r = int(li % size_t(rows)), c = int(li / size_t(rows));
            const int y_idx = r < v_shift ? r + v_shift : r - v_shift;
            const int x_idx = c < h_shift ? c + h_shift : c - h_shift;
            if (cnt[l] < p.source_min_area)
                continue;
            stp_island isl{};
            isl.label_idx = ids[l];
            isl.sign = ids[l] < 0 ? -1 : 1;
            isl.extremum_y_idx = y_idx;
            isl.extremum_x_idx = x_idx;
            isl.num_samples = cnt[l];
            isl.extremum_val = ext[l];
            const double* mm = &m[size_t(6) * l];
            const double xb = mm[0] / mm[5], yb = mm[1] / mm[5];
            const double xxb = mm[2] / mm[5] - xb * xb;
            const double yyb = mm[3] / mm[5] - yb * yb;
            const double xyb = mm[4] / mm[5] - xb * yb;
            estimate_moments_fit(isl, xb, yb, xxb, yyb, xyb, rms, p.analysis_n_sigma);
            res.islands.push_back(isl);
        }
    };
    make_islands(np, id_pos, ext_pos, idx_pos, cnt_pos, mp);
    make_islands(nn, id_neg, ext_neg, idx_neg, cnt_neg, mn);
    res.stats.num_islands = uint32_t(res.islands.size());
    return res;
}

} // namespace oracle
