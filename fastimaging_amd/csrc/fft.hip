// Hand-written Stockham FFTs for gfx950 with the semantics of the reference's
// FFTW calls (src/stp/common/fft.cpp:45-135 c2r, :227-323 c2c).
//
// Layout: a half-plane grid is (n/2+1) rows (v) x N columns (u), column-major,
// complex.  The 2-D c2r runs
//   U1, U2 : four-step c2c along u (N = L1 * L2), in place; element (v, x) of
//            the transformed grid sits in column slot (x % L1) * L2 + x / L1;
//   V      : per output column x, c2r along v of length n via one complex FFT
//            of length n/2 (even/odd packing); the imaginary parts of the v=0
//            and v=n/2 rows are ignored, as FFTW's c2r does.  Normalisation,
//            padding crop and grid correction of the imager are fused here.
// A real grid (the imager's sampling grid without W-projection) is stored packed,
// two real columns per complex one: U1, U2 run at half the length, and the V pass
// splits each column pair into the two u-transformed columns first (fft_v_c2r_real).
// Each transform runs in LDS with 64-wide waves (radix-8 Stockham stages after
// one radix-2 or radix-4 stage); a workgroup holds NT transforms of length L,
// TPT = 256/NT threads per transform.
#include <type_traits>

#include "common.h"
#include "fft.h"

namespace {

using stp_hip::NormSpec;
using stp_hip::OutPtrs;

constexpr int WG = 256;

template <int L>
struct Log2 {
    static constexpr int value = 1 + Log2<L / 2>::value;
};
template <>
struct Log2<1> {
    static constexpr int value = 0;
};

template <typename T>
__device__ __forceinline__ cplx<T> load_tw(const T* tw, int idx, bool inverse)
{
    cplx<T> w = { tw[2 * idx], tw[2 * idx + 1] };
    if (!inverse)
        w.y = -w.y;
    return w;
}

// Twiddle sources of lds_fft: tw(e) = exp(+-2 pi i e / L) for the length-L
// transform (+ for the inverse).
// From a global table of exp(+2 pi i m / (L * tws)).
template <typename T>
struct GlobalTw {
    const T* tw;
    int tws;
    bool inverse;
    __device__ cplx<T> operator()(int e) const { return load_tw(tw, e * tws, inverse); }
};

// exp(+-2 pi i m / n) from an LDS quarter table q[r] = exp(+2 pi i r / n), r < n/4
// (m = k n/4 + r: the entry times i^k, an exact rotation).
template <typename T>
__device__ __forceinline__ cplx<T> quarter_tw(const cplx<T>* q, int qlog, int m, bool inverse)
{
    const int k = m >> qlog;
    cplx<T> w = q[m & ((1 << qlog) - 1)];
    if (k & 1)
        w = { -w.y, w.x };
    if (k & 2)
        w = { -w.x, -w.y };
    if (!inverse)
        w.y = -w.y;
    return w;
}

// exp(+2 pi i m / n) from two small LDS tables: r = m mod n/4 = a 2^blog + b,
// W^r = hi[a] * lo[b] (hi[a] = exp(2 pi i a 2^blog / n), lo[b] = exp(2 pi i b / n)),
// then the exact rotation by i^(m div n/4).
template <typename T>
__device__ __forceinline__ cplx<T> split_tw(const cplx<T>* hi, const cplx<T>* lo, int qlog, int blog, int m)
{
    const int k = m >> qlog, r = m & ((1 << qlog) - 1);
    cplx<T> w = cmul(hi[r >> blog], lo[r & ((1 << blog) - 1)]);
    if (k & 1)
        w = { -w.y, w.x };
    if (k & 2)
        w = { -w.x, -w.y };
    return w;
}

// Twiddles of the c2r pass (inverse sign, table of n = 2M): tw(e) = exp(2 pi i 2e / n)
// for the length-M transform, at(m) = exp(2 pi i m / n) for the packing.  SPLIT: the
// two small tables of split_tw at t (lo, then hi); else a quarter table at t.
template <typename T, bool SPLIT>
struct C2rTw {
    const cplx<T>* t;
    int qlog, blog;
    __device__ cplx<T> at(int m) const
    {
        if constexpr (SPLIT)
            return split_tw(t + (1 << blog), t, qlog, blog, m);
        else
            return quarter_tw(t, qlog, m, true);
    }
    __device__ cplx<T> operator()(int e) const { return at(2 * e); }
};

// From an LDS table t[e] = exp(+2 pi i e / L), e < L.
template <typename T>
struct LdsTw {
    const cplx<T>* t;
    bool inverse;
    __device__ cplx<T> operator()(int e) const
    {
        cplx<T> w = t[e];
        if (!inverse)
            w.y = -w.y;
        return w;
    }
};

template <typename T>
__device__ __forceinline__ void bfly2(cplx<T>* v)
{
    const cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

// Radix-4 DFT, y_k = sum_r v_r w^{rk}, w = exp(+-i pi/2).
template <typename T>
__device__ __forceinline__ void bfly4(cplx<T>* v, bool inverse)
{
    const cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const cplx<T> a2 = cadd(v[1], v[3]);
    const cplx<T> d = csub(v[1], v[3]);
    // a3 = w * d with w = +i (inverse) or -i (forward)
    const cplx<T> a3 = inverse ? cplx<T>{ -d.y, d.x } : cplx<T>{ d.y, -d.x };
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

// Radix-8 DFT, y_k = sum_r v_r w^{rk}, w = exp(+-i pi/4): y_{2k} is the radix-4 DFT
// of a_r = v_r + v_{r+4}, y_{2k+1} that of b_r = (v_r - v_{r+4}) w^r (r < 4).
template <typename T>
__device__ __forceinline__ void bfly8(cplx<T>* v, bool inverse)
{
    const T h = T(0.70710678118654752440084436210484904L);  // cos(pi/4)
    cplx<T> a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        a[r] = cadd(v[r], v[r + 4]);
        b[r] = csub(v[r], v[r + 4]);
    }
    // w = (h, +-h), w^2 = +-i, w^3 = (-h, +-h)
    const cplx<T> d1 = b[1], d2 = b[2], d3 = b[3];
    b[1] = inverse ? cplx<T>{ h * (d1.x - d1.y), h * (d1.y + d1.x) } : cplx<T>{ h * (d1.x + d1.y), h * (d1.y - d1.x) };
    b[2] = inverse ? cplx<T>{ -d2.y, d2.x } : cplx<T>{ d2.y, -d2.x };
    b[3] = inverse ? cplx<T>{ -h * (d3.x + d3.y), h * (d3.x - d3.y) } : cplx<T>{ h * (d3.y - d3.x), -h * (d3.x + d3.y) };
    bfly4(a, inverse);
    bfly4(b, inverse);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[2 * k] = a[k];
        v[2 * k + 1] = b[k];
    }
}

// One Stockham stage of radix R on a length-L transform held at buf.
template <typename T, int L, int TPT, int R, typename TW>
__device__ __forceinline__ void stockham_stage(cplx<T>* buf, int lt, int Ns, const TW& tw, bool inverse)
{
    constexpr int NB = L / R;
    constexpr int BPT = (NB + TPT - 1) / TPT;
    cplx<T> v[BPT][R];
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        const int j = lt + b * TPT;
        if (j < NB) {
#pragma unroll
            for (int r = 0; r < R; r++)
                v[b][r] = buf[j + r * NB];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        const int j = lt + b * TPT;
        if (j < NB) {
            const int k = j & (Ns - 1);
            if (Ns > 1) {
                const int step = L / (Ns * R);
                if constexpr (R == 8) {
                    // three table twiddles, the other four as their products (a few
                    // ulp, fewer LDS loads)
                    const cplx<T> w1 = tw(k * step), w2 = tw(2 * k * step), w4 = tw(4 * k * step);
                    const cplx<T> w3 = cmul(w1, w2);
                    v[b][1] = cmul(v[b][1], w1);
                    v[b][2] = cmul(v[b][2], w2);
                    v[b][3] = cmul(v[b][3], w3);
                    v[b][4] = cmul(v[b][4], w4);
                    v[b][5] = cmul(v[b][5], cmul(w4, w1));
                    v[b][6] = cmul(v[b][6], cmul(w4, w2));
                    v[b][7] = cmul(v[b][7], cmul(w4, w3));
                } else {
#pragma unroll
                    for (int r = 1; r < R; r++)
                        v[b][r] = cmul(v[b][r], tw(r * k * step));
                }
            }
            if constexpr (R == 8)
                bfly8(v[b], inverse);
            else if constexpr (R == 4)
                bfly4(v[b], inverse);
            else
                bfly2(v[b]);
            const int idxD = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++)
                buf[idxD + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

// In-place unnormalised FFT of length L at buf, TPT cooperating threads (lt in [0, TPT)).
// tw: a twiddle source (exponents below 7L/8); inverse selects the + sign.
template <typename T, int L, int TPT, typename TW>
__device__ __forceinline__ void lds_fft(cplx<T>* buf, int lt, const TW& tw, bool inverse)
{
    // radix-8 stages after one radix-2 or radix-4 stage for the remainder of log2 L
    // mod 3 (fewer passes through LDS and barriers than radix 4 alone)
    constexpr int LOG = Log2<L>::value;
    int Ns = 1;
    if constexpr (LOG % 3 == 1) {
        stockham_stage<T, L, TPT, 2>(buf, lt, Ns, tw, inverse);
        Ns = 2;
    } else if constexpr (LOG % 3 == 2) {
        stockham_stage<T, L, TPT, 4>(buf, lt, Ns, tw, inverse);
        Ns = 4;
    }
    if constexpr (L >= 8) {
        for (; Ns < L; Ns *= 8)
            stockham_stage<T, L, TPT, 8>(buf, lt, Ns, tw, inverse);
    }
}

template <int L>
struct Shape {
    static constexpr int TPT = (L / 4 < 1) ? 1 : ((L / 4 > WG) ? WG : L / 4);
    static constexpr int NT = WG / TPT;
    static constexpr int PAD = 1;
};
// u passes: two radix-4 butterflies per thread and stage, so a workgroup holds
// twice the rows (longer contiguous runs of each grid column)
template <int L>
struct UShape {
    static constexpr int TPT = (L / 8 < 1) ? 1 : ((L / 8 > WG) ? WG : L / 8);
    static constexpr int NT = WG / TPT;
    static constexpr int PAD = 1;
};

// U1: for each u2 and block of NT rows, FFT over u1 (stride L2 columns), then
// multiply by exp(+-2 pi i k1 u2 / N).
template <typename T, int L1>
__global__ __launch_bounds__(WG) void fft_u1(cplx<T>* grids, int rows, int N, int L2, size_t grid_elems,
    const T* tw, int inverse)
{
    using S = UShape<L1>;
    __shared__ cplx<T> buf[S::NT * (L1 + S::PAD)];
    __shared__ cplx<T> s_tw[L1], s_post[L1];  // exp(+2 pi i e / L1); exp(+2 pi i k1 u2 / N)
    const int u2 = blockIdx.x;
    const int v0 = blockIdx.y * S::NT;
    cplx<T>* g = grids + size_t(blockIdx.z) * grid_elems;
    const int tid = threadIdx.x;
    // NT * L1 / WG elements per thread, all loads in flight before the LDS stores
    constexpr int EPT = S::NT * L1 / WG;
    static_assert(EPT * WG == S::NT * L1, "whole elements per thread");
    cplx<T> xv[EPT];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int idx = tid + e * WG;
        const int vv = idx % S::NT, u1 = idx / S::NT;
        const int v = v0 + vv;
        xv[e] = { T(0), T(0) };
        if (v < rows)
            xv[e] = g[size_t(v) + size_t(u1 * L2 + u2) * rows];
    }
    for (int e = tid; e < L1; e += WG) {
        s_tw[e] = load_tw(tw, e * (N / L1), true);
        s_post[e] = load_tw(tw, (e * u2) & (N - 1), true);
    }
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int idx = tid + e * WG;
        buf[(idx % S::NT) * (L1 + S::PAD) + idx / S::NT] = xv[e];
    }
    __syncthreads();
    const int t = tid / S::TPT, lt = tid % S::TPT;
    lds_fft<T, L1, S::TPT>(buf + t * (L1 + S::PAD), lt, LdsTw<T>{ s_tw, inverse != 0 }, inverse != 0);
    for (int idx = tid; idx < S::NT * L1; idx += WG) {
        const int vv = idx % S::NT, k1 = idx / S::NT;
        const int v = v0 + vv;
        if (v < rows) {
            cplx<T> x = buf[vv * (L1 + S::PAD) + k1];
            x = cmul(x, LdsTw<T>{ s_post, inverse != 0 }(k1));
            g[size_t(v) + size_t(k1 * L2 + u2) * rows] = x;
        }
    }
}

// U2: for each k1 and block of rows, FFT over u2 (contiguous column slots).
template <typename T, int L2>
__global__ __launch_bounds__(WG) void fft_u2(cplx<T>* grids, int rows, int N, int L1, size_t grid_elems,
    const T* tw, int inverse)
{
    using S = UShape<L2>;
    __shared__ cplx<T> buf[S::NT * (L2 + S::PAD)];
    __shared__ cplx<T> s_tw[L2];  // exp(+2 pi i e / L2)
    const int k1 = blockIdx.x;
    const int v0 = blockIdx.y * S::NT;
    cplx<T>* g = grids + size_t(blockIdx.z) * grid_elems;
    const int tid = threadIdx.x;
    constexpr int EPT = S::NT * L2 / WG;
    static_assert(EPT * WG == S::NT * L2, "whole elements per thread");
    cplx<T> xv[EPT];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int idx = tid + e * WG;
        const int vv = idx % S::NT, u2 = idx / S::NT;
        const int v = v0 + vv;
        xv[e] = { T(0), T(0) };
        if (v < rows)
            xv[e] = g[size_t(v) + size_t(k1 * L2 + u2) * rows];
    }
    for (int e = tid; e < L2; e += WG)
        s_tw[e] = load_tw(tw, e * (N / L2), true);
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int idx = tid + e * WG;
        buf[(idx % S::NT) * (L2 + S::PAD) + idx / S::NT] = xv[e];
    }
    __syncthreads();
    const int t = tid / S::TPT, lt = tid % S::TPT;
    lds_fft<T, L2, S::TPT>(buf + t * (L2 + S::PAD), lt, LdsTw<T>{ s_tw, inverse != 0 }, inverse != 0);
    for (int idx = tid; idx < S::NT * L2; idx += WG) {
        const int vv = idx % S::NT, k2 = idx / S::NT;
        const int v = v0 + vv;
        if (v < rows)
            g[size_t(v) + size_t(k1 * L2 + k2) * rows] = buf[vv * (L2 + S::PAD) + k2];
    }
}

// V: c2r along v of length n = 2M for output column x (natural order), reading
// column slot (x % L1) * L2 + x / L1.  Output value at padded (y, x) is
//   z(y, x) * factor / (gcf[ii] * gcf[jj])   (gcf optional)
// written to the cropped image_size x image_size output, as in
// normalise_image_beam_result_1D (imager.h:96-213).
template <typename T, typename OT, int M>
__global__ __launch_bounds__(WG) void fft_v_c2r(const cplx<T>* grids, size_t grid_elems, int N, int L1, int L2,
    const T* tw_n, OutPtrs<OT> outs, NormSpec ns)
{
    using S = Shape<M>;
    constexpr int n = 2 * M;
    constexpr int rows = M + 1;
    constexpr int QN = (n / 4 > 0) ? n / 4 : 1;  // quarter wave: r < n/4
    constexpr int QLOG = Log2<QN>::value;
    // two small tables (a complex multiply per twiddle) where the smaller LDS footprint
    // lets more workgroups share a CU (FP32: 80 -> 74 us at c2; FP64 from n = 2048: 68 ->
    // 58 us at c2); the quarter table for short FP64 columns
    constexpr bool SPLIT = std::is_same<T, float>::value || M >= 1024;
    constexpr int BLOG = QLOG < 5 ? QLOG : 5;
    constexpr int NTW = SPLIT ? (1 << BLOG) + (QN >> BLOG) : QN;
    constexpr int KI = (M / 2) / S::TPT + 1;      // pack pairs (k, M - k) per thread
    __shared__ cplx<T> buf[S::NT * (M + S::PAD)];
    __shared__ cplx<T> s_tw[NTW];
    const int tid = threadIdx.x;
    for (int r = tid; r < NTW; r += WG)
        s_tw[r] = load_tw(tw_n, (SPLIT && r >= (1 << BLOG)) ? (r - (1 << BLOG)) << BLOG : r, true);
    const C2rTw<T, SPLIT> tw{ s_tw, QLOG, BLOG };
    const int t = tid / S::TPT, lt = tid % S::TPT;
    const int gi = blockIdx.y;
    const cplx<T>* g = grids + size_t(gi) * grid_elems;
    cplx<T>* b = buf + t * (M + S::PAD);
    const int hi = ns.image_size / 2, hp = n / 2;
    const double* tp = ns.total_tab ? ns.total_tab[gi] : ns.total_ptr;
    const bool scale = tp != nullptr;
    const double factor = scale ? (1.0 / tp[0]) : 1.0;
    const int x = blockIdx.x * S::NT + t;
    const bool active = x < N;
    const int slot = active ? ((x % L1) * L2 + x / L1) : 0;
    const cplx<T>* col = g + size_t(slot) * rows;
    // the column's terms first (all loads in flight together)
    cplx<T> av[KI], cv[KI];
#pragma unroll
    for (int i = 0; i < KI; i++) {
        const int k = lt + i * S::TPT;
        if (active && k <= M / 2) {
            av[i] = col[k];
            cv[i] = col[M - k];
        }
    }
    __syncthreads();  // the table is ready
    // pack Z[k] = E[k] + i O[k]; thread handles the pair (k, M - k)
#pragma unroll
    for (int i = 0; i < KI; i++) {
        const int k = lt + i * S::TPT;
        if (!active || k > M / 2)
            continue;
        const int km = M - k;
        const cplx<T> a = av[i], c = cv[i];
        if (k == 0) {
            b[0] = { a.x + c.x, a.x - c.x }; // imaginary parts of v=0 and v=n/2 ignored
            continue;
        }
        // E[k] = X[k] + conj(X[M-k]);  O[k] = (X[k] - conj(X[M-k])) W^k, W = exp(+2 pi i / n)
        const cplx<T> e1 = { a.x + c.x, a.y - c.y };
        const cplx<T> d1 = { a.x - c.x, a.y + c.y };
        const cplx<T> o1 = cmul(d1, tw.at(k));
        b[k] = { e1.x - o1.y, e1.y + o1.x };
        if (km != k) {
            const cplx<T> e2 = { c.x + a.x, c.y - a.y };
            const cplx<T> d2 = { c.x - a.x, c.y + a.y };
            const cplx<T> o2 = cmul(d2, tw.at(km));
            b[km] = { e2.x - o2.y, e2.y + o2.x };
        }
    }
    __syncthreads();
    lds_fft<T, M, S::TPT>(b, lt, tw, true);
    // crop mapping of the padded column x (imager.h:99-150)
    int j = -1;
    if (active) {
        if (x < hi)
            j = x;
        else if (x >= N - (ns.image_size - hi))
            j = x - N + ns.image_size;
    }
    if (j < 0)
        return;
    const int joff = (j < hi) ? j : j - ns.image_size;
    OT* out = (outs.tab ? outs.tab[gi] : outs.p[gi]) + size_t(j) * ns.image_size;
    if (ns.image_size == n && !ns.gcf && (reinterpret_cast<uintptr_t>(out) & (2 * sizeof(OT) - 1)) == 0) {
        // no crop, no correction, aligned column: output rows 2m, 2m + 1 are z.x, z.y,
        // stored as one pair
        using V2 = typename std::conditional<std::is_same<OT, float>::value, float2, double2>::type;
        for (int m = lt; m < M; m += S::TPT) {
            const cplx<T> z = b[m];
            const double v0 = scale ? double(z.x) * factor : double(z.x);
            const double v1 = scale ? double(z.y) * factor : double(z.y);
            V2 pr;
            pr.x = OT(v0);
            pr.y = OT(v1);
            reinterpret_cast<V2*>(out)[m] = pr;
        }
        return;
    }
    for (int m = lt; m < M; m += S::TPT) {
        const cplx<T> z = b[m];
        const int ys[2] = { 2 * m, 2 * m + 1 };
        const T vals[2] = { z.x, z.y };
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int y = ys[q];
            int i;
            if (y < hi)
                i = y;
            else if (y >= n - (ns.image_size - hi))
                i = y - n + ns.image_size;
            else
                continue;
            double val = scale ? double(vals[q]) * factor : double(vals[q]);
            if (ns.gcf) {
                const int ioff = (i < hi) ? i : i - ns.image_size;
                val = val / (ns.gcf[hp + ioff] * ns.gcf[hp + joff]);
            }
            out[i] = OT(val);
        }
    }
}

// V for a real beam grid stored packed (square n x n transform): the grid holds
// z(v, u') = G(v, 2u') + i G(v, 2u' + 1), u' < U = n/2, already transformed along u'
// (length U, slot order of the L1 x L2 split) into Z(v, k).  Real-FFT split: with
// s = (Z(k) + conj Z(U-k)) / 2 and d = -i e^{2 pi i k / n} (Z(k) - conj Z(U-k)) / 2,
// the u-transform of the real rows is H(v, k) = s + d and H(v, U-k) = conj(s - d).
// A workgroup slot takes the column pair (k, U - k), k <= U/2: one c2r along v per
// column, and the image column n - c of each column c other than 0 and U from the
// point symmetry of a real even transform, B(y, n - c) = B(n - y mod n, c).  Output
// normalised, cropped and corrected as fft_v_c2r.
// Column buffers of fft_v_c2r_real: PAR (two columns in at most 40 KB of LDS; more
// would cost occupancy: FP64 at n = 4096 went 59 -> 68 us) gives each
// column of a pair its own threads, so both c2r run at once; otherwise a slot runs
// the pair's two c2r in turn.
template <typename T, int M>
struct RealVShape {
    using S = Shape<M>;
    static constexpr bool PAR = S::NT >= 2 || 2 * (M + S::PAD) * sizeof(cplx<T>) <= 40 * 1024;
    static constexpr int SLOTS = PAR ? (S::NT >= 2 ? S::NT : 2) : S::NT;
    static constexpr int PAIRS = PAR ? SLOTS / 2 : S::NT;
    static constexpr int THREADS = SLOTS * S::TPT;
};

template <typename T, typename OT, int M>
__global__ __launch_bounds__(2 * WG) void fft_v_c2r_real(const cplx<T>* grids, size_t grid_elems, int L1, int L2,
    const T* tw_n, OutPtrs<OT> outs, NormSpec ns)
{
    using S = Shape<M>;
    using RS = RealVShape<T, M>;
    constexpr int n = 2 * M;
    constexpr int U = M;             // packed columns
    constexpr int rows = M + 1;
    constexpr int QN = (n / 4 > 0) ? n / 4 : 1;
    constexpr int QLOG = Log2<QN>::value;
    // FP64 keeps the quarter table here up to n = 8192 (the split tables took this pass
    // from 49 to 54 us at c2: its occupancy is bound by registers, not LDS)
    constexpr bool SPLIT = std::is_same<T, float>::value || M >= 8192;
    constexpr int BLOG = QLOG < 5 ? QLOG : 5;
    constexpr int NTW = SPLIT ? (1 << BLOG) + (QN >> BLOG) : QN;
    constexpr int KI = (M / 2) / S::TPT + 1;
    __shared__ cplx<T> buf[RS::SLOTS * (M + S::PAD)];
    __shared__ cplx<T> s_tw[NTW];
    const int tid = threadIdx.x;
    for (int r = tid; r < NTW; r += RS::THREADS)
        s_tw[r] = load_tw(tw_n, (SPLIT && r >= (1 << BLOG)) ? (r - (1 << BLOG)) << BLOG : r, true);
    const C2rTw<T, SPLIT> tw{ s_tw, QLOG, BLOG };
    const int t = tid / S::TPT, lt = tid % S::TPT;
    const int pair = RS::PAR ? t >> 1 : t;
    const int first = RS::PAR ? (t & 1) : 0;  // the member this slot runs first (PAR: its only one)
    const int gi = blockIdx.y;
    const cplx<T>* g = grids + size_t(gi) * grid_elems;
    cplx<T>* b = buf + t * (M + S::PAD);
    const int hi = ns.image_size / 2, hp = n / 2;
    const double* tp = ns.total_tab ? ns.total_tab[gi] : ns.total_ptr;
    const bool scale = tp != nullptr;
    const double factor = scale ? (1.0 / tp[0]) : 1.0;
    const int x = blockIdx.x * RS::PAIRS + pair;
    const bool active = x <= U / 2;
    const int xb = active ? (U - x) % U : 0;
    const cplx<T>* ca = g + size_t(active ? (x % L1) * L2 + x / L1 : 0) * rows;
    const cplx<T>* cc = g + size_t((xb % L1) * L2 + xb / L1) * rows;
    // both columns' terms at rows k and M - k first (all loads in flight together)
    cplx<T> a0[KI], a1[KI], c0[KI], c1[KI];
#pragma unroll
    for (int i = 0; i < KI; i++) {
        const int k = lt + i * S::TPT;
        if (active && k <= M / 2) {
            a0[i] = ca[k];
            a1[i] = ca[M - k];
            c0[i] = cc[k];
            c1[i] = cc[M - k];
        }
    }
    __syncthreads();  // the table is ready
    const cplx<T> w = tw.at(active ? x : 0);
    // H(k) = s + d (member 0), conj(s - d) (member 1)
    auto split = [&](cplx<T> a, cplx<T> c, int member) -> cplx<T> {
        const cplx<T> sv = { (a.x + c.x) * T(0.5), (a.y - c.y) * T(0.5) };
        const cplx<T> q = { (a.x - c.x) * T(0.5), (a.y + c.y) * T(0.5) };
        const cplx<T> wq = cmul(w, q);
        const cplx<T> dv = { wq.y, -wq.x };  // -i * w * q
        return member == 0 ? cplx<T>{ sv.x + dv.x, sv.y + dv.y } : cplx<T>{ sv.x - dv.x, -(sv.y - dv.y) };
    };
    auto store = [&](int xc, bool refl) {
        const int xo = refl ? n - xc : xc;  // padded output column
        int j = -1;
        if (xo < hi)
            j = xo;
        else if (xo >= n - (ns.image_size - hi))
            j = xo - n + ns.image_size;
        if (j < 0)
            return;
        const int joff = (j < hi) ? j : j - ns.image_size;
        OT* out = (outs.tab ? outs.tab[gi] : outs.p[gi]) + size_t(j) * ns.image_size;
        if (ns.image_size == n && !ns.gcf && (reinterpret_cast<uintptr_t>(out) & (2 * sizeof(OT) - 1)) == 0) {
            // no crop, no correction, aligned column: rows 2p, 2p + 1 as one pair; the
            // reflected column's pair p holds B(n - 2p mod n), B(n - 2p - 1)
            using V2 = typename std::conditional<std::is_same<OT, float>::value, float2, double2>::type;
            for (int p = lt; p < M; p += S::TPT) {
                const T v0 = refl ? b[(M - p) & (M - 1)].x : b[p].x;
                const T v1 = refl ? b[M - 1 - p].y : b[p].y;
                V2 pr;
                pr.x = OT(scale ? double(v0) * factor : double(v0));
                pr.y = OT(scale ? double(v1) * factor : double(v1));
                reinterpret_cast<V2*>(out)[p] = pr;
            }
            return;
        }
        for (int m = lt; m < M; m += S::TPT) {
            const cplx<T> z = b[m];
            const int ys[2] = { 2 * m, 2 * m + 1 };
            const T vals[2] = { z.x, z.y };
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int y = refl ? (n - ys[q]) & (n - 1) : ys[q];
                int i;
                if (y < hi)
                    i = y;
                else if (y >= n - (ns.image_size - hi))
                    i = y - n + ns.image_size;
                else
                    continue;
                double val = scale ? double(vals[q]) * factor : double(vals[q]);
                if (ns.gcf) {
                    const int ioff = (i < hi) ? i : i - ns.image_size;
                    val = val / (ns.gcf[hp + ioff] * ns.gcf[hp + joff]);
                }
                out[i] = OT(val);
            }
        }
    };
#pragma unroll
    for (int mi = 0; mi < (RS::PAR ? 1 : 2); mi++) {
        const int member = first + mi;
        const int xc = member == 0 ? x : U - x;
        const bool on = active && (member == 0 || U - x != x);
        // pack Z[k] = E[k] + i O[k] of this column's c2r (as fft_v_c2r); thread pair (k, M - k)
#pragma unroll
        for (int i = 0; i < KI; i++) {
            const int k = lt + i * S::TPT;
            if (!on || k > M / 2)
                continue;
            const int km = M - k;
            const cplx<T> a = split(a0[i], c0[i], member), c = split(a1[i], c1[i], member);
            if (k == 0) {
                b[0] = { a.x + c.x, a.x - c.x };
                continue;
            }
            const cplx<T> e1 = { a.x + c.x, a.y - c.y };
            const cplx<T> d1 = { a.x - c.x, a.y + c.y };
            const cplx<T> o1 = cmul(d1, tw.at(k));
            b[k] = { e1.x - o1.y, e1.y + o1.x };
            if (km != k) {
                const cplx<T> e2 = { c.x + a.x, c.y - a.y };
                const cplx<T> d2 = { c.x - a.x, c.y + a.y };
                const cplx<T> o2 = cmul(d2, tw.at(km));
                b[km] = { e2.x - o2.y, e2.y + o2.x };
            }
        }
        __syncthreads();
        lds_fft<T, M, S::TPT>(b, lt, tw, true);
        if (on) {
            store(xc, false);
            if (xc != 0 && xc != U)
                store(xc, true);
        }
        __syncthreads();  // the next member's pack overwrites the column
    }
}

// Column c2c of length L (contiguous) reading column slot of a four-step row
// transform and writing natural column x of out.
template <typename T, int L>
__global__ __launch_bounds__(WG) void fft_cols_c2c(const cplx<T>* in, int rows, int N, int L1, int L2,
    cplx<T>* out, const T* tw_rows, int inverse)
{
    using S = Shape<L>;
    __shared__ cplx<T> buf[S::NT * (L + S::PAD)];
    // blockIdx.y: array of the batch (rows x N elements each)
    in += size_t(blockIdx.y) * rows * N;
    out += size_t(blockIdx.y) * rows * N;
    const int tid = threadIdx.x;
    const int t = tid / S::TPT, lt = tid % S::TPT;
    const int x = blockIdx.x * S::NT + t;
    const bool active = x < N;
    const int slot = active ? ((x % L1) * L2 + x / L1) : 0;
    cplx<T>* b = buf + t * (L + S::PAD);
    constexpr int EPT = (L + S::TPT - 1) / S::TPT;  // loads in flight together
    cplx<T> xv[EPT];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int r = lt + e * S::TPT;
        if (active && r < L)
            xv[e] = in[size_t(slot) * rows + r];
    }
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const int r = lt + e * S::TPT;
        if (active && r < L)
            b[r] = xv[e];
    }
    __syncthreads();
    lds_fft<T, L, S::TPT>(b, lt, GlobalTw<T>{ tw_rows, 1, inverse != 0 }, inverse != 0);
    if (active)
        for (int r = lt; r < L; r += S::TPT)
            out[size_t(x) * rows + r] = b[r];
}

// N = L1 * L2 for the four-step u transform: `req` when it is a valid split
// (both factors in [2, 128]), otherwise the balanced one.
template <typename T>
void split_len(int N, int& L1, int& L2, int req = 0)
{
    if (req >= 2 && req <= 128 && (req & (req - 1)) == 0 && N % req == 0 && N / req >= 2 && N / req <= 128) {
        L1 = req;
        L2 = N / req;
        return;
    }
    int lg = 0;
    while ((1 << lg) < N)
        lg++;
    L1 = 1 << ((lg + 1) / 2);
    L2 = N / L1;
}

#define STP_POW2_CASES(X) X(2) X(4) X(8) X(16) X(32) X(64) X(128)

template <typename T>
void launch_u_passes(stp_ctx* ctx, cplx<T>* grids, int nbatch, int rows, int N, size_t grid_elems, const T* tw, bool inverse,
    int req_l1 = 0)
{
    int L1, L2;
    split_len<T>(N, L1, L2, req_l1);
    using namespace stp_hip;
    switch (L1) {
#define U1CASE(LL)                                                                                          \
    case LL:                                                                                                \
        launch(ctx, "fft_u1", fft_u1<T, LL>, dim3(L2, div_up(rows, UShape<LL>::NT), nbatch), dim3(WG), 0,   \
            grids, rows, N, L2, grid_elems, tw, int(inverse));                                              \
        break;
        STP_POW2_CASES(U1CASE)
#undef U1CASE
    default:
        throw Error(STP_ERR_UNSUPPORTED, "FFT length not supported: " + std::to_string(N));
    }
    switch (L2) {
#define U2CASE(LL)                                                                                          \
    case LL:                                                                                                \
        launch(ctx, "fft_u2", fft_u2<T, LL>, dim3(L1, div_up(rows, UShape<LL>::NT), nbatch), dim3(WG), 0,   \
            grids, rows, N, L1, grid_elems, tw, int(inverse));                                              \
        break;
        STP_POW2_CASES(U2CASE)
#undef U2CASE
    default:
        throw Error(STP_ERR_UNSUPPORTED, "FFT length not supported: " + std::to_string(N));
    }
}

} // namespace

namespace stp_hip {

template <typename T, typename OT>
void fft_c2r_2d(stp_ctx* ctx, cplx<T>* grids, int nbatch, int n_rows, int N, const OutPtrs<OT>& outs, const NormSpec& ns)
{
    STP_REQUIRE(N >= 4 && (N & (N - 1)) == 0, "FFT column count must be a power of two >= 4");
    STP_REQUIRE(n_rows >= 4 && (n_rows & (n_rows - 1)) == 0, "FFT row count must be a power of two >= 4");
    const int rows = n_rows / 2 + 1;
    const size_t grid_elems = size_t(rows) * N;
    const Twiddles& twu = twiddles_for(ctx, N);
    const T* tw = std::is_same<T, double>::value ? (const T*)twu.d : (const T*)twu.f;
    launch_u_passes<T>(ctx, grids, nbatch, rows, N, grid_elems, tw, true, ns.u_l1);
    int L1, L2;
    split_len<T>(N, L1, L2, ns.u_l1);
    const Twiddles& twv = twiddles_for(ctx, n_rows);
    const T* tvn = std::is_same<T, double>::value ? (const T*)twv.d : (const T*)twv.f;
    STP_REQUIRE(nbatch >= 1 && nbatch <= (outs.tab ? 65535 : 4), "1 to 4 grids per FFT launch (65535 with an output table)");
    const int M = n_rows / 2;
    switch (M) {
#define VCASE(MM)                                                                                              \
    case MM:                                                                                                   \
        launch(ctx, "fft_v_c2r", fft_v_c2r<T, OT, MM>, dim3(div_up(N, Shape<MM>::NT), nbatch), dim3(WG), 0,   \
            (const cplx<T>*)grids, grid_elems, N, L1, L2, tvn, outs, ns);                                      \
        break;
        VCASE(2) VCASE(4) VCASE(8) VCASE(16) VCASE(32) VCASE(64) VCASE(128) VCASE(256) VCASE(512) VCASE(1024)
        VCASE(2048) VCASE(4096) VCASE(8192)
#undef VCASE
    default:
        throw Error(STP_ERR_UNSUPPORTED, "c2r length not supported: " + std::to_string(n_rows));
    }
}

template <typename T, typename OT>
void fft_c2r_2d_real(stp_ctx* ctx, cplx<T>* grids, int nbatch, int N, const OutPtrs<OT>& outs, const NormSpec& ns)
{
    STP_REQUIRE(N >= 8 && (N & (N - 1)) == 0, "real-grid FFT size must be a power of two >= 8");
    STP_REQUIRE(nbatch >= 1 && nbatch <= (outs.tab ? 65535 : 4), "1 to 4 grids per FFT launch (65535 with an output table)");
    const int rows = N / 2 + 1, U = N / 2;
    const size_t grid_elems = size_t(rows) * U;
    const Twiddles& twu = twiddles_for(ctx, U);
    const T* tw = std::is_same<T, double>::value ? (const T*)twu.d : (const T*)twu.f;
    launch_u_passes<T>(ctx, grids, nbatch, rows, U, grid_elems, tw, true, 0);
    int L1, L2;
    split_len<T>(U, L1, L2, 0);
    const Twiddles& twv = twiddles_for(ctx, N);
    const T* tvn = std::is_same<T, double>::value ? (const T*)twv.d : (const T*)twv.f;
    const int M = N / 2;
    switch (M) {
#define RVCASE(MM)                                                                                                    \
    case MM:                                                                                                          \
        launch(ctx, "fft_v_c2r_real", fft_v_c2r_real<T, OT, MM>, dim3(div_up(U / 2 + 1, RealVShape<T, MM>::PAIRS), nbatch), \
            dim3(RealVShape<T, MM>::THREADS), 0, (const cplx<T>*)grids, grid_elems, L1, L2, tvn, outs, ns);            \
        break;
        RVCASE(4) RVCASE(8) RVCASE(16) RVCASE(32) RVCASE(64) RVCASE(128) RVCASE(256) RVCASE(512) RVCASE(1024)
        RVCASE(2048) RVCASE(4096) RVCASE(8192)
#undef RVCASE
    default:
        throw Error(STP_ERR_UNSUPPORTED, "real-grid c2r length not supported: " + std::to_string(N));
    }
}

template <typename T>
void fft_c2c_2d(stp_ctx* ctx, cplx<T>* data, cplx<T>* out, int N, bool forward, int nbatch)
{
    STP_REQUIRE(N >= 4 && (N & (N - 1)) == 0, "c2c size must be a power of two >= 4");
    STP_REQUIRE(nbatch >= 1 && nbatch <= 65535, "1 to 65535 arrays per c2c launch");
    const Twiddles& twn = twiddles_for(ctx, N);
    const T* tw = std::is_same<T, double>::value ? (const T*)twn.d : (const T*)twn.f;
    // rows: transform along columns index (u) for every row, in place (four-step)
    launch_u_passes<T>(ctx, data, nbatch, N, N, size_t(N) * N, tw, !forward);
    int L1, L2;
    split_len<T>(N, L1, L2);
    switch (N) {
#define CCASE(LL)                                                                                           \
    case LL:                                                                                                \
        launch(ctx, "fft_cols_c2c", fft_cols_c2c<T, LL>, dim3(div_up(N, Shape<LL>::NT), nbatch), dim3(WG), 0, \
            (const cplx<T>*)data, N, N, L1, L2, out, tw, int(!forward));                                    \
        break;
        CCASE(4) CCASE(8) CCASE(16) CCASE(32) CCASE(64) CCASE(128) CCASE(256) CCASE(512) CCASE(1024)
        CCASE(2048) CCASE(4096)
#undef CCASE
    default:
        throw Error(STP_ERR_UNSUPPORTED, "c2c size not supported: " + std::to_string(N));
    }
}

template void fft_c2r_2d<double, double>(stp_ctx*, cplx<double>*, int, int, int, const OutPtrs<double>&, const NormSpec&);
template void fft_c2r_2d<float, float>(stp_ctx*, cplx<float>*, int, int, int, const OutPtrs<float>&, const NormSpec&);
template void fft_c2r_2d<float, double>(stp_ctx*, cplx<float>*, int, int, int, const OutPtrs<double>&, const NormSpec&);
template void fft_c2r_2d_real<double, double>(stp_ctx*, cplx<double>*, int, int, const OutPtrs<double>&, const NormSpec&);
template void fft_c2r_2d_real<float, float>(stp_ctx*, cplx<float>*, int, int, const OutPtrs<float>&, const NormSpec&);
template void fft_c2c_2d<double>(stp_ctx*, cplx<double>*, cplx<double>*, int, bool, int);

} // namespace stp_hip
