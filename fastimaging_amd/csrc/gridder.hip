// Convolutional gridding for gfx950: the scatter-add of convolve_to_grid
// (src/stp/gridder/gridder.h:240-997) as a tile-binned, LDS-accumulated gather.
//
//  1. grid_count   : per visibility, the reference's pre-processing (half-plane
//                    conversion gridder.cpp:11-37, rint centre/fraction
//                    gridder.h:377-386, bounds check gridder.cpp:202-217, uv shift
//                    gridder.h:399-413, oversampled index gridder.cpp:219-241);
//                    counts the (visibility pass, grid tile) pairs it touches and
//                    sums the good weights in the canonical order.
//  2. tile_scan    : exclusive scan of the per-tile counts.
//  3. grid_scatter : writes one 40-byte entry per (pass, tile) pair, binned by tile.
//  4. grid_tiles   : one workgroup per TILE_R x TILE_C tile accumulates every
//                    entry's taps that fall inside it in LDS (ds_add_f64), then
//                    writes the whole tile once: no global atomics, no memset.
// Per tap the arithmetic is the reference's: kval = K * w, grid += vis * kval,
// beam += kval (gridder.h:854-858); the oversampled kernel is the separable
// product ky[row] * kx[col] of normalised 1-D kernels (gridder.h:84-102).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "gridder.h"
#include "fft.h"
#include "kernel_tables.h"
#include "reduce.h"

namespace stp_hip {
namespace {

constexpr int TILE_R = 64;  // grid rows (v) per tile: 64 contiguous complex per column
constexpr int TILE_C = 32;  // default grid columns (u) per tile; GridderArgs::tile_cols overrides
constexpr int GWG = 256;

struct GridGeom {
    int N, rows, half, s, ks, os;
    int halfplane, shift, exact, beam;
    int bpack = 0; // beam grid packed real (rows x N/2 complex: columns 2u', 2u'+1)
    int tiles_r, tiles_c;
    int tc;       // tile width in grid columns (TILE_R rows per tile)
    double scale;
    long long n;
    int kfunc;
    double kparam, kwg, kws;
    // W-projection: visibility k of the pass is perm[k] (|w| order) and lies in plane[k]
    int wproj = 0;
    const uint32_t* perm = nullptr;
    const uint16_t* plane = nullptr;
};

struct Entry {
    int gx, gy;   // kernel centre on the (shifted) grid for this pass
    int cp;       // cp_x | cp_y << 16 (oversampled kernel indices); W-projection: bit 31 = conjugate kernel
    int idx;      // visibility index << 1 | pass; W-projection: the W sub-kernel (plane, cp_y, cp_x)
    double vr, vi, w; // W-projection: vr, vi are the visibility times w
};
static_assert(sizeof(Entry) == 40, "entry layout");

struct Prep {
    int good, dup;
    int gx, gy, cpx, cpy;
    double fx, fy;
    double vr, vi, w;
    double wl;   // w (wavelengths) after the half-plane conversion (W-projection)
};

// The inputs of one visibility (loads only, so that several can be in flight).
struct VisIn {
    double u, v, vr, vi, w, wl;
};

__device__ __forceinline__ VisIn load_vis(const GridGeom& g, const double* uv, const double* vis, const double* wts, long long k)
{
    VisIn r;
    const long long i = g.perm ? (long long)g.perm[k] : k;
    r.u = uv[i];
    r.v = uv[g.n + i];
    r.vr = vis[2 * i];
    r.vi = vis[2 * i + 1];
    r.w = wts[i];
    r.wl = g.wproj ? uv[2 * g.n + i] : 0.0;
    return r;
}

__device__ __forceinline__ Prep prep_in(const GridGeom& g, const VisIn& r)
{
    Prep p;
    double u = r.u * g.scale;
    double v = r.v * g.scale;
    p.vr = r.vr;
    p.vi = r.vi;
    p.w = r.w;
    p.dup = 0;
    p.wl = r.wl;
    p.good = (fabs(u) < 1.0e9 && fabs(v) < 1.0e9) ? 1 : 0; // also rejects NaN
    if (!p.good)
        u = v = 0.0;
    if (g.halfplane) {
        if (v < 0.0) {
            p.dup = (v > -(g.s + 1)) ? 1 : 0;
            u = -u;
            v = -v;
            p.vi = -p.vi;
            p.wl = -p.wl; // gridder.cpp:39-66
        } else {
            p.dup = (v < (g.s + 1)) ? 1 : 0;
        }
    }
    const int ku = int(rint(u)), kv = int(rint(v));
    p.fx = u - ku;
    p.fy = v - kv;
    int kx = ku + g.half, ky = kv + g.half;
    if ((kx - g.s) <= 0 || (ky - g.s) <= 0 || (kx + g.s) >= g.N || (ky + g.s) >= g.N)
        p.good = 0;
    if (g.shift) {
        kx = (kx < g.half) ? kx + g.half : kx - g.half;
        ky = (ky < g.half) ? ky + g.half : ky - g.half;
    }
    p.gx = kx;
    p.gy = ky;
    const int rmax = g.os / 2;
    int cx = int(rint(p.fx * g.os)), cy = int(rint(p.fy * g.os));
    cx = cx > rmax ? rmax : (cx < -rmax ? -rmax : cx);
    cy = cy > rmax ? rmax : (cy < -rmax ? -rmax : cy);
    p.cpx = cx + rmax;
    p.cpy = cy + rmax;
    return p;
}

__device__ __forceinline__ Prep prep_vis(const GridGeom& g, const double* uv, const double* vis, const double* wts, long long k)
{
    return prep_in(g, load_vis(g, uv, vis, wts, k));
}

// Centre / kernel indices of pass `pass` (1 = conjugate copy, gridder.h:794-805).
__device__ __forceinline__ void pass_geom(const GridGeom& g, const Prep& p, int pass, int& gx, int& gy, int& cpx, int& cpy)
{
    gx = p.gx;
    gy = p.gy;
    cpx = p.cpx;
    cpy = p.cpy;
    if (pass) {
        gx = g.N - p.gx;
        gy = g.N - p.gy;
        if (g.os > 1) {
            cpx = g.os - p.cpx;
            cpy = g.os - p.cpy;
        }
    }
}

// Entry of visibility position k, pass `pass`.  The conjugate pass conjugates the
// visibility and negates w (gridder.h:794-805); a negative w selects the
// conjugate W-kernel (gridder.h:826-843).
__device__ __forceinline__ Entry make_entry(const GridGeom& g, const Prep& p, int pass, long long k)
{
    int gx, gy, cpx, cpy;
    pass_geom(g, p, pass, gx, gy, cpx, cpy);
    Entry e;
    e.gx = gx;
    e.gy = gy;
    e.cp = cpx | (cpy << 16);
    e.vr = p.vr;
    e.vi = pass ? -p.vi : p.vi;
    e.w = p.w;
    if (g.wproj) {
        const double wl = pass ? -p.wl : p.wl;
        if (wl < 0.0)
            e.cp |= int(0x80000000u);
        // the W tile kernel takes the sub-kernel index (plane, cp_y, cp_x) and the
        // visibility times its weight
        const int cs = (g.os / 2) * 2 + 1;
        e.idx = (int(g.plane[k]) * cs + cpy) * cs + cpx;
        e.vr = e.vr * p.w;
        e.vi = e.vi * p.w;
    } else {
        e.idx = int((k << 1) | pass);
    }
    return e;
}

// Grid row of kernel row i, or -1 if dropped.  Oversampled path: rows < 0 skipped,
// rows >= N wrapped (gridder.h:847-853); exact path wraps both ways (:977-983).
__device__ __forceinline__ int tap_row(const GridGeom& g, int gy, int i)
{
    int r = gy - g.s + i;
    if (r < 0) {
        if (!g.exact)
            return -1;
        r += g.N;
    }
    if (r >= g.N)
        r -= g.N;
    return r < g.rows ? r : -1;
}

// Grid column of kernel column j: (gx + j - s) mod N (gridder.h:807-930, :970-974).
__device__ __forceinline__ int tap_col(const GridGeom& g, int gx, int j)
{
    int c = gx - g.s + j;
    if (c < 0)
        c += g.N;
    if (c >= g.N)
        c -= g.N;
    return c;
}

// The tiles of one axis that hold kept taps of a footprint: three intervals of
// tile indices [lo[k], hi[k]] (empty when lo > hi), one per contiguous piece of
// the raw range [centre - s, centre - s + ks - 1]: below 0, inside, at or above N.
// Pieces of a small grid can share tiles; visit_tiles skips the repeats.
struct TileSpan {
    int lo[3], hi[3];
};

__host__ __device__ __forceinline__ void span_set(TileSpan& sp, int k, int lo, int hi, int w)
{
    const bool on = lo <= hi;
    sp.lo[k] = on ? lo / w : 1;
    sp.hi[k] = on ? hi / w : 0;
}

// Rows: the oversampled path skips rows < 0 and wraps rows >= N, the exact path
// also wraps rows < 0; rows beyond `last` (the half plane) are dropped.
// Columns wrap both ways (the same sets as tap_row / tap_col over every tap).
__host__ __device__ __forceinline__ void footprint_spans(int N, int last, int s, int ks, int exact, int tile_r, int tile_c,
    int gx, int gy, TileSpan& rs, TileSpan& cs)
{
    int a = gy - s, b = a + ks - 1;
    span_set(rs, 0, a + N, (a < 0 && exact) ? min(min(b, -1) + N, last) : -1, tile_r);
    span_set(rs, 1, max(a, 0), min(min(b, N - 1), last), tile_r);
    span_set(rs, 2, max(a, N) - N, (b >= N) ? min(b - N, last) : -1, tile_r);
    a = gx - s;
    b = a + ks - 1;
    span_set(cs, 0, a + N, (a < 0) ? min(b, -1) + N : -1, tile_c);
    span_set(cs, 1, max(a, 0), min(b, N - 1), tile_c);
    span_set(cs, 2, max(a, N) - N, (b >= N) ? b - N : -1, tile_c);
}

// Whether tile t lies in one of the intervals before k.
__host__ __device__ __forceinline__ bool span_seen(const TileSpan& sp, int k, int t)
{
    bool seen = false;
#pragma unroll
    for (int q = 0; q < 2; q++)
        seen = seen || (q < k && t >= sp.lo[q] && t <= sp.hi[q]);
    return seen;
}

// Calls f(tile column, tile row) once for every distinct tile of the footprint.
template <typename F>
__device__ __forceinline__ void visit_tiles(const GridGeom& g, int gx, int gy, F f)
{
    TileSpan rs, cs;
    footprint_spans(g.N, g.rows - 1, g.s, g.ks, g.exact, TILE_R, g.tc, gx, gy, rs, cs);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        for (int ct = cs.lo[a]; ct <= cs.hi[a]; ct++) {
            if (span_seen(cs, a, ct))
                continue;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                for (int rt = rs.lo[b]; rt <= rs.hi[b]; rt++)
                    if (!span_seen(rs, b, rt))
                        f(ct, rt);
            }
        }
    }
}

// Largest number of distinct tiles one footprint can touch along an axis of
// length N (kept indices 0..last), over every centre the gridder can produce.
int max_span_tiles(int N, int last, int s, int ks, int exact, int w, bool rows)
{
    int best = 1;
    for (int c = -1; c <= N + 1; c++) {
        TileSpan rs, cs;
        footprint_spans(N, last, s, ks, exact, w, w, rows ? 0 : c, rows ? c : 0, rs, cs);
        const TileSpan& sp = rows ? rs : cs;
        int cnt = 0;
        for (int k = 0; k < 3; k++)
            for (int t = sp.lo[k]; t <= sp.hi[k]; t++)
                if (!span_seen(sp, k, t))
                    cnt++;
        best = std::max(best, cnt);
    }
    return best;
}

// Memoised per geometry: the scan over every centre takes tens of microseconds
// of host time per call at N = 4096.
size_t footprint_tiles_max(const GridGeom& g)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int, int, int>, size_t> memo;
    const auto key = std::make_tuple(g.N, g.rows, g.s, g.ks, g.exact, g.tc);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = memo.find(key);
        if (it != memo.end())
            return it->second;
    }
    const size_t v = size_t(max_span_tiles(g.N, g.rows - 1, g.s, g.ks, g.exact, TILE_R, true)) *
        size_t(max_span_tiles(g.N, g.rows - 1, g.s, g.ks, g.exact, g.tc, false));
    std::lock_guard<std::mutex> lk(mu);
    memo[key] = v;
    return v;
}

// wpart: per block the canonical sum of the good weights and the good count.
__global__ __launch_bounds__(GWG) void grid_count_kernel(GridGeom g, const double* uv, const double* vis,
    const double* wts, uint32_t* counts, double* wpart)
{
#pragma clang fp contract(off)
    __shared__ double lds[2 * CANON_LANES];
    const long long base = (long long)blockIdx.x * CANON_BLOCK;
    double wsum = 0.0, nsum = 0.0;  // good weights, good count (exact in double)
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const long long i = base + (long long)j * CANON_LANES + threadIdx.x;
        if (i >= g.n)
            break;
        const Prep p = prep_vis(g, uv, vis, wts, i);
        if (!p.good)
            continue;
        wsum = wsum + p.w;
        nsum = nsum + 1.0;
        for (int pass = 0; pass <= p.dup; pass++) {
            int gx, gy, cpx, cpy;
            pass_geom(g, p, pass, gx, gy, cpx, cpy);
            visit_tiles(g, gx, gy, [&](int ct, int rt) { atomicAdd(&counts[ct * g.tiles_r + rt], 1u); });
        }
    }
    canon_tree(wsum, nsum, lds);
    if (threadIdx.x == 0) {
        wpart[2 * blockIdx.x] = wsum;
        wpart[2 * blockIdx.x + 1] = nsum;
    }
}

__global__ __launch_bounds__(GWG) void grid_scatter_kernel(GridGeom g, const double* uv, const double* vis,
    const double* wts, uint32_t* cursor, Entry* entries, uint32_t cap)
{
    const long long i = (long long)blockIdx.x * GWG + threadIdx.x;
    if (i >= g.n)
        return;
    const Prep p = prep_vis(g, uv, vis, wts, i);
    if (!p.good)
        return;
    for (int pass = 0; pass <= p.dup; pass++) {
        const Entry e = make_entry(g, p, pass, i);
        visit_tiles(g, e.gx, e.gy, [&](int ct, int rt) {
            const uint32_t slot = atomicAdd(&cursor[ct * g.tiles_r + rt], 1u);
            if (slot < cap)
                entries[slot] = e;
        });
    }
}

// Kernel functors on the device (exact gridding only; conv_func.cpp:10-202).
// Spheroidal function of the PSWF functor (conv_func.cpp:112-182): F. Schwab's
// rational approximation (Indirect Imaging, 1984), M = 6, alpha = 1; the same
// coefficients and evaluation order as the host table (kernel_tables.cpp).
__device__ double dev_spheroidal(double nu)
{
    constexpr double p0[5] = { 8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1 };
    constexpr double p1[5] = { 4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2 };
    constexpr double q0[3] = { 1.0, 8.212018e-1, 2.078043e-1 };
    constexpr double q1[3] = { 1.0, 9.599102e-1, 2.918724e-1 };
    nu = fabs(nu);
    const bool part = nu >= 0.75 && nu <= 1.0;
    const double nuend = (nu >= 0.0 && nu < 0.75) ? 0.75 : ((nu <= 1.0) ? 1.0 : 0.0);
    const double d = pow(nu, 2.0) - pow(nuend, 2.0);
    double top = part ? p1[0] : p0[0], bot = part ? q1[0] : q0[0];
    for (int k = 1; k < 5; k++)
        top += (part ? p1[k] : p0[k]) * pow(d, double(k));
    for (int k = 1; k < 3; k++)
        bot += (part ? q1[k] : q0[k]) * pow(d, double(k));
    return bot > 0.0 ? top / bot : 0.0;
}

__device__ double dev_kernel(const GridGeom& g, double r)
{
    double v;
    switch (g.kfunc) {
    case STP_KERNEL_TOPHAT:
        return (fabs(r) < g.kparam) ? 1.0 : 0.0;
    case STP_KERNEL_TRIANGLE:
        return fmax(1.0 - fabs(r) / g.kparam, 0.0);
    case STP_KERNEL_SINC: {
        double x = r * (M_PI / g.kws);
        if (x == 0.0)
            x = 1.0e-20;
        v = sin(x) / x;
        break;
    }
    case STP_KERNEL_GAUSSIAN:
        v = exp((-1.0 / (g.kwg * g.kwg)) * (r * r));
        break;
    case STP_KERNEL_PSWF: { // conv_func.cpp:185-196
        const double nu = r / g.kparam;
        v = dev_spheroidal(nu) * (1.0 - pow(nu, 2.0));
        break;
    }
    default: { // GaussianSinc
        double x = r * (M_PI / g.kws);
        if (x == 0.0)
            x = 1.0e-20;
        v = exp((-1.0 / (g.kwg * g.kwg)) * (r * r)) * (sin(x) / x);
        break;
    }
    }
    if (g.kparam > 0.0 && fabs(r) > g.kparam)
        v = 0.0;
    return v;
}

// Last canonical level of the (good weights, good count) block pairs: the sampling
// grid total 2 * sum(weights) and the good count.
// Also copies the number of gridding entries (the end of the last tile's range).
__global__ __launch_bounds__(CANON_LANES) void grid_total_kernel(const double* part, unsigned m, double* total,
    unsigned long long* ngood, const uint32_t* entries_end, unsigned long long* nentries)
{
    __shared__ double lds[2 * CANON_LANES];
    double a, b;
    canon_last_level(part, m, a, b, lds);
    if (threadIdx.x == 0) {
        total[0] = a * 2.0;
        if (ngood)
            *ngood = (unsigned long long)b;
        if (nentries)
            *nentries = *entries_end;
    }
}

// ---------------------------------------------------------------------------
// Binning by per-block LDS counting sort (no global atomics): block b takes
// visibilities [4096 b, 4096 b + 4096) and writes its entry count of every tile
// to counts[b * ntiles + tile] (one contiguous row per block).  A column scan over
// the blocks turns each row into the block's cursor inside every tile and gives
// the tile totals, whose exclusive scan is the per-tile entry ranges.
constexpr int BWG = 1024;
constexpr int VPB = CANON_BLOCK; // visibilities per binning block (also the canonical sum block)
constexpr int MAX_LDS_TILES = 24576;

__device__ __forceinline__ void bin_count_body(const GridGeom& g, const double* uv, const double* vis, const double* wts,
    uint32_t* counts, double* wpart, int bid)
{
#pragma clang fp contract(off)
    extern __shared__ uint32_t hist[];
    __shared__ double wl[VPB];
    __shared__ unsigned long long wng[BWG / 64];
    const int ntiles = g.tiles_r * g.tiles_c;
    for (int t = threadIdx.x; t < ntiles; t += BWG)
        hist[t] = 0;
    __syncthreads();
    const long long base = (long long)bid * VPB;
    unsigned long long ng = 0;
    VisIn in[VPB / BWG];
#pragma unroll
    for (int j = 0; j < VPB / BWG; j++) {
        const long long i = base + threadIdx.x + j * BWG;
        if (i < g.n)
            in[j] = load_vis(g, uv, vis, wts, i);
    }
#pragma unroll
    for (int j = 0; j < VPB / BWG; j++) {
        const int local = threadIdx.x + j * BWG;
        const long long i = base + local;
        double wgood = 0.0;
        if (i < g.n) {
            const Prep p = prep_in(g, in[j]);
            if (p.good) {
                wgood = p.w;
                ng++;
                for (int pass = 0; pass <= p.dup; pass++) {
                    int gx, gy, cpx, cpy;
                    pass_geom(g, p, pass, gx, gy, cpx, cpy);
                    visit_tiles(g, gx, gy, [&](int ct, int rt) { atomicAdd(&hist[ct * g.tiles_r + rt], 1u); });
                }
            }
        }
        wl[local] = wgood;
    }
    __syncthreads();
    // canonical sum of the good weights: lane t adds elements t, t+256, ..., t+15*256
    double a = 0.0;
    if (threadIdx.x < CANON_LANES)
        for (int j = 0; j < CANON_PER_LANE; j++)
            a = a + wl[threadIdx.x + j * CANON_LANES];
    __syncthreads();
    if (threadIdx.x < CANON_LANES)
        wl[threadIdx.x] = a;
    for (int off = CANON_LANES / 2; off >= 1; off >>= 1) {
        __syncthreads();
        if (threadIdx.x < off)
            wl[threadIdx.x] = wl[threadIdx.x] + wl[threadIdx.x + off];
    }
    // the block's good count (exact in double) beside its weight sum
    for (int off = 32; off >= 1; off >>= 1)
        ng += __shfl_down(ng, off, 64);
    if ((threadIdx.x & 63) == 0)
        wng[threadIdx.x >> 6] = ng;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0;
        for (int k = 0; k < BWG / 64; k++)
            c += wng[k];
        wpart[2 * bid] = wl[0];
        wpart[2 * bid + 1] = double(c);
    }
    for (int t = threadIdx.x; t < ntiles; t += BWG)
        counts[size_t(bid) * ntiles + t] = hist[t];
}

__global__ __launch_bounds__(BWG) void grid_bin_count_kernel(GridGeom g, const double* uv, const double* vis,
    const double* wts, uint32_t* counts, double* wpart)
{
    bin_count_body(g, uv, vis, wts, counts, wpart, int(blockIdx.x));
}

// Per tile t: counts[b][t] -> entries of tile t in blocks before b (in place); totals[t].
// A workgroup takes 64 tiles (lanes) x 16 slices of the blocks (waves).
__device__ __forceinline__ void colscan_body(uint32_t* counts, int nblk, int ntiles, uint32_t* totals, int bid)
{
    __shared__ uint32_t part[16][64];
    const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int t = bid * 64 + lane;
    const int per = (nblk + 15) / 16, b0 = min(nblk, sl * per), b1 = min(nblk, b0 + per);
    uint32_t s = 0;
    if (t < ntiles)
        for (int b = b0; b < b1; b++)
            s += counts[size_t(b) * ntiles + t];
    part[sl][lane] = s;
    __syncthreads();
    if (sl == 0) {
        uint32_t run = 0;
        for (int k = 0; k < 16; k++) {
            const uint32_t v = part[k][lane];
            part[k][lane] = run;
            run += v;
        }
        if (t < ntiles)
            totals[t] = run;
    }
    __syncthreads();
    if (t >= ntiles)
        return;
    uint32_t run = part[sl][lane];
    for (int bb = b0; bb < b1; bb += 16) {
        uint32_t v[16];
#pragma unroll
        for (int k = 0; k < 16; k++)
            v[k] = bb + k < b1 ? counts[size_t(bb + k) * ntiles + t] : 0u;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (bb + k < b1)
                counts[size_t(bb + k) * ntiles + t] = run;
            run += v[k];
        }
    }
}

__global__ __launch_bounds__(1024) void bin_colscan_kernel(uint32_t* counts, int nblk, int ntiles, uint32_t* totals)
{
    colscan_body(counts, nblk, ntiles, totals, int(blockIdx.x));
}

// Exclusive scan of the tile totals into out[0..ntiles) (LDS) by the BWG threads of a
// block, each taking a run of at most TSCAN_PER tiles; returns the total.
constexpr int TSCAN_PER = (MAX_LDS_TILES + BWG - 1) / BWG;
__device__ uint32_t scan_tile_totals(const uint32_t* totals, int ntiles, uint32_t* out, uint32_t* wsum)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (ntiles + BWG - 1) / BWG;
    const int i0 = min(ntiles, t * per), i1 = min(ntiles, i0 + per);
    uint32_t v[TSCAN_PER];
#pragma unroll
    for (int j = 0; j < TSCAN_PER; j++)
        v[j] = i0 + j < i1 ? totals[i0 + j] : 0u;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < TSCAN_PER; j++)
        s += v[j];
    uint32_t x = s;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    uint32_t total = 0;
    for (int k = 0; k < BWG / 64; k++) {
        total += wsum[k];
        if (k < w)
            x += wsum[k];
    }
    uint32_t run = x - s;
#pragma unroll
    for (int j = 0; j < TSCAN_PER; j++) {
        if (i0 + j < i1) {
            out[i0 + j] = run;
            run += v[j];
        }
    }
    return total;
}

// Block b's entries go to cursors tile offset + block offset; the tile offsets (the
// scan of the tile totals) are computed by every block in LDS, and block 0 also
// writes them to tile_off[0..ntiles] for the tile kernel.
__device__ __forceinline__ void bin_scatter_body(const GridGeom& g, const double* uv, const double* vis, const double* wts,
    const uint32_t* totals, const uint32_t* block_off, Entry* entries, uint32_t cap, uint32_t* tile_off, int bid)
{
    extern __shared__ uint32_t cur[];
    __shared__ uint32_t wsum[BWG / 64];
    const int ntiles = g.tiles_r * g.tiles_c;
    const uint32_t total = scan_tile_totals(totals, ntiles, cur, wsum);
    __syncthreads();
    if (bid == 0) {
        for (int t = threadIdx.x; t < ntiles; t += BWG)
            tile_off[t] = cur[t];
        if (threadIdx.x == 0)
            tile_off[ntiles] = total;
    }
    for (int t = threadIdx.x; t < ntiles; t += BWG)
        cur[t] += block_off[size_t(bid) * ntiles + t];
    __syncthreads();
    const long long base = (long long)bid * VPB;
    VisIn in[VPB / BWG];
#pragma unroll
    for (int j = 0; j < VPB / BWG; j++) {
        const long long i = base + threadIdx.x + j * BWG;
        if (i < g.n)
            in[j] = load_vis(g, uv, vis, wts, i);
    }
#pragma unroll
    for (int j = 0; j < VPB / BWG; j++) {
        const long long i = base + threadIdx.x + j * BWG;
        if (i >= g.n)
            break;
        const Prep p = prep_in(g, in[j]);
        if (!p.good)
            continue;
        for (int pass = 0; pass <= p.dup; pass++) {
            const Entry e = make_entry(g, p, pass, i);
            visit_tiles(g, e.gx, e.gy, [&](int ct, int rt) {
                const uint32_t slot = atomicAdd(&cur[ct * g.tiles_r + rt], 1u);
                if (slot < cap)
                    entries[slot] = e;
            });
        }
    }
}

__global__ __launch_bounds__(BWG) void grid_bin_scatter_kernel(GridGeom g, const double* uv, const double* vis,
    const double* wts, const uint32_t* totals, const uint32_t* block_off, Entry* entries, uint32_t cap, uint32_t* tile_off)
{
    bin_scatter_body(g, uv, vis, wts, totals, block_off, entries, cap, tile_off, int(blockIdx.x));
}

// ---------------------------------------------------------------------------
// Tile accumulation: one workgroup per TILE_R x TILE_C tile accumulates the taps
// of every entry binned to it in LDS, then writes the tile once (no global
// atomics, no memset).  Lane mapping for supports up to 31: a wave takes
// 64 / ks entries at a time, lane (entry e, kernel column j) walks the ks kernel
// rows, so one load of kx[j] serves ks taps and the ks LDS atomics of a lane are
// independent.  Wider kernels take one entry per wave, lanes striding columns.
//
// FP64: ds_add_f64 per channel.  FP32 (compensated): each channel keeps an
// unevaluated pair hi + lo.  The returning atomic add on hi yields the old value,
// so the rounding error of the sum it stored is known exactly (Knuth's TwoSum on
// old and x) and is itself added to lo.
constexpr int TWG = 512;
constexpr int TEB = 256;                  // entries staged per batch
constexpr int KV_DOUBLES = 3584;          // exact mode: per-entry kernel vectors (28 KB)

constexpr int RG = 4;                     // kernel rows per group of LDS atomics

// FP64 accumulation (ds_add_f64) of FP64 or FP32 taps.
struct AccF64 {
    using S = double;
    static constexpr int arrays = 3;      // re, im, beam
    template <int R, typename T>
    __device__ static void add_rows(double* base, int plane, const int* aa, const T* xr, const T* xi,
        const T* xb, bool beam)
    {
#pragma unroll
        for (int u = 0; u < R; u++) {
            if (aa[u] < 0)
                continue;
            unsafeAtomicAdd(&base[aa[u]], double(xr[u]));
            unsafeAtomicAdd(&base[plane + aa[u]], double(xi[u]));
            if (beam)
                unsafeAtomicAdd(&base[2 * plane + aa[u]], double(xb[u]));
        }
    }
    __device__ static double value(const double* base, int plane, int ch, int a) { return base[ch * plane + a]; }
};
// Error of the rounded sum s = fl(old + x) (TwoSum).
__device__ __forceinline__ float two_sum_err(float old, float x)
{
    const float sum = old + x;
    const float bp = sum - old;
    return (old - (sum - bp)) + (x - bp);
}
// Two-float accumulation of FP32 taps: hi planes (re, im, beam), then lo planes.
struct AccTwoFloat {
    using S = float;
    static constexpr int arrays = 6;
    template <int R, typename T>
    __device__ static void add_rows(float* base, int plane, const int* aa, const T* xr, const T* xi,
        const T* xb, bool beam)
    {
        float orr[R], oi[R], ob[R];
        // returning adds on hi: each yields the old value, so the stored sum is fl(old + x)
#pragma unroll
        for (int u = 0; u < R; u++) {
            orr[u] = oi[u] = ob[u] = 0.f;
            if (aa[u] < 0)
                continue;
            orr[u] = atomicAdd(&base[aa[u]], xr[u]);
            oi[u] = atomicAdd(&base[plane + aa[u]], xi[u]);
            if (beam)
                ob[u] = atomicAdd(&base[2 * plane + aa[u]], xb[u]);
        }
        // the exact rounding errors go to the lo planes
#pragma unroll
        for (int u = 0; u < R; u++) {
            if (aa[u] < 0)
                continue;
            const float er = two_sum_err(orr[u], xr[u]), ei = two_sum_err(oi[u], xi[u]);
            if (er != 0.f)
                atomicAdd(&base[3 * plane + aa[u]], er);
            if (ei != 0.f)
                atomicAdd(&base[4 * plane + aa[u]], ei);
            if (beam) {
                const float eb = two_sum_err(ob[u], xb[u]);
                if (eb != 0.f)
                    atomicAdd(&base[5 * plane + aa[u]], eb);
            }
        }
    }
    __device__ static float value(const float* base, int plane, int ch, int a)
    {
        return base[ch * plane + a] + base[(3 + ch) * plane + a];
    }
};

// Writes a tile's sums once: the vis grid complex; the beam grid complex (imaginary
// part 0) or, with g.bpack, packed real: cell (r, c) is component c & 1 of element
// (r, c / 2) of a rows x N/2 complex array (tile columns come in even pairs).
template <typename T, typename Acc>
__device__ __forceinline__ void write_tile(const GridGeom& g, const typename Acc::S* acc, int plane, int S, int r0, int c0,
    cplx<T>* vgrid, cplx<T>* bgrid)
{
    for (int k = threadIdx.x; k < g.tc * TILE_R; k += TWG) {
        const int lr = k % TILE_R, lc = k / TILE_R;
        const int r = r0 + lr, c = c0 + lc;
        if (r < g.rows && c < g.N) {
            const size_t o = size_t(r) + size_t(c) * g.rows;
            const int a = lc * S + lr;
            vgrid[o] = { T(Acc::value(acc, plane, 0, a)), T(Acc::value(acc, plane, 1, a)) };
            if (g.beam && !g.bpack)
                bgrid[o] = { T(Acc::value(acc, plane, 2, a)), T(0) };
        }
    }
    if (g.beam && g.bpack)
        for (int k = threadIdx.x; k < (g.tc / 2) * TILE_R; k += TWG) {
            const int lr = k % TILE_R, lp = k / TILE_R;
            const int r = r0 + lr, c = c0 + 2 * lp;
            if (r < g.rows && c < g.N) {
                const int a = 2 * lp * S + lr;
                bgrid[size_t(r) + size_t(c / 2) * g.rows] = { T(Acc::value(acc, plane, 2, a)), T(Acc::value(acc, plane, 2, a + S)) };
            }
        }
}

template <typename T, typename Acc>
__global__ __launch_bounds__(TWG) void grid_tiles_kernel(GridGeom g, const Entry* entries, const uint32_t* offsets, int ostride,
    uint32_t cap, const double* ktab, const double* uv, cplx<T>* vgrid, cplx<T>* bgrid)
{
#pragma clang fp contract(off)
    using AS = typename Acc::S;
    extern __shared__ double dyn[];
    __shared__ Entry ebuf[TEB];
    const int ks = g.ks, s = g.s;
    const int S = TILE_R + (ks <= 15 ? ks : 1);     // column stride of the LDS tile
    const int plane = g.tc * S;
    AS* acc = reinterpret_cast<AS*>(dyn);
    const int acc_bytes = Acc::arrays * plane * int(sizeof(AS));
    double* kv = dyn + (acc_bytes + 7) / 8;          // kernel table, or exact-mode vectors then sums
    const int cs = (g.os / 2) * 2 + 1;
    T* kt = reinterpret_cast<T*>(kv);
    double* ksum = kv + KV_DOUBLES;
    if (!g.exact)
        for (int k = threadIdx.x; k < cs * ks; k += TWG)
            kt[k] = T(ktab[k]);
    for (int k = threadIdx.x; k < Acc::arrays * plane; k += TWG)
        acc[k] = AS(0);
    const int t = blockIdx.x;
    const int rt = t % g.tiles_r, ct = t / g.tiles_r;
    const int r0 = rt * TILE_R, c0 = ct * g.tc;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int epw = ks <= 64 ? 64 / ks : 1;          // entries per wave pass
    const int slot = ks <= 64 ? lane / ks : 0;
    const int jl = ks <= 64 ? lane - slot * ks : lane;
    const bool lane_on = slot < epw;
    const uint32_t eb = min(offsets[size_t(t) * ostride], cap), ee = min(offsets[size_t(t + 1) * ostride], cap);
    const int batch = g.exact ? min(TEB, KV_DOUBLES / (2 * ks)) : TEB;
    for (uint32_t e0 = eb; e0 < ee; e0 += uint32_t(batch)) {
        const int cnt = int(min(uint32_t(batch), ee - e0));
        __syncthreads();
        if (int(threadIdx.x) < cnt) {
            const Entry E = entries[e0 + threadIdx.x];
            ebuf[threadIdx.x] = E;
            if (g.exact) {
                // exact kernel of this visibility pass (make_kernel_array, conv_func.h:279-303)
                const long long vi = E.idx >> 1;
                double uu = uv[vi] * g.scale, vv = uv[g.n + vi] * g.scale;
                if (g.halfplane && vv < 0.0) {
                    uu = -uu;
                    vv = -vv;
                }
                double fx = uu - int(rint(uu)), fy = vv - int(rint(vv));
                if (E.idx & 1) {
                    fx = -fx;
                    fy = -fy;
                }
                double* ky = &kv[2 * ks * threadIdx.x];
                double* kx = ky + ks;
                double sx = 0.0, sy = 0.0;
                for (int q = 0; q < ks; q++) {
                    ky[q] = dev_kernel(g, double(q - s) - fy);
                    kx[q] = dev_kernel(g, double(q - s) - fx);
                    sy += ky[q];
                    sx += kx[q];
                }
                ksum[threadIdx.x] = sx * sy;
            }
        }
        __syncthreads();
        for (int q0 = wave * epw; q0 < cnt; q0 += (TWG / 64) * epw) {
            const int q = q0 + slot;
            if (!lane_on || q >= cnt)
                continue;
            const Entry& E = ebuf[q];
            const int cpx = E.cp & 0xffff, cpy = E.cp >> 16;
            const T w = T(E.w), vr = T(E.vr), vim = T(E.vi);
            const double* kyd = g.exact ? &kv[2 * ks * q] : nullptr;
            const double inv = g.exact ? ksum[q] : 1.0;
            for (int j = jl; j < ks; j += (ks <= 64 ? ks : 64)) {
                const int c = tap_col(g, E.gx, j);
                const int lc = c - c0;
                if (lc < 0 || lc >= g.tc)
                    continue;
                const double kxd = g.exact ? kyd[ks + j] : 0.0;
                const T kx = g.exact ? T(0) : kt[cpx * ks + j];
                // RG kernel rows at a time: their atomics are issued back to back
                for (int i0 = 0; i0 < ks; i0 += RG) {
                    int aa[RG];
                    T xr[RG], xi[RG], xb[RG];
#pragma unroll
                    for (int u = 0; u < RG; u++) {
                        const int i = i0 + u;
                        aa[u] = -1;
                        xr[u] = xi[u] = xb[u] = T(0);
                        if (i >= ks)
                            continue;
                        const int r = tap_row(g, E.gy, i);
                        const int lr = r - r0;
                        if (r < 0 || lr < 0 || lr >= TILE_R)
                            continue;
                        T K;
                        if (g.exact)
                            K = T((kyd[i] * kxd) / inv);
                        else
                            K = kt[cpy * ks + i] * kx;
                        const T kw = K * w; // gridder.h:854-858
                        aa[u] = lc * S + lr;
                        xr[u] = vr * kw;
                        xi[u] = vim * kw;
                        xb[u] = kw;
                    }
                    Acc::template add_rows<RG, T>(acc, plane, aa, xr, xi, xb, g.beam != 0);
                }
            }
        }
    }
    __syncthreads();
    write_tile<T, Acc>(g, acc, plane, S, r0, c0, vgrid, bgrid);
}

// Oversampled-kernel tiles with ks <= KSM taps per side (the common case): the
// same lane mapping and tap arithmetic as grid_tiles_kernel, specialised so a
// lane preloads its entry's ks kernel-row values and computes row validity
// without branches; taps that miss the tile issue no LDS atomic (the LDS FP64
// atomic rate bounds this kernel).
template <typename T, typename Acc, int KSM>
__device__ __forceinline__ void tiles_os_body(const GridGeom& g, const Entry* entries, const uint32_t* offsets, int ostride,
    uint32_t cap, const double* ktab, cplx<T>* vgrid, cplx<T>* bgrid, int t)
{
#pragma clang fp contract(off)
    using AS = typename Acc::S;
    extern __shared__ double dyn[];
    __shared__ Entry ebuf[TEB];
    const int ks = g.ks, s = g.s;
    const int S = TILE_R + (ks <= 15 ? ks : 1);     // column stride of the LDS tile
    const int plane = g.tc * S;
    AS* acc = reinterpret_cast<AS*>(dyn);
    const int acc_bytes = Acc::arrays * plane * int(sizeof(AS));
    T* kt = reinterpret_cast<T*>(dyn + (acc_bytes + 7) / 8);
    const int cs = (g.os / 2) * 2 + 1;
    for (int k = threadIdx.x; k < cs * ks; k += TWG)
        kt[k] = T(ktab[k]);
    for (int k = threadIdx.x; k < Acc::arrays * plane; k += TWG)
        acc[k] = AS(0);
    const int rt = t % g.tiles_r, ct = t / g.tiles_r;
    const int r0 = rt * TILE_R, c0 = ct * g.tc;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int epw = 64 / ks;                         // entries per wave pass
    const int slot = lane / ks;
    const int jl = lane - slot * ks;
    const bool lane_on = slot < epw;
    const uint32_t eb = min(offsets[size_t(t) * ostride], cap), ee = min(offsets[size_t(t + 1) * ostride], cap);
    for (uint32_t e0 = eb; e0 < ee; e0 += uint32_t(TEB)) {
        const int cnt = int(min(uint32_t(TEB), ee - e0));
        __syncthreads();
        if (int(threadIdx.x) < cnt)
            ebuf[threadIdx.x] = entries[e0 + threadIdx.x];
        __syncthreads();
        for (int q0 = wave * epw; q0 < cnt; q0 += (TWG / 64) * epw) {
            const int q = q0 + slot;
            if (!lane_on || q >= cnt)
                continue;
            const Entry& E = ebuf[q];
            const int lc = tap_col(g, E.gx, jl) - c0;
            if (lc < 0 || lc >= g.tc)
                continue;
            const int cpx = E.cp & 0xffff, cpy = E.cp >> 16;
            const T w = T(E.w), vr = T(E.vr), vim = T(E.vi);
            const T kx = kt[cpx * ks + jl];
            T ky[KSM];
#pragma unroll
            for (int i = 0; i < KSM; i++)
                ky[i] = i < ks ? kt[cpy * ks + i] : T(0);
            const int rb = E.gy - s, cb = lc * S;
#pragma unroll
            for (int i0 = 0; i0 < KSM; i0 += RG) {
                int aa[RG];
                T xr[RG], xi[RG], xb[RG];
#pragma unroll
                for (int u = 0; u < RG; u++) {
                    const int i = i0 + u;
                    // tap_row: rows < 0 skipped, rows >= N wrapped, rows >= rows dropped
                    int r = rb + i;
                    bool ok = i < ks && r >= 0;
                    r = r >= g.N ? r - g.N : r;
                    const int lr = r - r0;
                    ok = ok && r < g.rows && lr >= 0 && lr < TILE_R;
                    const T kw = (ky[i] * kx) * w; // gridder.h:854-858
                    aa[u] = ok ? cb + lr : -1;
                    xr[u] = vr * kw;
                    xi[u] = vim * kw;
                    xb[u] = kw;
                }
                Acc::template add_rows<RG, T>(acc, plane, aa, xr, xi, xb, g.beam != 0);
            }
        }
    }
    __syncthreads();
    write_tile<T, Acc>(g, acc, plane, S, r0, c0, vgrid, bgrid);
}

template <typename T, typename Acc, int KSM>
__global__ __launch_bounds__(TWG) void grid_tiles_os_kernel(GridGeom g, const Entry* entries, const uint32_t* offsets,
    int ostride, uint32_t cap, const double* ktab, cplx<T>* vgrid, cplx<T>* bgrid)
{
    tiles_os_body<T, Acc, KSM>(g, entries, offsets, ostride, cap, ktab, vgrid, bgrid, int(blockIdx.x));
}

// ---------------------------------------------------------------------------
// W-projection (gridder.h:311-367,467-486,767-842; aw_projection.cpp:19-210,366-474)
//
//  1. |w| order: stable LSD radix sort of the 64-bit keys bits(|w|) with the
//     visibility index as payload (the reference's arma::sort_index, gridder.h:324,
//     with ties kept in index order).
//  2. planes: good flags in |w| order, their prefix sum, then the equal-count plane
//     boundaries and plane |w| (mean or median) of average_w_planes
//     (gridder.cpp:68-146) on the device.
//  3. W-kernels per plane: the image-domain kernel aa(x) aa(y) exp(-2 pi i w (n-1)) / n
//     on an (workarea * os)^2 array, a forward c2c FFT, and the (os+1)^2 oversampled
//     (2s+1)^2 kernels cut from it, each normalised by its real sum.
//  4. gridding: the entries of every plane binned by tile together; the tile
//     kernel reads each entry's complex kernel from its plane's cache, so the grid
//     is written once for all planes.
constexpr int RS_KEYS = 1024;          // keys per radix-sort block

__global__ void wsort_keys_kernel(const double* uvw, long long n, unsigned long long* key, uint32_t* idx)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    key[i] = (unsigned long long)__double_as_longlong(fabs(uvw[2 * n + i]));
    idx[i] = uint32_t(i);
}

__global__ __launch_bounds__(256) void rs_hist_kernel(const unsigned long long* key, long long n, int shift, uint32_t* counts,
    int nblk)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * RS_KEYS;
    for (int t = threadIdx.x; t < RS_KEYS; t += 256) {
        const long long i = base + t;
        if (i < n)
            atomicAdd(&h[(key[i] >> shift) & 0xff], 1u);
    }
    __syncthreads();
    counts[size_t(threadIdx.x) * nblk + blockIdx.x] = h[threadIdx.x];
}

// One wave per block of RS_KEYS keys, in order: lanes with equal digits find each
// other by ballots on the 8 digit bits, rank among themselves by lane, and take
// consecutive slots after the digit's running cursor, so the scatter is stable.
__global__ __launch_bounds__(64) void rs_scatter_kernel(const unsigned long long* kin, const uint32_t* iin, long long n, int shift,
    const uint32_t* offs, int nblk, unsigned long long* kout, uint32_t* iout)
{
    __shared__ uint32_t run[256];
    const int lane = threadIdx.x;
    for (int d = lane; d < 256; d += 64)
        run[d] = offs[size_t(d) * nblk + blockIdx.x];
    __syncthreads();
    const long long base = (long long)blockIdx.x * RS_KEYS;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int c = 0; c < RS_KEYS; c += 64) {
        const long long i = base + c + lane;
        const bool valid = i < n;
        const unsigned long long k = valid ? kin[i] : 0ull;
        const uint32_t d = uint32_t(k >> shift) & 0xffu;
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = uint32_t(__popcll(peers & below));
        const uint32_t cnt = uint32_t(__popcll(peers));
        uint32_t slot = 0;
        if (valid)
            slot = run[d];
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0)
            run[d] = slot + cnt;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            kout[slot + rank] = k;
            iout[slot + rank] = iin[i];
        }
    }
}

// The inputs in |w| order, gathered once: uv (u, v, w columns), vis and weights of
// sorted position k from original position perm[k], so the later passes read them
// in sequence instead of gathering each time.
__global__ void wpermute_kernel(const double* uvw, const double* vis, const double* wts, const uint32_t* perm, long long n,
    double* uvw_s, double* vis_s, double* wts_s)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n)
        return;
    const long long i = perm[k];
    const double u = uvw[i], v = uvw[n + i], w = uvw[2 * n + i];
    const double vr = vis[2 * i], vi = vis[2 * i + 1], wt = wts[i];
    uvw_s[k] = u;
    uvw_s[n + k] = v;
    uvw_s[2 * n + k] = w;
    vis_s[2 * k] = vr;
    vis_s[2 * k + 1] = vi;
    wts_s[k] = wt;
}

__global__ void wgood_kernel(GridGeom g, const double* uvw, const double* vis, const double* wts, uint32_t* flag)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= g.n)
        return;
    flag[k] = prep_vis(g, uvw, vis, wts, k).good ? 1u : 0u;
}

// Smallest k in [lo, hi) with G[k] >= target (G non-decreasing), or hi, found by
// one full wave: 64 probes per step narrow the range 64-fold, so a search over 2^21
// positions takes four dependent loads instead of 21.  Every lane returns the result.
__device__ unsigned long long lower_bound_wave(const uint32_t* G, unsigned long long lo, unsigned long long hi,
    unsigned long long target)
{
    const unsigned long long lane = threadIdx.x & 63;
    while (hi - lo > 64) {
        const unsigned long long step = (hi - lo + 63) / 64;
        const unsigned long long q = lo + lane * step;
        const unsigned long long m = __ballot(q >= hi || (unsigned long long)G[q] >= target);
        const int f = m ? __builtin_ctzll(m) : 64;
        if (f == 0)
            return lo;
        // the answer lies in (probe f-1, probe f]
        const unsigned long long nlo = lo + (unsigned long long)(f - 1) * step + 1;
        const unsigned long long nhi = f == 64 ? hi : min(hi, lo + (unsigned long long)f * step);
        lo = nlo;
        hi = nhi;
    }
    const unsigned long long q = lo + lane;
    const unsigned long long m = __ballot(q >= hi || (unsigned long long)G[q] >= target);
    return m ? lo + (unsigned long long)__builtin_ctzll(m) : hi;
}

// average_w_planes (gridder.cpp:68-146) over the |w|-sorted good flags: G is their
// exclusive prefix sum (G[n] = number of good visibilities).  One wave (the planes
// are sequential: each starts where the last stopped).  Writes first[0..P]
// (first[P] = n), first[P+1] (where the last plane's count stopped), the good count
// of each plane, and for the median variant the plane |w|.
__global__ __launch_bounds__(64) void wplane_bounds_kernel(const uint32_t* G, long long n, int P, int median,
    const unsigned long long* skey, uint32_t* first, uint32_t* count, double* w_plane)
{
    const bool lead = threadIdx.x == 0;
    const unsigned long long num_gvis = G[n];
    unsigned long long begin = 0;
    auto wv = [&](unsigned long long k) { return __longlong_as_double((long long)skey[k]); };
    for (int p = 0; p < P; p++) {
        const unsigned long long plane_size = (unsigned long long)ceil(double(num_gvis - begin) / double(P - p));
        unsigned long long k = begin, counter = 0;
        if (plane_size > 0 && begin < (unsigned long long)n) {
            const unsigned long long target = (unsigned long long)G[begin] + plane_size;
            k = lower_bound_wave(G, begin + 1, (unsigned long long)n, target);
            counter = (unsigned long long)G[k] - G[begin];
        }
        if (lead) {
            first[p] = uint32_t(begin);
            count[p] = uint32_t(counter);
        }
        if (median) {
            unsigned long long middle = plane_size / 2;
            if ((begin + plane_size) > num_gvis)
                middle = (num_gvis - begin) / 2;
            double v = 0.0;
            if (counter > 0) {
                // the good element with counter == middle: the first position whose
                // inclusive good count reaches middle + 1
                const unsigned long long mid = lower_bound_wave(G, begin + 1, k, (unsigned long long)G[begin] + middle + 1) - 1;
                v = (plane_size % 2 == 0 && mid > 0) ? (wv(mid - 1) + wv(mid)) / 2.0 : wv(mid);
            }
            if (lead)
                w_plane[p] = v;
        }
        begin = k;
    }
    // the last plane grids every position up to n (gridder.h:532-533); its count
    // stopped at `begin`
    if (lead) {
        first[P] = uint32_t(n);
        first[P + 1] = uint32_t(begin);
    }
}

// Mean |w| of each plane's counted good visibilities: those in [first[p],
// first[p+1]), except for the last plane, whose count stops at first[P+1] (with
// rejected visibilities before it, the reference's plane size, computed from a
// position rather than a good count, stops short of n).  WMS blocks per plane sum
// fixed slices (part[p * WMS + slice]); a second kernel adds the slices in order.
constexpr int WMS = 32;
__global__ __launch_bounds__(256) void wplane_sum_kernel(const unsigned long long* skey, const uint32_t* flag,
    const uint32_t* first, double* part)
{
    __shared__ double red[256];
    const int p = blockIdx.x, sl = blockIdx.y;
    const unsigned long long b = first[p];
    const unsigned long long end = p + 1 == int(gridDim.x) ? first[p + 2] : first[p + 1];
    const unsigned long long len = end > b ? end - b : 0, per = (len + WMS - 1) / WMS;
    const unsigned long long s0 = b + min(len, sl * per), s1 = b + min(len, (sl + 1) * per);
    double sum = 0.0;
    for (unsigned long long k = s0 + threadIdx.x; k < s1; k += 256)
        if (flag[k])
            sum += __longlong_as_double((long long)skey[k]);
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[p * WMS + sl] = red[0];
}

__global__ void wplane_mean_kernel(const double* part, const uint32_t* count, int P, double* w_plane)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P)
        return;
    double sum = 0.0;
    for (int sl = 0; sl < WMS; sl++)
        sum += part[p * WMS + sl];
    w_plane[p] = count[p] > 0 ? sum / double(count[p]) : 0.0;
}

__global__ void wplane_ids_kernel(const uint32_t* first, int P, long long n, uint16_t* plane)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n)
        return;
    int lo = 0, hi = P - 1; // last plane p with first[p] <= k
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if ((long long)first[mid] <= k)
            lo = mid;
        else
            hi = mid - 1;
    }
    plane[k] = uint16_t(lo);
}

// generate_image_domain_convolution_kernel (aw_projection.cpp:50-111): the
// array x array image-domain kernel, zero outside the +-workarea/2 corners.
// blockIdx.y: plane p0 + blockIdx.y, written to array blockIdx.y of comb.
__global__ void wk_fill_kernel(cplx<double>* comb, int arr, int half, const double* aa, const double* w_plane, int p0,
    double scaled_cell)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)arr * arr)
        return;
    const int p = p0 + int(blockIdx.y);
    comb += size_t(blockIdx.y) * arr * arr;
    const int i = int(e % arr), j = int(e / arr);
    // distance of the sample from the origin, folded: i and arr - i map to px = i
    const int pi = (i < half) ? i : ((i > arr - half) ? arr - i : -1);
    const int pj = (j < half) ? j : ((j > arr - half) ? arr - j : -1);
    cplx<double> v{ 0.0, 0.0 };
    if (pi >= 0 && pj >= 0) {
        // combine_2d_kernel_value (aw_projection.cpp:455-474)
        const double a = aa[half + pi] * aa[half + pj];
        if (pi == 0 && pj == 0) {
            v.x = a;
        } else {
            const double dx = double(pi) * scaled_cell, dy = double(pj) * scaled_cell;
            const double r2 = dx * dx + dy * dy;
            if (r2 >= 1.0) {
                v.x = a;
            } else {
                const double nn = sqrt(1.0 - r2);
                const double ph = (-2.0 * M_PI) * w_plane[p] * (nn - 1.0);
                double sn, cs;
                sincos(ph, &sn, &cs);
                v.x = ((cs / nn) * aa[half + pi]) * aa[half + pj];
                v.y = ((sn / nn) * aa[half + pi]) * aa[half + pj];
            }
        }
    }
    comb[e] = v;
}

// Each sub-kernel is stored column-major with WPAD zero columns on both sides
// (ks rows x (ks + 2 WPAD) columns), so that the W tile kernel's column runs
// starting up to WPAD columns outside the kernel read zeros.
constexpr int WPAD = W_KERNEL_PAD_COLS;

// ---------------------------------------------------------------------------
// Hankel W-kernels (hankel_opt; aw_projection.cpp:211-287,380-411, aw_projection.h:127-213)
//
// Per plane: a 1-D profile (the radial screen times the AA kernel, or the
// projection of the 2-D image-domain kernel onto one axis), its transform (the
// DHT, or the forward DFT), the truncated support, a natural cubic or linear
// spline through the transform at radii 0, 1, ..., the spline sampled on the half
// quadrant {(ii, j): ii <= j < ko}, ko = (s + 1) os, at radius sqrt(ii^2 + j^2), and
// the (os+1)^2 cuts of that radially symmetric kernel, each normalised by its real
// sum.  One block per plane for the sequential parts (the spline's tridiagonal
// solve), one thread per output for the transforms.
struct HankelArgs {
    int n;          // transform length: arr / 2 (DHT) or arr (projection slice)
    int half;       // workarea / 2
    int arr;        // workarea * os
    int os, support;
    int proj_slice, cubic;
    double scaled_cell, trunc_frac;
    int kh_stride;  // half-quadrant values per plane (for the maximum support)
};

// combine_2d_kernel_value (aw_projection.cpp:455-474) at pixel offsets (px, py)
__device__ cplx<double> wk_combine(const double* aa, int half, int px, int py, double w, double sc)
{
    const double dx = double(px) * sc, dy = double(py) * sc;
    const double r2 = dx * dx + dy * dy;
    if (r2 >= 1.0)
        return { aa[half + px] * aa[half + py], 0.0 };
    const double nn = sqrt(1.0 - r2);
    const double ph = (-2.0 * M_PI) * w * (nn - 1.0);
    double sn, cs;
    sincos(ph, &sn, &cs);
    return { ((cs / nn) * aa[half + px]) * aa[half + py], ((sn / nn) * aa[half + px]) * aa[half + py] };
}

// Profiles of one plane per block: prof[p * n + i].
__global__ __launch_bounds__(256) void wk_hankel_profile_kernel(HankelArgs h, const double* aa, const double* w_plane,
    cplx<double>* prof)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x;
    const double w = w_plane[p];
    cplx<double>* out = prof + size_t(p) * h.n;
    if (!h.proj_slice) {
        // the radial screen along x times aa(x) (aw_projection.cpp:224-239)
        for (int i = threadIdx.x; i < h.n; i += 256) {
            cplx<double> v{ 0.0, 0.0 };
            if (i < h.half) {
                const double dx = double(i) * h.scaled_cell;
                const double r2 = dx * dx;
                double tr = 1.0, ti = 0.0;
                if (r2 < 1.0) {
                    const double nn = sqrt(1.0 - r2);
                    const double ph = (-2.0 * M_PI) * w * (nn - 1.0);
                    double sn, cs;
                    sincos(ph, &sn, &cs);
                    tr = cs / nn;
                    ti = sn / nn;
                }
                v = { tr * aa[h.half + i], ti * aa[h.half + i] };
            }
            out[i] = v;
        }
        return;
    }
    // generate_projected_image_domain_kernel (aw_projection.cpp:113-171): column sums
    // of the 2-D kernel, each sum in the reference's order
    for (int j = threadIdx.x; j < h.n; j += 256) {
        cplx<double> v{ 0.0, 0.0 };
        if (j == 0) {
            // twice the axis values (1 .. half-1) plus the centre
            double sr = 0.0, si = 0.0;
            for (int i = 1; i < h.half; i++) {
                const cplx<double> c = wk_combine(aa, h.half, i, 0, w, h.scaled_cell);
                sr += c.x;
                si += c.y;
            }
            v = { sr * 2.0 + aa[h.half] * aa[h.half], si * 2.0 };
        } else {
            const int jj = j < h.half ? j : (j > h.arr - h.half ? h.arr - j : -1);
            if (jj > 0) {
                v = wk_combine(aa, h.half, jj, 0, w, h.scaled_cell);
                for (int i = 1; i < h.half; i++) {
                    const cplx<double> c = wk_combine(aa, h.half, i, jj, w, h.scaled_cell);
                    v.x += c.x * 2.0;
                    v.y += c.y * 2.0;
                }
            }
        }
        out[j] = v;
    }
}

// y[p * n + i] = sum_j T(j, i) prof[p * n + j], j in order: T = the DHT (tab[j * n + i])
// or the forward DFT (tab = exp(-2 pi i m / n) interleaved, m = j i mod n).
__global__ __launch_bounds__(256) void wk_hankel_transform_kernel(HankelArgs h, const double* tab,
    const cplx<double>* prof, cplx<double>* y)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x, p = blockIdx.y;
    if (i >= h.n)
        return;
    const cplx<double>* x = prof + size_t(p) * h.n;
    double ar = 0.0, ai = 0.0;
    if (!h.proj_slice) {
        for (int j = 0; j < h.n; j++) {
            const double d = tab[size_t(j) * h.n + i];
            ar += d * x[j].x;
            ai += d * x[j].y;
        }
    } else {
        unsigned long long m = 0;
        for (int j = 0; j < h.n; j++) {
            const double c = tab[2 * m], sn = tab[2 * m + 1];
            ar += x[j].x * c - x[j].y * sn;
            ai += x[j].x * sn + x[j].y * c;
            m += (unsigned long long)i;
            if (m >= (unsigned long long)h.n)
                m -= (unsigned long long)h.n;
        }
    }
    y[size_t(p) * h.n + i] = { ar, ai };
}

// Truncated support (aw_projection.h:132-148) and spline coefficients of one plane:
// lane 0 the real part, lane 1 the imaginary part.  coef[(p * 2 + part) * 5 + {a, b,
// c, d, r}][k]; the natural cubic spline on the knots 0, 1, ... solves its
// tridiagonal system (rows 0 and n-1: 2 b = 0; else (1/3, 4/3, 1/3) against the
// second differences) by forward elimination and back substitution.
__global__ __launch_bounds__(64) void wk_hankel_spline_kernel(HankelArgs h, const cplx<double>* y, double* coef, int* psup)
{
#pragma clang fp contract(off)
    const int p = blockIdx.x, part = threadIdx.x;
    const int n = h.n;
    const cplx<double>* yp = y + size_t(p) * n;
    if (part == 0) {
        int ts = h.support;
        if (h.trunc_frac > 0.0) {
            const double min_value = hypot(yp[0].x, yp[0].y) * h.trunc_frac;
            int idx = h.support * h.os;
            const double scale = 1.0 / double(h.os);
            while (idx > h.os) {
                if (hypot(yp[idx].x, yp[idx].y) > min_value)
                    break;
                idx -= h.os;
            }
            ts = min(int(ceil(double(idx) * scale)), h.support);
        }
        psup[p] = ts;
    }
    if (part > 1)
        return;
    double* a = coef + (size_t(p) * 2 + part) * 5 * n;
    double* b = a + n;
    double* c = b + n;
    double* d = c + n;
    double* r = d + n;
    auto yv = [&](int k) { return part ? yp[k].y : yp[k].x; };
    if (!h.cubic) {
        for (int k = 0; k + 1 < n; k++) {
            c[k] = (yv(k + 1) - yv(k)) / 1.0;
            a[k] = b[k] = 0.0;
        }
        c[n - 1] = c[n - 2];
        a[n - 1] = b[n - 1] = 0.0;
        return;
    }
    const double lo = 1.0 / 3.0 * 1.0, di = 2.0 / 3.0 * 2.0, up = 1.0 / 3.0 * 1.0;
    d[0] = 2.0;
    r[0] = 0.0;
    for (int k = 1; k < n; k++) {
        const bool last = k == n - 1;
        const double lk = last ? 0.0 : lo, dk = last ? 2.0 : di;
        const double rk = last ? 0.0 : (yv(k + 1) - yv(k)) / 1.0 - (yv(k) - yv(k - 1)) / 1.0;
        const double m = lk / d[k - 1];
        d[k] = dk - m * (k == 1 ? 0.0 : up);
        r[k] = rk - m * r[k - 1];
    }
    b[n - 1] = r[n - 1] / d[n - 1];
    for (int k = n - 2; k >= 0; k--)
        b[k] = (r[k] - (k == 0 ? 0.0 : up) * b[k + 1]) / d[k];
    for (int k = 0; k + 1 < n; k++) {
        a[k] = 1.0 / 3.0 * (b[k + 1] - b[k]) / 1.0;
        c[k] = (yv(k + 1) - yv(k)) / 1.0 - 1.0 / 3.0 * (2.0 * b[k] + b[k + 1]) * 1.0;
    }
    a[n - 1] = 0.0;
    c[n - 1] = 3.0 * a[n - 2] * 1.0 * 1.0 + 2.0 * b[n - 2] * 1.0 + c[n - 2];
}

// The spline on the half quadrant of one plane (RadialInterpolate's loop): row ii
// holds j = ii .. ko-1 at kh[p * stride + ii ko - ii (ii - 1) / 2 + j - ii].
__global__ __launch_bounds__(256) void wk_hankel_quad_kernel(HankelArgs h, const cplx<double>* y, const double* coef,
    const int* psup, cplx<double>* kh)
{
#pragma clang fp contract(off)
    const int p = blockIdx.y, n = h.n;
    const int ko = (psup[p] * 2 + 2) * h.os / 2;
    const long long total = (long long)ko * (ko + 1) / 2;
    const cplx<double>* yp = y + size_t(p) * n;
    const double* cr = coef + size_t(p) * 2 * 5 * n;
    const double* ci = cr + 5 * size_t(n);
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < total; q += (long long)gridDim.x * 256) {
        // row ii: the largest with ii ko - ii (ii - 1) / 2 <= q
        int ii = int((double(2 * ko + 1) - sqrt(double(2 * ko + 1) * double(2 * ko + 1) - 8.0 * double(q))) / 2.0);
        ii = max(0, min(ii, ko - 1));
        while (ii > 0 && (long long)ii * ko - (long long)ii * (ii - 1) / 2 > q)
            ii--;
        while (ii + 1 < ko && (long long)(ii + 1) * ko - (long long)(ii + 1) * ii / 2 <= q)
            ii++;
        const int j = int(q - ((long long)ii * ko - (long long)ii * (ii - 1) / 2)) + ii;
        const double nx = ii == 0 ? double(j) : sqrt(double((long long)ii * ii + (long long)j * j));
        int k = int(nx);
        if (k >= n)
            k = n - 1;
        const double hh = nx - double(k);
        double vr, vi;
        if (h.cubic) {
            vr = ((cr[k] * hh + cr[n + k]) * hh + cr[2 * n + k]) * hh + yp[k].x;
            vi = ((ci[k] * hh + ci[n + k]) * hh + ci[2 * n + k]) * hh + yp[k].y;
        } else {
            vr = cr[2 * n + k] * hh + yp[k].x;
            vi = ci[2 * n + k] * hh + yp[k].y;
        }
        kh[size_t(p) * h.kh_stride + q] = { vr, vi };
    }
}

// generate_kernel_cache, Hankel branch (aw_projection.cpp:380-411): cut (x, y) of the
// radially symmetric kernel, element (row tx, column ty) at offsets
// cx = -ts os + os/2 - x + ty os, cy = ... - y + tx os, divided by the cut's real sum;
// stored like wk_extract_kernel's cuts (padded columns, truncated kernels centred).
template <typename KT>
__global__ __launch_bounds__(256) void wk_hankel_extract_kernel(HankelArgs h, const cplx<double>* kh, const int* psup,
    int p0, cplx<KT>* cache)
{
    __shared__ double red[256];
    const int p = p0 + int(blockIdx.y);
    const int cs = (h.os / 2) * 2 + 1;
    const int x = blockIdx.x % cs, yy = blockIdx.x / cs;
    const int ks = 2 * h.support + 1, kw = ks + 2 * WPAD;
    const int ts = psup[p], ko = (ts * 2 + 2) * h.os / 2;
    const int lo = h.support - ts, hi = h.support + ts;
    const long long os = h.os, op = h.os / 2;
    const cplx<double>* khp = kh + size_t(p) * h.kh_stride;
    auto value = [&](int i, int j) -> cplx<double> {
        const long long cx = -(long long)ts * os + op - x + (long long)(j - lo) * os;
        const long long cy = -(long long)ts * os + op - yy + (long long)(i - lo) * os;
        const long long dcx = cx < 0 ? -cx : cx, dcy = cy < 0 ? -cy : cy;
        const long long vi = dcx > dcy ? dcy * ko - ((dcy * (dcy + 1)) >> 1) + dcx : dcx * ko - ((dcx * (dcx + 1)) >> 1) + dcy;
        return khp[vi];
    };
    double sum = 0.0;
    for (int e = threadIdx.x; e < ks * ks; e += 256) {
        const int i = e % ks, j = e / ks;
        if (i >= lo && i <= hi && j >= lo && j <= hi)
            sum += value(i, j).x;
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const double ksum = red[0];
    cplx<KT>* out = cache + (size_t(p) * cs * cs + blockIdx.x) * size_t(ks) * kw;
    for (int e = threadIdx.x; e < ks * kw; e += 256) {
        const int i = e % ks, j = e / ks - WPAD;
        if (j < 0 || j >= ks || i < lo || i > hi || j < lo || j > hi) {
            out[e] = { KT(0), KT(0) };
            continue;
        }
        const cplx<double> v = value(i, j);
        out[e] = { KT(v.x / ksum), KT(v.y / ksum) };
    }
}

// Truncated support of each plane of a batch (aw_projection.cpp:176-179,197-209):
// the largest s <= support whose sample conv(s * os, 0) exceeds trunc_frac times
// |conv(0, 0)|.  A plane truncated to 0 (the reference asserts there, gridder.h:584)
// keeps support 1.
__global__ void wk_trunc_kernel(const cplx<double>* conv, int arr, int os, int support, double trunc_frac, int nb,
    int* plane_support)
{
    const int b = blockIdx.x * 64 + threadIdx.x;
    if (b >= nb)
        return;
    const cplx<double>* c = conv + size_t(b) * arr * arr;
    const double min_value = hypot(c[0].x, c[0].y) * trunc_frac;
    long long i = min((long long)support * os, (long long)(arr >> 1));
    while (i > 0) {
        if (hypot(c[i].x, c[i].y) > min_value)
            break;
        i -= os;
    }
    plane_support[b] = max(1, min(int(i / os), support));
}

// generate_kernel_cache, non-Hankel branch (aw_projection.cpp:415-443): kernel
// (cp_x = x, cp_y = y), element (row i, column j) = conv[cxi(x, j) + cyi(y, i) * arr],
// normalised by the sum of its real parts.  One block per (x, y); blockIdx.y:
// array blockIdx.y of conv, written to plane blockIdx.y of cache.  With a truncated
// plane (plane_support non-NULL) the kernel of support s_p sits centred in the
// (2 support + 1)^2 slot, normalised over its own taps, zero outside: gridding
// those zeros adds nothing, as the reference's smaller loop adds nothing.

template <typename KT>
__global__ __launch_bounds__(256) void wk_extract_kernel(const cplx<double>* conv, int arr, int os, int support,
    const int* plane_support, cplx<KT>* cache)
{
    __shared__ double red[256];
    conv += size_t(blockIdx.y) * arr * arr;
    const int cs = (os / 2) * 2 + 1;
    const int x = blockIdx.x % cs, y = blockIdx.x / cs;
    const int ks = 2 * support + 1;
    const int sp = plane_support ? plane_support[blockIdx.y] : support;
    const int lo = support - sp, hi = support + sp;
    const long long ctr = arr / 2, op = os / 2;
    const long long oc = ctr - (long long)support * os + op;
    double sum = 0.0;
    for (int e = threadIdx.x; e < ks * ks; e += 256) {
        const int i = e % ks, j = e / ks;
        if (i < lo || i > hi || j < lo || j > hi)
            continue;
        const long long cxi = (oc - x + (long long)j * os + ctr) % arr;
        const long long cyi = (oc - y + (long long)i * os + ctr) % arr;
        sum += conv[cxi + cyi * arr].x;
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const double inv = 1.0 / red[0];
    const int kw = ks + 2 * WPAD;
    cplx<KT>* out = cache + (size_t(blockIdx.y) * cs * cs + blockIdx.x) * ks * kw;
    for (int e = threadIdx.x; e < ks * kw; e += 256) {
        const int i = e % ks, j = e / ks - WPAD;
        if (j < 0 || j >= ks || i < lo || i > hi || j < lo || j > hi) {
            out[e] = { KT(0), KT(0) };
            continue;
        }
        const long long cxi = (oc - x + (long long)j * os + ctr) % arr;
        const long long cyi = (oc - y + (long long)i * os + ctr) % arr;
        const cplx<double> v = conv[cxi + cyi * arr];
        out[e] = { KT(v.x * inv), KT(v.y * inv) };
    }
}

// W-projection tile accumulation without atomics.  The tile's 64 rows are the
// lanes and its columns are split among the workgroup's eight waves (CPW columns
// each), so every grid cell belongs to exactly one lane, which keeps the cell's
// sums in registers.  Every wave walks the tile's entries (staged in LDS) and adds
// the taps of its own columns.  Per tap (gridder.h:826-861): kval = K * w with K
// the plane's complex kernel (conjugated for w < 0), grid += vis * kval,
// beam += kval, summed in FP64 on every grid precision.
//
// Phases: the binning blocks hold the visibilities in |w| order, so a run of
// blocks covers about one w-plane.  One launch per run of blocks (a phase) keeps
// the W-kernels that all tiles read during the launch to about one plane's
// cache, which stays in L2, instead of every plane at once.  A phase takes tile
// t's entries of blocks [b0, b1) (the block cursors of the binning give their
// range) and continues the FP64 running sums of the phase before (acc), in the
// same order as one launch over all entries.
struct WPhase {
    const uint32_t* bcur = nullptr; // bcur[b * ntiles + t]: entries of tile t in blocks < b (null: whole tile)
    const uint32_t* btot = nullptr; // entries of tile t
    int b0 = 0, b1 = 0, nblk = 0;
    int first = 1, last = 1;        // first: sums start at zero; last: write the grids
    cplx<double>* acc_v = nullptr;  // FP64 running sums between phases (rows x N, grid layout)
    cplx<double>* acc_b = nullptr;
};
constexpr int WTW = 8;   // waves per W tile workgroup
constexpr double W_PHASE_BYTES = 6.0 * (1 << 20); // W-kernel cache read per phase (measured best on c3)

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uni(double x)
{
    const unsigned long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane(int(b & 0xffffffffu)), hi = __builtin_amdgcn_readfirstlane(int(b >> 32));
    return __longlong_as_double((long long)((unsigned long long)(unsigned)hi << 32 | (unsigned)lo));
}

constexpr uint32_t W_OOB = 0x80000000u;
constexpr int BUF_DWORD3 = 0x00020000; // raw (untyped) buffer access

// One W-kernel value per lane through a range-checked buffer resource over the
// entry's padded sub-kernel (scalar base, scalar column offset): lanes whose row
// is outside [0, ks) carry offset W_OOB and read zero without a memory access.
template <typename KT>
__device__ __forceinline__ cplx<KT> load_kval(const cplx<KT>* sub, int bytes, uint32_t voff, int soff)
{
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)sub, (short)0, bytes, BUF_DWORD3);
    if constexpr (std::is_same<KT, double>::value) {
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        const u4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
        return { __longlong_as_double((long long)((unsigned long long)v.y << 32 | v.x)),
            __longlong_as_double((long long)((unsigned long long)v.w << 32 | v.z)) };
    } else {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        const u2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0);
        return { __uint_as_float(v.x), __uint_as_float(v.y) };
    }
}

__device__ __forceinline__ double flip_if(double x, bool neg)
{
    return neg ? -x : x;
}

template <typename T, typename KT, int CPW, int EB>
__global__ __launch_bounds__(WTW * 64) void grid_wtiles_kernel(GridGeom g, const Entry* __restrict__ entries,
    const uint32_t* __restrict__ offsets, int ostride, uint32_t cap, const cplx<KT>* __restrict__ cache,
    cplx<T>* __restrict__ vgrid, cplx<T>* __restrict__ bgrid, WPhase ph)
{
    const int ks = g.ks, s = g.s, N = g.N;
    const int t = blockIdx.x;
    const int rt = t % g.tiles_r, ct = t / g.tiles_r;
    const int r0 = rt * TILE_R, c0 = ct * g.tc;
    const int wave = uni(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const int r = r0 + lane;
    const int wc = c0 + wave * CPW; // first grid column of this wave (uniform: the loads' buffer resources are scalar)
    uint32_t eb = offsets[size_t(t) * ostride], ee = offsets[size_t(t + 1) * ostride];
    if (ph.bcur) {
        const size_t nt = size_t(g.tiles_r) * g.tiles_c;
        const uint32_t base = eb;
        eb = base + (ph.b0 < ph.nblk ? ph.bcur[size_t(ph.b0) * nt + t] : ph.btot[t]);
        ee = base + (ph.b1 < ph.nblk ? ph.bcur[size_t(ph.b1) * nt + t] : ph.btot[t]);
    }
    eb = uni(int(min(eb, cap)));
    ee = uni(int(min(ee, cap)));
    // FP32 grids take the sums at the last phase; FP64 grids are their own running sums
    constexpr bool narrow = !std::is_same<T, double>::value;
    if (!ph.first && !(narrow && ph.last) && eb >= ee)
        return; // no entry of this tile in the phase: its sums stay as they are
    const bool cell = r < g.rows;
    double ar[CPW], ai[CPW], br[CPW], bi[CPW];
#pragma unroll
    for (int u = 0; u < CPW; u++) {
        ar[u] = ai[u] = br[u] = bi[u] = 0.0;
        const int c = wc + u;
        if (!ph.first && cell && c < N && c < c0 + g.tc) {
            const size_t o = size_t(r) + size_t(c) * g.rows;
            const cplx<double> v = ph.acc_v[o];
            ar[u] = v.x;
            ai[u] = v.y;
            if (g.beam) {
                const cplx<double> b = ph.acc_b[o];
                br[u] = b.x;
                bi[u] = b.y;
            }
        }
    }
    static_assert(CPW - 1 <= WPAD, "column runs must stay within the zero padding");
    const int kw = ks + 2 * WPAD;
    const int sub_bytes = int(size_t(ks) * kw * sizeof(cplx<KT>));
    const int col_bytes = int(ks * sizeof(cplx<KT>));
    // EB entries at a time (scalar loads; every value uniform): all their kernel
    // values are loaded before any is summed
    for (uint32_t e0 = eb; e0 < ee; e0 += EB) {
        cplx<KT> kv[EB][CPW]; // in the cache precision until summed
        double ca[EB], cb[EB], cw[EB];
        bool conj[EB], hit[EB];
        // the batch's entries: their scalar loads go out together (the empty asm
        // takes every field in scalar registers, so one wait covers the batch)
        Entry ent[EB];
#pragma unroll
        for (int b = 0; b < EB; b++)
            ent[b] = entries[min(e0 + b, ee - 1)];
#pragma unroll
        for (int b = 0; b < EB; b++)
            asm volatile("" ::"s"(ent[b].gx), "s"(ent[b].gy), "s"(ent[b].cp), "s"(ent[b].idx), "s"(ent[b].vr),
                "s"(ent[b].vi), "s"(ent[b].w));
#pragma unroll
        for (int b = 0; b < EB; b++) {
            const Entry& en = ent[b];
            conj[b] = en.cp < 0;
            ca[b] = en.vr;
            cb[b] = en.vi;
            cw[b] = en.w;
            // kernel column of the wave's first column, mod N; the wave's columns run
            // on from it and wrap past N
            int j0 = wc - (en.gx - s);
            j0 = j0 < 0 ? j0 + N : (j0 >= N ? j0 - N : j0);
            const bool wrap = j0 > N - CPW;
            // byte offset of each column in the padded sub-kernel: the run
            // j0 - (wrap ? N : 0) + u (columns outside [0, ks) are padding zeros),
            // or, when the run wraps inside the kernel (grids narrower than
            // ks + CPW - 1), column by column with the first padding column for
            // columns outside
            int soff[CPW];
            bool any;
            if (!(wrap && j0 < ks)) {
                const int jv = wrap ? j0 - N : j0;
                any = jv < ks && jv + CPW > 0;
#pragma unroll
                for (int u = 0; u < CPW; u++)
                    soff[u] = (jv + u + WPAD) * col_bytes;
            } else {
                any = true;
#pragma unroll
                for (int u = 0; u < CPW; u++) {
                    const int j = j0 + u >= N ? j0 + u - N : j0 + u;
                    soff[u] = (j < ks ? j + WPAD : 0) * col_bytes;
                }
            }
            hit[b] = e0 + b < ee && any;
            if (!hit[b])
                continue;
            // kernel row of my grid row: direct, or wrapped from >= N (rows < 0 are skipped)
            int i = r - (en.gy - s);
            if (i < 0)
                i += N;
            const uint32_t voff = (i < ks && cell) ? uint32_t(i) * uint32_t(sizeof(cplx<KT>)) : W_OOB;
            const cplx<KT>* K = cache + size_t(uint32_t(en.idx)) * (size_t(ks) * kw);
#pragma unroll
            for (int u = 0; u < CPW; u++)
                kv[b][u] = load_kval<KT>(K, sub_bytes, voff, soff[u]);
        }
#pragma unroll
        for (int b = 0; b < EB; b++) {
            if (!hit[b])
                continue;
            // per tap (gridder.h:826-861): kval = K w, conjugated for w < 0; grid += vis kval,
            // beam += kval, as FMAs with the entry's coefficients (vis w, w) and their signs
            const double sa = flip_if(ca[b], conj[b]), sb = flip_if(cb[b], !conj[b]), sw = flip_if(cw[b], conj[b]);
#pragma unroll
            for (int u = 0; u < CPW; u++) {
                const double kr = double(kv[b][u].x), ki = double(kv[b][u].y);
                ar[u] = fma(ca[b], kr, ar[u]);
                ar[u] = fma(sb, ki, ar[u]);
                ai[u] = fma(cb[b], kr, ai[u]);
                ai[u] = fma(sa, ki, ai[u]);
                br[u] = fma(cw[b], kr, br[u]);
                bi[u] = fma(sw, ki, bi[u]);
            }
        }
    }
    if (cell) {
#pragma unroll
        for (int u = 0; u < CPW; u++) {
            const int c = wc + u;
            if (c < N && c < c0 + g.tc) {
                const size_t o = size_t(r) + size_t(c) * g.rows;
                if (narrow && !ph.last) {
                    ph.acc_v[o] = { ar[u], ai[u] };
                    if (g.beam)
                        ph.acc_b[o] = { br[u], bi[u] };
                } else {
                    vgrid[o] = { T(ar[u]), T(ai[u]) };
                    if (g.beam)
                        bgrid[o] = { T(br[u]), T(bi[u]) };
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused batches: the snapshots of a batch are gridded by one launch per stage.
// Each stage's workgroups are numbered across the snapshots; pre[k] is the first
// workgroup of snapshot k (pre[0] = 0, non-decreasing, snapshots without work
// in a stage repeat the next start) and a workgroup finds its snapshot by a
// binary search over pre (uniform, so scalar loads).
struct BItem {
    GridGeom g;
    const double* uv;
    const double* vis;
    const double* wts;
    const double* ktab;        // oversampled 1-D kernel table
    uint32_t* counts;          // nblk x ntiles block counts, then block cursors
    uint32_t* totals;          // entries per tile
    uint32_t* toffs;           // tile entry offsets (ntiles + 1)
    Entry* entries;
    double* wpart;             // canonical weight / good-count pairs, one per block
    void* vgrid;
    void* bgrid;
    double* total;
    unsigned long long* ngood;
    uint32_t cap;
    int nblk;
};

__device__ __forceinline__ int batch_item(const uint32_t* __restrict__ pre, int K, uint32_t b)
{
    int lo = 0, hi = K - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(BWG) void batch_count_kernel(const BItem* __restrict__ items, const uint32_t* __restrict__ pre,
    int K)
{
    const int k = batch_item(pre, K, blockIdx.x);
    const BItem& it = items[k];
    bin_count_body(it.g, it.uv, it.vis, it.wts, it.counts, it.wpart, int(blockIdx.x - pre[k]));
}

__global__ __launch_bounds__(1024) void batch_colscan_kernel(const BItem* __restrict__ items, const uint32_t* __restrict__ pre,
    int K)
{
    const int k = batch_item(pre, K, blockIdx.x);
    const BItem& it = items[k];
    colscan_body(it.counts, it.nblk, it.g.tiles_r * it.g.tiles_c, it.totals, int(blockIdx.x - pre[k]));
}

__global__ __launch_bounds__(BWG) void batch_scatter_kernel(const BItem* __restrict__ items, const uint32_t* __restrict__ pre,
    int K)
{
    const int k = batch_item(pre, K, blockIdx.x);
    const BItem& it = items[k];
    bin_scatter_body(it.g, it.uv, it.vis, it.wts, it.totals, it.counts, it.entries, it.cap, it.toffs,
        int(blockIdx.x - pre[k]));
}

template <typename T, typename Acc, int KSM>
__global__ __launch_bounds__(TWG) void batch_tiles_kernel(const BItem* __restrict__ items, const uint32_t* __restrict__ pre,
    int K)
{
    const int k = batch_item(pre, K, blockIdx.x);
    const BItem& it = items[k];
    tiles_os_body<T, Acc, KSM>(it.g, it.entries, it.toffs, 1, it.cap, it.ktab, (cplx<T>*)it.vgrid, (cplx<T>*)it.bgrid,
        int(blockIdx.x - pre[k]));
}

// One workgroup per snapshot: the last canonical level of its weight pairs (at
// most CANON_BLOCK blocks), sample_grid_total and the good count.
__global__ __launch_bounds__(CANON_LANES) void batch_total_kernel(const BItem* __restrict__ items)
{
    __shared__ double lds[2 * CANON_LANES];
    const BItem& it = items[blockIdx.x];
    double a, b;
    canon_last_level(it.wpart, unsigned(it.nblk), a, b, lds);
    if (threadIdx.x == 0) {
        it.total[0] = a * 2.0;
        if (it.ngood)
            *it.ngood = (unsigned long long)b;
    }
}

} // namespace

// Grid geometry of one call (validated); the kernel functor in kf.
static GridGeom make_geom(const GridderArgs& a, size_t n, bool beam, KernelFn& kf)
{
    const int N = a.image_size;
    const bool wproj = a.wproj.num_wplanes > 0;
    // W-projection grids with max_wpconv_support (gridder.h:260-268)
    const int support = wproj ? int(a.wproj.max_wpconv_support) : a.support;
    STP_REQUIRE(N >= 4 && (N % 2) == 0, "image_size must be even and >= 4");
    STP_REQUIRE(support >= 1 && support <= 127, "kernel support must be in [1, 127]");
    STP_REQUIRE(2 * support + 1 <= N, "kernel wider than the grid");
    STP_REQUIRE(n < (size_t(1) << 30), "at most 2^30 visibilities per call");
    if (wproj) {
        STP_REQUIRE(!a.kernel_exact, "W-projection requires kernel_exact == false");
        STP_REQUIRE(a.wproj.num_wplanes <= 65535, "at most 65535 w-planes");
        STP_REQUIRE(n < (size_t(1) << 31), "too many visibilities for W-projection");
    }
    int os = a.oversampling;
    if (a.kernel_exact) {
        os = 1;
    } else {
        os = (os >> 1) << 1; // gridder.h:297-303
        if (os == 0)
            os = 1;
    }
    STP_REQUIRE(os <= 1024, "oversampling must be <= 1024");
    GridGeom g;
    g.N = N;
    g.half = N / 2;
    g.rows = a.halfplane ? g.half + 1 : N;
    g.s = support;
    g.ks = 2 * support + 1;
    g.os = os;
    g.halfplane = a.halfplane;
    g.shift = a.shift_uv;
    g.exact = a.kernel_exact;
    g.beam = beam;
    g.bpack = beam && a.beam_packed ? 1 : 0;
    STP_REQUIRE(!g.bpack || (!wproj && a.halfplane), "a packed real beam grid needs the half-plane grid without W-projection");
    g.tiles_r = int(div_up(g.rows, TILE_R));
    g.tc = a.tile_cols > 0 ? a.tile_cols : TILE_C;
    STP_REQUIRE(g.tc == 16 || g.tc == 32 || g.tc == 64, "tile width must be 16, 32 or 64 columns");
    g.tiles_c = int(div_up(N, g.tc));
    g.scale = a.uv_scale;
    g.n = (long long)n;
    kf = kernel_fn(a.kernel);
    g.kfunc = kf.kind;
    g.kparam = kf.param;
    g.kwg = kf.width_g;
    g.kws = kf.width_s;
    return g;
}

template <typename T>
void grid_visibilities(stp_ctx* ctx, const GridderArgs& a, size_t n, const double* d_uv, const double* d_vis,
    const double* d_w, cplx<T>* d_vis_grid, cplx<T>* d_sampling_grid, double* d_total, unsigned long long* d_ngood)
{
    const bool wproj = a.wproj.num_wplanes > 0;
    KernelFn kf;
    GridGeom g = make_geom(a, n, a.generate_beam && d_sampling_grid != nullptr, kf);
    const int N = g.N, os = g.os;
    const int ntiles = g.tiles_r * g.tiles_c;
    // kernel table (oversampled 1-D kernels), computed on the host once per configuration
    const double* d_tab = nullptr;
    const int cs = (os / 2) * 2 + 1;
    if (wproj) {
        d_tab = dev_table(ctx, "k1d:none", std::vector<double>(1, 0.0));
    } else if (!a.kernel_exact) {
        char key[160];
        snprintf(key, sizeof(key), "k1d:%d:%.17g:%.17g:%.17g:%d:%d", kf.kind, kf.param, kf.width_g, kf.width_s, a.support, os);
        d_tab = has_table(ctx, key) ? dev_table(ctx, key, {}) : dev_table(ctx, key, kernel_1d_table(kf, a.support, os));
    } else {
        d_tab = dev_table(ctx, "k1d:none", std::vector<double>(1, 0.0));
    }

    const size_t nblk = std::max<size_t>(1, (n + VPB - 1) / VPB);
    double* wpart = (double*)ws_get(ctx, "grid_wpart", sizeof(double) * (2 * nblk + 2));
    double* wscr = (double*)ws_get(ctx, "grid_wscr", sizeof(double) * canon_scratch_doubles(nblk));
    // worst-case entries: the most tiles one footprint can touch (every centre
    // checked on the host), times two passes
    const size_t cap64 = std::max<size_t>(1, 2 * footprint_tiles_max(g) * n);
    STP_REQUIRE(cap64 < (size_t(1) << 32), "too many gridding entries for one call");
    const uint32_t cap = uint32_t(cap64);
    Entry* entries = (Entry*)ws_get(ctx, "grid_entries", sizeof(Entry) * cap64);

    // W-kernel cache in the grid precision: FP64 kernels on the FP64 path, FP32
    // kernels (half the bytes the W tile kernel reads per tap) on the FP32 paths
    using KT = T;
    cplx<KT>* wcache = nullptr;
    if (wproj) {
        const int P = int(a.wproj.num_wplanes);
        // 1. |w| order
        unsigned long long* k0 = (unsigned long long*)ws_get(ctx, "w_key0", sizeof(unsigned long long) * n);
        unsigned long long* k1 = (unsigned long long*)ws_get(ctx, "w_key1", sizeof(unsigned long long) * n);
        uint32_t* i0 = (uint32_t*)ws_get(ctx, "w_idx0", sizeof(uint32_t) * n);
        uint32_t* i1 = (uint32_t*)ws_get(ctx, "w_idx1", sizeof(uint32_t) * n);
        const size_t nrb = std::max<size_t>(1, div_up(n, RS_KEYS));
        uint32_t* rcounts = (uint32_t*)ws_get(ctx, "w_rcounts", sizeof(uint32_t) * 256 * nrb);
        uint32_t* roffs = (uint32_t*)ws_get(ctx, "w_roffs", sizeof(uint32_t) * (256 * nrb + 1));
        launch(ctx, "w_keys", wsort_keys_kernel, dim3(unsigned(div_up(n, 256))), dim3(256), 0, d_uv, (long long)n, k0, i0);
        for (int shift = 0; shift < 64; shift += 8) {
            launch(ctx, "w_sort_hist", rs_hist_kernel, dim3(unsigned(nrb)), dim3(256), 0, (const unsigned long long*)k0,
                (long long)n, shift, rcounts, int(nrb));
            scan_u32(ctx, rcounts, 256 * nrb, roffs);
            launch(ctx, "w_sort_scatter", rs_scatter_kernel, dim3(unsigned(nrb)), dim3(64), 0, (const unsigned long long*)k0,
                (const uint32_t*)i0, (long long)n, shift, (const uint32_t*)roffs, int(nrb), k1, i1);
            std::swap(k0, k1);
            std::swap(i0, i1);
        }
        {
            double* uvw_s = (double*)ws_get(ctx, "w_uvw_sorted", sizeof(double) * 3 * n);
            double* vis_s = (double*)ws_get(ctx, "w_vis_sorted", sizeof(double) * 2 * n);
            double* wts_s = (double*)ws_get(ctx, "w_wts_sorted", sizeof(double) * n);
            launch(ctx, "w_permute", wpermute_kernel, dim3(unsigned(div_up(n, 256))), dim3(256), 0, d_uv, d_vis, d_w,
                (const uint32_t*)i0, (long long)n, uvw_s, vis_s, wts_s);
            d_uv = uvw_s;
            d_vis = vis_s;
            d_w = wts_s;
            g.perm = nullptr;  // position k of the pass is sorted position k
        }
        // 2. planes
        uint32_t* flag = (uint32_t*)ws_get(ctx, "w_flag", sizeof(uint32_t) * n);
        uint32_t* G = (uint32_t*)ws_get(ctx, "w_gscan", sizeof(uint32_t) * (n + 1));
        uint32_t* first = (uint32_t*)ws_get(ctx, "w_first", sizeof(uint32_t) * (P + 2));
        uint32_t* pcount = (uint32_t*)ws_get(ctx, "w_pcount", sizeof(uint32_t) * P);
        double* w_plane = (double*)ws_get(ctx, "w_plane", sizeof(double) * P);
        uint16_t* plane_of = (uint16_t*)ws_get(ctx, "w_planeof", sizeof(uint16_t) * n);
        launch(ctx, "w_good", wgood_kernel, dim3(unsigned(div_up(n, 256))), dim3(256), 0, g, d_uv, d_vis, d_w, flag);
        scan_u32(ctx, flag, n, G);
        launch(ctx, "w_planes", wplane_bounds_kernel, dim3(1), dim3(64), 0, (const uint32_t*)G, (long long)n, P,
            int(a.wproj.wplanes_median != 0), (const unsigned long long*)k0, first, pcount, w_plane);
        if (!a.wproj.wplanes_median) {
            double* wpart = (double*)ws_get(ctx, "w_plane_part", sizeof(double) * size_t(P) * WMS);
            launch(ctx, "w_plane_sum", wplane_sum_kernel, dim3(unsigned(P), unsigned(WMS)), dim3(256), 0,
                (const unsigned long long*)k0, (const uint32_t*)flag, (const uint32_t*)first, wpart);
            launch(ctx, "w_plane_mean", wplane_mean_kernel, dim3(unsigned(div_up(P, 256))), dim3(256), 0,
                (const double*)wpart, (const uint32_t*)pcount, P, w_plane);
        }
        launch(ctx, "w_plane_ids", wplane_ids_kernel, dim3(unsigned(div_up(n, 256))), dim3(256), 0, (const uint32_t*)first, P,
            (long long)n, plane_of);
        g.plane = plane_of;
        g.wproj = 1;
        // 3. W-kernels: workarea (gridder.h:467-486), image-domain AA kernel, one FFT per plane
        const int us = a.wproj.undersampling_opt > 0 ? (1 << (a.wproj.undersampling_opt - 1)) : 0;
        int workarea = 2;
        if (us > 0)
            while (workarea / us < g.ks)
                workarea *= 2;
        else
            while (workarea < N)
                workarea *= 2;
        const int arr = workarea * os;
        STP_REQUIRE(arr <= 4096, "W-kernel work array (workarea * oversampling) must be <= 4096");
        const double scaling = double(N) / double(workarea);
        char key[192];
        snprintf(key, sizeof(key), "waa:%d:%.17g:%.17g:%.17g:%d:%d", kf.kind, kf.param, kf.width_g, kf.width_s, workarea,
            a.analytic_gcf ? 1 : 0);
        const double* d_aa = has_table(ctx, key) ? dev_table(ctx, key, {})
                                                 : dev_table(ctx, key, img_dom_kernel(kf, size_t(workarea), false, a.analytic_gcf));
        // planes in groups of PB: one fill, one batched 2-D c2c and one extraction per
        // group, the image-domain arrays of a group at most 256 MB each
        const size_t per_plane = size_t(cs) * cs * size_t(g.ks) * (g.ks + 2 * WPAD);
        STP_REQUIRE(size_t(P) * cs * cs < (size_t(1) << 31), "too many W sub-kernels (planes x oversampling^2)");
        wcache = (cplx<KT>*)ws_get(ctx, "w_cache", sizeof(cplx<KT>) * per_plane * P);
        // kernel_trunc_perc > 0: every plane grids with its own truncated support
        const bool trunc = a.wproj.kernel_trunc_perc > 0.0;
        if (a.wproj.hankel_opt) {
            // Hankel W-kernels: profile, DHT or DFT, spline, half quadrant, cuts
            HankelArgs h;
            h.proj_slice = a.wproj.hankel_proj_slice != 0;
            h.cubic = a.wproj.interp_type == 2;
            h.n = h.proj_slice ? arr : arr / 2;
            h.half = workarea / 2;
            h.arr = arr;
            h.os = os;
            h.support = g.s;
            h.scaled_cell = a.cell_rad * scaling;
            h.trunc_frac = trunc ? a.wproj.kernel_trunc_perc / 100.0 : 0.0;
            const long long ko_max = (long long)(g.s + 1) * os;
            h.kh_stride = int(ko_max * (ko_max + 1) / 2);
            STP_REQUIRE(h.n >= 3, "Hankel W-kernels need a transform of at least 3 points");
            STP_REQUIRE(h.n <= 8192, "Hankel W-kernel transform longer than 8192 points (workarea * oversampling)");
            const double* tab;
            char tkey[64];
            if (!h.proj_slice) {
                snprintf(tkey, sizeof(tkey), "dht:%d", h.n);
                tab = has_table(ctx, tkey) ? dev_table(ctx, tkey, {}) : dev_table(ctx, tkey, dht_table(size_t(h.n)));
            } else {
                snprintf(tkey, sizeof(tkey), "dft:%d", h.n);
                tab = has_table(ctx, tkey) ? dev_table(ctx, tkey, {}) : dev_table(ctx, tkey, dft_twiddles(size_t(h.n)));
            }
            cplx<double>* prof = (cplx<double>*)ws_get(ctx, "w_hk_prof", sizeof(cplx<double>) * size_t(h.n) * P);
            cplx<double>* yv = (cplx<double>*)ws_get(ctx, "w_hk_y", sizeof(cplx<double>) * size_t(h.n) * P);
            double* coef = (double*)ws_get(ctx, "w_hk_coef", sizeof(double) * 10 * size_t(h.n) * P);
            cplx<double>* kh = (cplx<double>*)ws_get(ctx, "w_hk_quad", sizeof(cplx<double>) * size_t(h.kh_stride) * P);
            int* psup = (int*)ws_get(ctx, "w_psupport", sizeof(int) * P);
            launch(ctx, "w_hankel_profile", wk_hankel_profile_kernel, dim3(unsigned(P)), dim3(256), 0, h, d_aa,
                (const double*)w_plane, prof);
            launch(ctx, "w_hankel_transform", wk_hankel_transform_kernel, dim3(unsigned(div_up(h.n, 256)), unsigned(P)),
                dim3(256), 0, h, tab, (const cplx<double>*)prof, yv);
            launch(ctx, "w_hankel_spline", wk_hankel_spline_kernel, dim3(unsigned(P)), dim3(64), 0, h, (const cplx<double>*)yv,
                coef, psup);
            launch(ctx, "w_hankel_quad", wk_hankel_quad_kernel,
                dim3(unsigned(std::min<long long>(1024, div_up((long long)h.kh_stride, 256))), unsigned(P)), dim3(256), 0, h,
                (const cplx<double>*)yv, (const double*)coef, (const int*)psup, kh);
            launch(ctx, "w_kernel_extract", wk_hankel_extract_kernel<KT>, dim3(unsigned(cs * cs), unsigned(P)), dim3(256), 0,
                h, (const cplx<double>*)kh, (const int*)psup, 0, wcache);
        } else {
            const size_t arr2 = size_t(arr) * arr;
            const int PB = int(std::max<size_t>(1, std::min<size_t>(size_t(P), (size_t(1) << 28) / (arr2 * sizeof(cplx<double>)))));
            cplx<double>* comb = (cplx<double>*)ws_get(ctx, "w_comb", sizeof(cplx<double>) * arr2 * PB);
            cplx<double>* conv = (cplx<double>*)ws_get(ctx, "w_conv", sizeof(cplx<double>) * arr2 * PB);
            int* psup = trunc ? (int*)ws_get(ctx, "w_psupport", sizeof(int) * P) : nullptr;
            for (int p0 = 0; p0 < P; p0 += PB) {
                const int nb = std::min(PB, P - p0);
                launch(ctx, "w_kernel_fill", wk_fill_kernel, dim3(unsigned(div_up(arr2, 256)), unsigned(nb)), dim3(256), 0, comb,
                    arr, workarea / 2, d_aa, (const double*)w_plane, p0, a.cell_rad * scaling);
                fft_c2c_2d<double>(ctx, comb, conv, arr, true, nb);
                if (trunc)
                    launch(ctx, "w_kernel_trunc", wk_trunc_kernel, dim3(div_up(nb, 64)), dim3(64), 0, (const cplx<double>*)conv, arr,
                        os, g.s, a.wproj.kernel_trunc_perc / 100.0, nb, psup + p0);
                launch(ctx, "w_kernel_extract", wk_extract_kernel<KT>, dim3(unsigned(cs * cs), unsigned(nb)), dim3(256), 0,
                    (const cplx<double>*)conv, arr, os, g.s, (const int*)(trunc ? psup + p0 : nullptr),
                    wcache + size_t(p0) * per_plane);
            }
        }
    }

    const uint32_t* offsets;
    int ostride;
    const uint32_t* bcur = nullptr; // block cursors inside each tile (block binning only)
    const uint32_t* btot = nullptr;
    if (ntiles <= MAX_LDS_TILES) {
        const size_t m = size_t(ntiles) * nblk;
        uint32_t* counts = (uint32_t*)ws_get(ctx, "grid_bcounts", sizeof(uint32_t) * m);
        uint32_t* totals = (uint32_t*)ws_get(ctx, "grid_btotals", sizeof(uint32_t) * ntiles);
        uint32_t* toffs = (uint32_t*)ws_get(ctx, "grid_boffsets", sizeof(uint32_t) * (ntiles + 1));
        const size_t hist_bytes = sizeof(uint32_t) * size_t(ntiles);
        HIP_CHECK(hipFuncSetAttribute((const void*)grid_bin_count_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(hist_bytes)));
        HIP_CHECK(hipFuncSetAttribute((const void*)grid_bin_scatter_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(hist_bytes)));
        launch(ctx, "grid_count", grid_bin_count_kernel, dim3(unsigned(nblk)), dim3(BWG), hist_bytes, g, d_uv, d_vis, d_w, counts,
            wpart);
        launch(ctx, "grid_colscan", bin_colscan_kernel, dim3(div_up(ntiles, 64)), dim3(1024), 0, counts, int(nblk), int(ntiles),
            totals);
        launch(ctx, "grid_scatter", grid_bin_scatter_kernel, dim3(unsigned(nblk)), dim3(BWG), hist_bytes, g, d_uv, d_vis, d_w,
            (const uint32_t*)totals, (const uint32_t*)counts, entries, cap, toffs);
        offsets = toffs;
        ostride = 1;
        bcur = counts;
        btot = totals;
    } else {
        // very large grids: global per-tile counters
        uint32_t* counts = (uint32_t*)ws_get(ctx, "grid_counts", sizeof(uint32_t) * ntiles);
        uint32_t* offs = (uint32_t*)ws_get(ctx, "grid_offsets", sizeof(uint32_t) * (ntiles + 1));
        uint32_t* cursor = (uint32_t*)ws_get(ctx, "grid_cursor", sizeof(uint32_t) * ntiles);
        fill_u32(ctx, counts, 0u, ntiles);
        launch(ctx, "grid_count", grid_count_kernel, dim3(unsigned(nblk)), dim3(GWG), 0, g, d_uv, d_vis, d_w, counts, wpart);
        exclusive_scan_u32(ctx, counts, ntiles, offs, cursor);
        launch(ctx, "grid_scatter", grid_scatter_kernel, dim3(div_up(n, GWG)), dim3(GWG), 0, g, d_uv, d_vis, d_w, cursor,
            entries, cap);
        offsets = offs;
        ostride = 1;
    }
    if (wproj) {
        // eight waves split the tile's columns: 2, 4 or 8 columns per wave; 4 entries'
        // kernel values in flight per lane
        // eight waves split the tile's columns (2, 4 or 8 columns per wave); two
        // entries' kernel values in flight per lane (more registers cost occupancy)
        auto kern = g.tc == 16 ? grid_wtiles_kernel<T, KT, 2, 2>
                               : (g.tc == 32 ? grid_wtiles_kernel<T, KT, 4, 2> : grid_wtiles_kernel<T, KT, 8, 2>);
        // phases whose W-kernels (about 6 MB of cache each: one FP64 plane, two FP32
        // planes) are read by all tiles while mostly held in L2
        const int P = int(a.wproj.num_wplanes);
        const double plane_bytes = double(cs) * cs * g.ks * (g.ks + 2 * WPAD) * sizeof(cplx<KT>);
        int nph = bcur ? int(std::min<double>(double(nblk), std::ceil(double(P) * plane_bytes / W_PHASE_BYTES))) : 1;
        if (a.w_phases > 0)
            nph = bcur ? std::min(a.w_phases, int(nblk)) : 1;
        nph = std::max(1, nph);
        WPhase ph;
        if (nph > 1) {
            ph.bcur = bcur;
            ph.btot = btot;
            ph.nblk = int(nblk);
            if (std::is_same<T, double>::value) {
                ph.acc_v = (cplx<double>*)d_vis_grid;
                ph.acc_b = (cplx<double>*)d_sampling_grid;
            } else {
                const size_t cells = size_t(g.rows) * N;
                ph.acc_v = (cplx<double>*)ws_get(ctx, "w_acc", sizeof(cplx<double>) * cells * (g.beam ? 2 : 1));
                ph.acc_b = g.beam ? ph.acc_v + cells : nullptr;
            }
        }
        for (int f = 0; f < nph; f++) {
            ph.b0 = int(size_t(f) * nblk / nph);
            ph.b1 = int(size_t(f + 1) * nblk / nph);
            ph.first = f == 0;
            ph.last = f == nph - 1;
            launch(ctx, "grid_wtiles", kern, dim3(ntiles), dim3(WTW * 64), 0, g, (const Entry*)entries, offsets, ostride, cap,
                (const cplx<KT>*)wcache, d_vis_grid, d_sampling_grid, ph);
        }
    } else {
        STP_REQUIRE(a.kernel_exact || cs * g.ks <= KV_DOUBLES, "oversampled kernel table too large");
        const bool twofloat = std::is_same<T, float>::value && a.twofloat;
        const int S = TILE_R + (g.ks <= 15 ? g.ks : 1);
        const size_t acc_bytes = twofloat ? 6 * size_t(g.tc) * S * sizeof(float) : 3 * size_t(g.tc) * S * sizeof(double);
        const size_t kv_bytes = a.kernel_exact ? sizeof(double) * (KV_DOUBLES + TEB) : sizeof(double) * size_t(cs) * g.ks;
        if (!a.kernel_exact && g.ks <= 16) {
            // oversampled kernels up to 16 taps: the specialised tile kernel
            const size_t os_acc = (twofloat ? 6 * sizeof(float) : 3 * sizeof(double)) * (size_t(g.tc) * S);
            const size_t os_bytes = ((os_acc + 7) / 8) * 8 + kv_bytes;
            auto kern = g.ks <= 8 ? (twofloat ? grid_tiles_os_kernel<T, AccTwoFloat, 8> : grid_tiles_os_kernel<T, AccF64, 8>)
                                  : (twofloat ? grid_tiles_os_kernel<T, AccTwoFloat, 16> : grid_tiles_os_kernel<T, AccF64, 16>);
            HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(os_bytes)));
            launch(ctx, "grid_tiles", kern, dim3(ntiles), dim3(TWG), os_bytes, g, (const Entry*)entries, offsets, ostride, cap,
                d_tab, d_vis_grid, d_sampling_grid);
        } else {
            const size_t dyn_bytes = ((acc_bytes + 7) / 8) * 8 + kv_bytes;
            auto kern = twofloat ? grid_tiles_kernel<T, AccTwoFloat> : grid_tiles_kernel<T, AccF64>;
            HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(dyn_bytes)));
            launch(ctx, "grid_tiles", kern, dim3(ntiles), dim3(TWG), dyn_bytes, g, (const Entry*)entries, offsets, ostride, cap,
                d_tab, d_uv, d_vis_grid, d_sampling_grid);
        }
    }
    // sample_grid_total = 2 * sum(good weights) (gridder.h:415-423); the good count
    size_t mfin = 0;
    const double* fin = canon_reduce_to_block(ctx, wpart, nblk, wscr, &mfin);
    launch(ctx, "grid_total", grid_total_kernel, dim3(1), dim3(CANON_LANES), 0, fin, unsigned(mfin), d_total, d_ngood,
        offsets + size_t(ntiles) * ostride, a.d_nentries);
}

bool grid_batch_eligible(const GridderArgs& a, size_t n)
{
    if (a.wproj.num_wplanes > 0 || a.kernel_exact || a.support < 1 || 2 * a.support + 1 > 16)
        return false;
    const int N = a.image_size;
    const int tc = a.tile_cols > 0 ? a.tile_cols : TILE_C;
    const int rows = a.halfplane ? N / 2 + 1 : N;
    const long long ntiles = (long long)div_up(rows, TILE_R) * div_up(N, tc);
    const int os = std::max(1, (a.oversampling >> 1) << 1);
    const int cs = (os / 2) * 2 + 1;
    return N >= 8 && ntiles <= MAX_LDS_TILES && std::max<size_t>(1, div_up(n, VPB)) <= size_t(CANON_BLOCK) &&
           cs * (2 * a.support + 1) <= KV_DOUBLES;
}

template <typename T>
BatchPlan grid_batch_bin(stp_ctx* ctx, const std::vector<BatchGridItem>& items, const std::vector<uint8_t>& extra)
{
    const int K = int(items.size());
    STP_REQUIRE(K >= 1 && K <= 65535, "1 to 65535 snapshots per fused batch");
    const bool twofloat = std::is_same<T, float>::value && items[0].a.twofloat;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    // per snapshot: geometry, kernel table, workspace offsets inside one pool
    struct Place {
        size_t counts, totals, toffs, wpart, entries;
    };
    std::vector<BItem> bi(K);
    std::vector<Place> pl(K);
    // tile launches per class: kernel width (up to 8 or 16 taps) x tile width (16, 32,
    // 64 columns), so each launch reserves the LDS its own tiles need
    constexpr int NCLS = BatchPlan::CLASSES;
    std::vector<uint32_t> pre_b(K), pre_c(K), pre_t(size_t(NCLS) * K);
    size_t nb = 0, nc = 0, nt[NCLS] = {}, pool = 0, hist_bytes = 4, tile_bytes[NCLS] = {};
    for (int k = 0; k < K; k++) {
        const BatchGridItem& it = items[k];
        STP_REQUIRE(grid_batch_eligible(it.a, it.n), "snapshot not eligible for a fused batch");
        STP_REQUIRE((std::is_same<T, float>::value && it.a.twofloat) == twofloat, "one tile-sum form per fused batch");
        KernelFn kf;
        BItem& b = bi[k];
        b.g = make_geom(it.a, it.n, it.a.generate_beam && it.bgrid != nullptr, kf);
        const GridGeom& g = b.g;
        const int ntiles = g.tiles_r * g.tiles_c;
        const size_t nblk = std::max<size_t>(1, div_up(it.n, VPB));
        const size_t cap64 = std::max<size_t>(1, 2 * footprint_tiles_max(g) * it.n);
        STP_REQUIRE(cap64 < (size_t(1) << 32), "too many gridding entries for one snapshot");
        char key[160];
        snprintf(key, sizeof(key), "k1d:%d:%.17g:%.17g:%.17g:%d:%d", kf.kind, kf.param, kf.width_g, kf.width_s, it.a.support,
            g.os);
        b.ktab = has_table(ctx, key) ? dev_table(ctx, key, {}) : dev_table(ctx, key, kernel_1d_table(kf, it.a.support, g.os));
        b.uv = it.uv;
        b.vis = it.vis;
        b.wts = it.wts;
        b.vgrid = it.vgrid;
        b.bgrid = g.beam ? it.bgrid : nullptr;
        b.total = it.total;
        b.ngood = it.ngood;
        b.cap = uint32_t(cap64);
        b.nblk = int(nblk);
        Place& p = pl[k];
        p.counts = pool;
        pool = al(pool + sizeof(uint32_t) * nblk * ntiles);
        p.totals = pool;
        pool = al(pool + sizeof(uint32_t) * ntiles);
        p.toffs = pool;
        pool = al(pool + sizeof(uint32_t) * (ntiles + 1));
        p.wpart = pool;
        pool = al(pool + sizeof(double) * (2 * nblk + 2));
        p.entries = pool;
        pool = al(pool + sizeof(Entry) * cap64);
        pre_b[k] = uint32_t(nb);
        nb += nblk;
        pre_c[k] = uint32_t(nc);
        nc += div_up(ntiles, 64);
        const int cls = (g.ks <= 8 ? 0 : 3) + (g.tc == 16 ? 0 : (g.tc == 32 ? 1 : 2));
        for (int c = 0; c < NCLS; c++)
            pre_t[size_t(c) * K + k] = uint32_t(nt[c]);
        nt[cls] += ntiles;
        hist_bytes = std::max(hist_bytes, sizeof(uint32_t) * size_t(ntiles));
        const int S = TILE_R + (g.ks <= 15 ? g.ks : 1);
        const size_t acc = (twofloat ? 6 * sizeof(float) : 3 * sizeof(double)) * (size_t(g.tc) * S);
        const int cs = (g.os / 2) * 2 + 1;
        tile_bytes[cls] = std::max(tile_bytes[cls], ((acc + 7) / 8) * 8 + sizeof(double) * size_t(cs) * g.ks);
    }
    STP_REQUIRE(nb < (size_t(1) << 31) && *std::max_element(nt, nt + NCLS) < (size_t(1) << 31),
        "too many workgroups for one fused batch");
    uint8_t* base = (uint8_t*)ws_get(ctx, "bat_pool", pool);
    for (int k = 0; k < K; k++) {
        bi[k].counts = (uint32_t*)(base + pl[k].counts);
        bi[k].totals = (uint32_t*)(base + pl[k].totals);
        bi[k].toffs = (uint32_t*)(base + pl[k].toffs);
        bi[k].wpart = (double*)(base + pl[k].wpart);
        bi[k].entries = (Entry*)(base + pl[k].entries);
    }
    // tables: items, the stage prefixes, the caller's tables; one upload from the
    // page-locked stage once the previous batch's upload has left it
    const size_t o_items = 0, o_b = al(sizeof(BItem) * K), o_c = o_b + al(4 * K), o_t = o_c + al(4 * K);
    const size_t o_x = o_t + al(4 * size_t(NCLS) * K), tab_bytes = o_x + std::max<size_t>(extra.size(), 1);
    hipEvent_t& up = ctx->batch_tab_events[ctx->ws_slot];
    if (!up)
        HIP_CHECK(hipEventCreateWithFlags(&up, hipEventDisableTiming));
    HIP_CHECK(hipEventSynchronize(up));
    uint8_t* h = (uint8_t*)host_stage(ctx, "bat_tabs#" + std::to_string(ctx->ws_slot), tab_bytes);
    uint8_t* d = (uint8_t*)ws_get(ctx, "bat_tabs", tab_bytes);
    memcpy(h + o_items, bi.data(), sizeof(BItem) * K);
    memcpy(h + o_b, pre_b.data(), 4 * K);
    memcpy(h + o_c, pre_c.data(), 4 * K);
    memcpy(h + o_t, pre_t.data(), 4 * size_t(NCLS) * K);
    if (!extra.empty())
        memcpy(h + o_x, extra.data(), extra.size());
    HIP_CHECK(hipMemcpyAsync(d, h, tab_bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipEventRecord(up, ctx->stream));
    const BItem* d_items = (const BItem*)(d + o_items);
    const uint32_t* d_pre_b = (const uint32_t*)(d + o_b);
    HIP_CHECK(hipFuncSetAttribute((const void*)batch_count_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(hist_bytes)));
    HIP_CHECK(hipFuncSetAttribute((const void*)batch_scatter_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(hist_bytes)));
    launch(ctx, "batch_count", batch_count_kernel, dim3(unsigned(nb)), dim3(BWG), hist_bytes, d_items, d_pre_b, K);
    launch(ctx, "batch_colscan", batch_colscan_kernel, dim3(unsigned(nc)), dim3(1024), 0, d_items, (const uint32_t*)(d + o_c), K);
    launch(ctx, "batch_scatter", batch_scatter_kernel, dim3(unsigned(nb)), dim3(BWG), hist_bytes, d_items, d_pre_b, K);
    BatchPlan plan;
    plan.d_tab = d;
    plan.d_extra = d + o_x;
    plan.items = K;
    plan.twofloat = twofloat;
    plan.pre_t_off = o_t;
    for (int c = 0; c < NCLS; c++) {
        plan.tiles[c] = nt[c];
        plan.tile_lds[c] = tile_bytes[c];
    }
    return plan;
}

template <typename T>
void grid_batch_tiles(stp_ctx* ctx, const BatchPlan& plan)
{
    const BItem* d_items = (const BItem*)plan.d_tab;
    const int K = plan.items;
    for (int c = 0; c < BatchPlan::CLASSES; c++) {
        if (plan.tiles[c] == 0)
            continue;
        auto kern = c < 3 ? (plan.twofloat ? batch_tiles_kernel<T, AccTwoFloat, 8> : batch_tiles_kernel<T, AccF64, 8>)
                           : (plan.twofloat ? batch_tiles_kernel<T, AccTwoFloat, 16> : batch_tiles_kernel<T, AccF64, 16>);
        HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(plan.tile_lds[c])));
        launch(ctx, "batch_tiles", kern, dim3(unsigned(plan.tiles[c])), dim3(TWG), plan.tile_lds[c], d_items,
            (const uint32_t*)((const uint8_t*)plan.d_tab + plan.pre_t_off) + size_t(c) * K, K);
    }
    launch(ctx, "batch_total", batch_total_kernel, dim3(unsigned(K)), dim3(CANON_LANES), 0, d_items);
}

template <typename T>
const void* grid_batch(stp_ctx* ctx, const std::vector<BatchGridItem>& items, const std::vector<uint8_t>& extra)
{
    const BatchPlan plan = grid_batch_bin<T>(ctx, items, extra);
    grid_batch_tiles<T>(ctx, plan);
    return plan.d_extra;
}

template BatchPlan grid_batch_bin<double>(stp_ctx*, const std::vector<BatchGridItem>&, const std::vector<uint8_t>&);
template BatchPlan grid_batch_bin<float>(stp_ctx*, const std::vector<BatchGridItem>&, const std::vector<uint8_t>&);
template void grid_batch_tiles<double>(stp_ctx*, const BatchPlan&);
template void grid_batch_tiles<float>(stp_ctx*, const BatchPlan&);
template const void* grid_batch<double>(stp_ctx*, const std::vector<BatchGridItem>&, const std::vector<uint8_t>&);
template const void* grid_batch<float>(stp_ctx*, const std::vector<BatchGridItem>&, const std::vector<uint8_t>&);

size_t gridder_entry_capacity(int N, bool halfplane, int support, bool exact, int tile_cols, size_t n)
{
    GridGeom g;
    g.N = N;
    g.rows = halfplane ? N / 2 + 1 : N;
    g.s = support;
    g.ks = 2 * support + 1;
    g.exact = exact ? 1 : 0;
    g.tc = tile_cols > 0 ? tile_cols : TILE_C;
    return std::max<size_t>(1, 2 * footprint_tiles_max(g) * n);
}

template void grid_visibilities<double>(stp_ctx*, const GridderArgs&, size_t, const double*, const double*,
    const double*, cplx<double>*, cplx<double>*, double*, unsigned long long*);
template void grid_visibilities<float>(stp_ctx*, const GridderArgs&, size_t, const double*, const double*,
    const double*, cplx<float>*, cplx<float>*, double*, unsigned long long*);

} // namespace stp_hip
