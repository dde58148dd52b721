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
#include <cmath>

#include "gridder.h"
#include "kernel_tables.h"
#include "reduce.h"

namespace stp_hip {
namespace {

constexpr int TILE_R = 64;  // grid rows (v) per tile: 64 contiguous complex per column
constexpr int TILE_C = 32;  // grid columns (u) per tile
constexpr int GWG = 256;
constexpr int MAXT = 8;     // max distinct tiles per dimension of one footprint

struct GridGeom {
    int N, rows, half, s, ks, os;
    int halfplane, shift, exact, beam;
    int tiles_r, tiles_c;
    double scale;
    long long n;
    int kfunc;
    double kparam, kwg, kws;
};

struct Entry {
    int gx, gy;   // kernel centre on the (shifted) grid for this pass
    int cp;       // cp_x | cp_y << 16 (oversampled kernel indices)
    int idx;      // visibility index << 1 | pass
    double vr, vi, w;
};
static_assert(sizeof(Entry) == 40, "entry layout");

struct Prep {
    int good, dup;
    int gx, gy, cpx, cpy;
    double fx, fy;
    double vr, vi, w;
};

__device__ __forceinline__ Prep prep_vis(const GridGeom& g, const double* uv, const double* vis, const double* wts, long long i)
{
    Prep p;
    double u = uv[i] * g.scale;
    double v = uv[g.n + i] * g.scale;
    p.vr = vis[2 * i];
    p.vi = vis[2 * i + 1];
    p.w = wts[i];
    p.dup = 0;
    p.good = (fabs(u) < 1.0e9 && fabs(v) < 1.0e9) ? 1 : 0; // also rejects NaN
    if (!p.good)
        u = v = 0.0;
    if (g.halfplane) {
        if (v < 0.0) {
            p.dup = (v > -(g.s + 1)) ? 1 : 0;
            u = -u;
            v = -v;
            p.vi = -p.vi;
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

__device__ __forceinline__ void add_unique(int* set, int& cnt, int v)
{
    for (int k = 0; k < cnt; k++)
        if (set[k] == v)
            return;
    if (cnt < MAXT)
        set[cnt++] = v;
}

__device__ __forceinline__ void footprint_tiles(const GridGeom& g, int gx, int gy, int* rt, int& nrt, int* ct, int& nct)
{
    nrt = nct = 0;
    for (int i = 0; i < g.ks; i++) {
        const int r = tap_row(g, gy, i);
        if (r >= 0)
            add_unique(rt, nrt, r / TILE_R);
    }
    for (int j = 0; j < g.ks; j++)
        add_unique(ct, nct, tap_col(g, gx, j) / TILE_C);
}

__global__ __launch_bounds__(GWG) void grid_count_kernel(GridGeom g, const double* uv, const double* vis,
    const double* wts, uint32_t* counts, double* wpart, unsigned long long* ngood)
{
#pragma clang fp contract(off)
    __shared__ double lds[2 * CANON_LANES];
    const long long base = (long long)blockIdx.x * CANON_BLOCK;
    double wsum = 0.0, nsum = 0.0;
    unsigned long long ng = 0;
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const long long i = base + (long long)j * CANON_LANES + threadIdx.x;
        if (i >= g.n)
            break;
        const Prep p = prep_vis(g, uv, vis, wts, i);
        if (!p.good)
            continue;
        wsum = wsum + p.w;
        ng++;
        for (int pass = 0; pass <= p.dup; pass++) {
            int gx, gy, cpx, cpy;
            pass_geom(g, p, pass, gx, gy, cpx, cpy);
            int rt[MAXT], ct[MAXT], nrt, nct;
            footprint_tiles(g, gx, gy, rt, nrt, ct, nct);
            for (int a = 0; a < nct; a++)
                for (int b = 0; b < nrt; b++)
                    atomicAdd(&counts[ct[a] * g.tiles_r + rt[b]], 1u);
        }
    }
    canon_tree(wsum, nsum, lds);
    if (threadIdx.x == 0) {
        wpart[2 * blockIdx.x] = wsum;
        wpart[2 * blockIdx.x + 1] = 0.0;
    }
    if (ngood) {
        // wave-level sum of the good count, one atomic per wave
        for (int off = 32; off >= 1; off >>= 1)
            ng += __shfl_down(ng, off, 64);
        if ((threadIdx.x & 63) == 0 && ng)
            atomicAdd(ngood, ng);
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
        int gx, gy, cpx, cpy;
        pass_geom(g, p, pass, gx, gy, cpx, cpy);
        int rt[MAXT], ct[MAXT], nrt, nct;
        footprint_tiles(g, gx, gy, rt, nrt, ct, nct);
        Entry e;
        e.gx = gx;
        e.gy = gy;
        e.cp = cpx | (cpy << 16);
        e.idx = int((i << 1) | pass);
        e.vr = p.vr;
        e.vi = pass ? -p.vi : p.vi;
        e.w = p.w;
        for (int a = 0; a < nct; a++)
            for (int b = 0; b < nrt; b++) {
                const uint32_t slot = atomicAdd(&cursor[ct[a] * g.tiles_r + rt[b]], 1u);
                if (slot < cap)
                    entries[slot] = e;
            }
    }
}

// Kernel functors on the device (exact gridding only; conv_func.cpp:10-202).
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

__global__ void scale_total_kernel(const double* s, double* out)
{
    out[0] = s[0] * 2.0;
}

__global__ void zero_u64_kernel(unsigned long long* p)
{
    p[0] = 0ull;
}

// ---------------------------------------------------------------------------
// Binning by per-block LDS counting sort (no global atomics): block b takes
// visibilities [4096 b, 4096 b + 4096); counts[tile * nblk + b] holds its
// entries per tile, so one exclusive scan of the flat array yields both the
// per-tile entry ranges and each block's write cursor inside every tile.
constexpr int BWG = 1024;
constexpr int VPB = CANON_BLOCK; // visibilities per binning block (also the canonical sum block)
constexpr int MAX_LDS_TILES = 24576;

__global__ __launch_bounds__(BWG) void grid_bin_count_kernel(GridGeom g, const double* uv, const double* vis, const double* wts,
    uint32_t* counts, int nblk, double* wpart, unsigned long long* ngood)
{
#pragma clang fp contract(off)
    extern __shared__ uint32_t hist[];
    __shared__ double wl[VPB];
    const int ntiles = g.tiles_r * g.tiles_c;
    for (int t = threadIdx.x; t < ntiles; t += BWG)
        hist[t] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * VPB;
    unsigned long long ng = 0;
    for (int j = 0; j < VPB / BWG; j++) {
        const int local = threadIdx.x + j * BWG;
        const long long i = base + local;
        double wgood = 0.0;
        if (i < g.n) {
            const Prep p = prep_vis(g, uv, vis, wts, i);
            if (p.good) {
                wgood = p.w;
                ng++;
                for (int pass = 0; pass <= p.dup; pass++) {
                    int gx, gy, cpx, cpy;
                    pass_geom(g, p, pass, gx, gy, cpx, cpy);
                    int rt[MAXT], ct[MAXT], nrt, nct;
                    footprint_tiles(g, gx, gy, rt, nrt, ct, nct);
                    for (int a = 0; a < nct; a++)
                        for (int b = 0; b < nrt; b++)
                            atomicAdd(&hist[ct[a] * g.tiles_r + rt[b]], 1u);
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
    __syncthreads();
    if (threadIdx.x == 0) {
        wpart[2 * blockIdx.x] = wl[0];
        wpart[2 * blockIdx.x + 1] = 0.0;
    }
    for (int t = threadIdx.x; t < ntiles; t += BWG)
        counts[size_t(t) * nblk + blockIdx.x] = hist[t];
    if (ngood) {
        for (int off = 32; off >= 1; off >>= 1)
            ng += __shfl_down(ng, off, 64);
        if ((threadIdx.x & 63) == 0 && ng)
            atomicAdd(ngood, ng);
    }
}

__global__ __launch_bounds__(BWG) void grid_bin_scatter_kernel(GridGeom g, const double* uv, const double* vis, const double* wts,
    const uint32_t* flat_off, int nblk, Entry* entries, uint32_t cap)
{
    extern __shared__ uint32_t cur[];
    const int ntiles = g.tiles_r * g.tiles_c;
    for (int t = threadIdx.x; t < ntiles; t += BWG)
        cur[t] = flat_off[size_t(t) * nblk + blockIdx.x];
    __syncthreads();
    const long long base = (long long)blockIdx.x * VPB;
    for (int j = 0; j < VPB / BWG; j++) {
        const long long i = base + threadIdx.x + j * BWG;
        if (i >= g.n)
            break;
        const Prep p = prep_vis(g, uv, vis, wts, i);
        if (!p.good)
            continue;
        for (int pass = 0; pass <= p.dup; pass++) {
            int gx, gy, cpx, cpy;
            pass_geom(g, p, pass, gx, gy, cpx, cpy);
            int rt[MAXT], ct[MAXT], nrt, nct;
            footprint_tiles(g, gx, gy, rt, nrt, ct, nct);
            Entry e;
            e.gx = gx;
            e.gy = gy;
            e.cp = cpx | (cpy << 16);
            e.idx = int((i << 1) | pass);
            e.vr = p.vr;
            e.vi = pass ? -p.vi : p.vi;
            e.w = p.w;
            for (int a = 0; a < nct; a++)
                for (int b = 0; b < nrt; b++) {
                    const uint32_t slot = atomicAdd(&cur[ct[a] * g.tiles_r + rt[b]], 1u);
                    if (slot < cap)
                        entries[slot] = e;
                }
        }
    }
}

// Accumulates every entry binned to this tile.  One wave per entry, lane = kernel tap
// (row fastest); LDS column stride TILE_R + ks keeps the taps of one entry on distinct
// banks; entries are staged through LDS 256 at a time.
constexpr int EBUF = GWG;

template <typename T>
__global__ __launch_bounds__(GWG) void grid_tiles_kernel(GridGeom g, const Entry* entries, const uint32_t* offsets, int ostride,
    uint32_t cap, const double* ktab, int ktab_lds, const double* uv, cplx<T>* vgrid, cplx<T>* bgrid)
{
    extern __shared__ double sm[];
    __shared__ Entry ebuf[EBUF];
    const int S = TILE_R + (g.ks <= 15 ? g.ks : 1);
    double* acc_r = sm;
    double* acc_i = sm + TILE_C * S;
    double* acc_b = sm + 2 * TILE_C * S;
    double* kt = sm + 3 * TILE_C * S;
    const int cs = (g.os / 2) * 2 + 1;
    const double* K1 = ktab;
    if (ktab_lds) {
        for (int k = threadIdx.x; k < cs * g.ks; k += GWG)
            kt[k] = ktab[k];
        K1 = kt;
    }
    const int t = blockIdx.x;
    const int rt = t % g.tiles_r, ct = t / g.tiles_r;
    const int r0 = rt * TILE_R, c0 = ct * TILE_C;
    for (int k = threadIdx.x; k < TILE_C * S; k += GWG) {
        acc_r[k] = 0.0;
        acc_i[k] = 0.0;
        acc_b[k] = 0.0;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t eb = min(offsets[size_t(t) * ostride], cap), ee = min(offsets[size_t(t + 1) * ostride], cap);
    const int ntaps = g.ks * g.ks;
    for (uint32_t e0 = eb; e0 < ee; e0 += EBUF) {
        const int cnt = int(min(uint32_t(EBUF), ee - e0));
        __syncthreads();
        if (int(threadIdx.x) < cnt)
            ebuf[threadIdx.x] = entries[e0 + threadIdx.x];
        __syncthreads();
        for (int q = wave; q < cnt; q += GWG / 64) {
            const Entry E = ebuf[q];
            const int cpx = E.cp & 0xffff, cpy = E.cp >> 16;
            double fx = 0.0, fy = 0.0, ksum = 1.0;
            if (g.exact) {
                // exact kernel for this visibility (make_kernel_array, conv_func.h:279-303)
                const long long vi = E.idx >> 1;
                const double u = uv[vi] * g.scale, v = uv[g.n + vi] * g.scale;
                double uu = u, vv = v;
                if (g.halfplane && v < 0.0) {
                    uu = -u;
                    vv = -v;
                }
                fx = uu - int(rint(uu));
                fy = vv - int(rint(vv));
                if (E.idx & 1) {
                    fx = -fx;
                    fy = -fy;
                }
                double sx = 0.0, sy = 0.0;
                for (int qq = 0; qq < g.ks; qq++) {
                    sx += dev_kernel(g, double(qq - g.s) - fx);
                    sy += dev_kernel(g, double(qq - g.s) - fy);
                }
                ksum = sx * sy;
            }
            for (int tap = lane; tap < ntaps; tap += 64) {
                const int i = tap % g.ks, j = tap / g.ks;
                const int r = tap_row(g, E.gy, i);
                if (r < 0)
                    continue;
                const int c = tap_col(g, E.gx, j);
                const int lr = r - r0, lc = c - c0;
                if (lr < 0 || lr >= TILE_R || lc < 0 || lc >= TILE_C)
                    continue;
                double K;
                if (g.exact)
                    K = (dev_kernel(g, double(i - g.s) - fy) * dev_kernel(g, double(j - g.s) - fx)) / ksum;
                else
                    K = K1[cpy * g.ks + i] * K1[cpx * g.ks + j];
                const double kw = K * E.w;
                const int a = lc * S + lr;
                unsafeAtomicAdd(&acc_r[a], E.vr * kw);
                unsafeAtomicAdd(&acc_i[a], E.vi * kw);
                if (g.beam)
                    unsafeAtomicAdd(&acc_b[a], kw);
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TILE_C * TILE_R; k += GWG) {
        const int lr = k % TILE_R, lc = k / TILE_R;
        const int r = r0 + lr, c = c0 + lc;
        if (r < g.rows && c < g.N) {
            const size_t o = size_t(r) + size_t(c) * g.rows;
            const int a = lc * S + lr;
            vgrid[o] = { T(acc_r[a]), T(acc_i[a]) };
            if (g.beam)
                bgrid[o] = { T(acc_b[a]), T(0) };
        }
    }
}

} // namespace

template <typename T>
void grid_visibilities(stp_ctx* ctx, const GridderArgs& a, size_t n, const double* d_uv, const double* d_vis,
    const double* d_w, cplx<T>* d_vis_grid, cplx<T>* d_sampling_grid, double* d_total, unsigned long long* d_ngood)
{
    const int N = a.image_size;
    STP_REQUIRE(N >= 4 && (N % 2) == 0, "image_size must be even and >= 4");
    STP_REQUIRE(a.support >= 1 && a.support <= 64, "kernel support must be in [1, 64]");
    STP_REQUIRE(2 * a.support + 1 <= N, "kernel wider than the grid");
    STP_REQUIRE(n < (size_t(1) << 30), "at most 2^30 visibilities per call");
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
    g.s = a.support;
    g.ks = 2 * a.support + 1;
    g.os = os;
    g.halfplane = a.halfplane;
    g.shift = a.shift_uv;
    g.exact = a.kernel_exact;
    g.beam = a.generate_beam && d_sampling_grid != nullptr;
    g.tiles_r = int(div_up(g.rows, TILE_R));
    g.tiles_c = int(div_up(N, TILE_C));
    g.scale = a.uv_scale;
    g.n = (long long)n;
    const KernelFn kf = kernel_fn(a.kernel);
    g.kfunc = kf.kind;
    g.kparam = kf.param;
    g.kwg = kf.width_g;
    g.kws = kf.width_s;
    if (a.kernel_exact)
        STP_REQUIRE(kf.kind != STP_KERNEL_PSWF, "exact gridding with PSWF is not implemented on the GPU");

    const int ntiles = g.tiles_r * g.tiles_c;
    // kernel table (oversampled 1-D kernels), computed on the host once per configuration
    const double* d_tab = nullptr;
    const int cs = (os / 2) * 2 + 1;
    if (!a.kernel_exact) {
        char key[160];
        snprintf(key, sizeof(key), "k1d:%d:%.17g:%.17g:%.17g:%d:%d", kf.kind, kf.param, kf.width_g, kf.width_s, a.support, os);
        d_tab = has_table(ctx, key) ? dev_table(ctx, key, {}) : dev_table(ctx, key, kernel_1d_table(kf, a.support, os));
    } else {
        d_tab = dev_table(ctx, "k1d:none", std::vector<double>(1, 0.0));
    }

    const size_t nblk = std::max<size_t>(1, (n + VPB - 1) / VPB);
    double* wpart = (double*)ws_get(ctx, "grid_wpart", sizeof(double) * (2 * nblk + 2));
    double* wscr = (double*)ws_get(ctx, "grid_wscr", sizeof(double) * canon_scratch_doubles(nblk));
    double* wres = (double*)ws_get(ctx, "grid_wres", sizeof(double) * 2);
    // worst-case entries: every footprint spans at most (ceil(ks/T)+2) tiles per axis, two passes
    const size_t per_pass = size_t(std::min(MAXT, int(div_up(g.ks, TILE_R)) + 2)) * size_t(std::min(MAXT, int(div_up(g.ks, TILE_C)) + 2));
    const size_t cap64 = std::max<size_t>(1, 2 * per_pass * n);
    STP_REQUIRE(cap64 < (size_t(1) << 32), "too many gridding entries for one call");
    const uint32_t cap = uint32_t(cap64);
    Entry* entries = (Entry*)ws_get(ctx, "grid_entries", sizeof(Entry) * cap64);
    if (d_ngood)
        launch(ctx, "fill", zero_u64_kernel, dim3(1), dim3(1), 0, d_ngood);

    const uint32_t* offsets;
    int ostride;
    if (ntiles <= MAX_LDS_TILES) {
        const size_t m = size_t(ntiles) * nblk;
        uint32_t* counts = (uint32_t*)ws_get(ctx, "grid_bcounts", sizeof(uint32_t) * m);
        uint32_t* flat = (uint32_t*)ws_get(ctx, "grid_boffsets", sizeof(uint32_t) * (m + 1));
        const size_t hist_bytes = sizeof(uint32_t) * size_t(ntiles);
        HIP_CHECK(hipFuncSetAttribute((const void*)grid_bin_count_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(hist_bytes)));
        HIP_CHECK(hipFuncSetAttribute((const void*)grid_bin_scatter_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(hist_bytes)));
        launch(ctx, "grid_count", grid_bin_count_kernel, dim3(unsigned(nblk)), dim3(BWG), hist_bytes, g, d_uv, d_vis, d_w, counts,
            int(nblk), wpart, d_ngood);
        scan_u32(ctx, counts, m, flat);
        launch(ctx, "grid_scatter", grid_bin_scatter_kernel, dim3(unsigned(nblk)), dim3(BWG), hist_bytes, g, d_uv, d_vis, d_w,
            (const uint32_t*)flat, int(nblk), entries, cap);
        offsets = flat;
        ostride = int(nblk);
    } else {
        // very large grids: global per-tile counters
        uint32_t* counts = (uint32_t*)ws_get(ctx, "grid_counts", sizeof(uint32_t) * ntiles);
        uint32_t* offs = (uint32_t*)ws_get(ctx, "grid_offsets", sizeof(uint32_t) * (ntiles + 1));
        uint32_t* cursor = (uint32_t*)ws_get(ctx, "grid_cursor", sizeof(uint32_t) * ntiles);
        fill_u32(ctx, counts, 0u, ntiles);
        launch(ctx, "grid_count", grid_count_kernel, dim3(unsigned(nblk)), dim3(GWG), 0, g, d_uv, d_vis, d_w, counts, wpart,
            d_ngood);
        exclusive_scan_u32(ctx, counts, ntiles, offs, cursor);
        launch(ctx, "grid_scatter", grid_scatter_kernel, dim3(div_up(n, GWG)), dim3(GWG), 0, g, d_uv, d_vis, d_w, cursor,
            entries, cap);
        offsets = offs;
        ostride = 1;
    }
    const int S = TILE_R + (g.ks <= 15 ? g.ks : 1);
    const int ktab_lds = (!a.kernel_exact && cs * g.ks <= 1024) ? 1 : 0;
    const size_t tile_bytes = sizeof(double) * (3 * size_t(TILE_C) * S + (ktab_lds ? size_t(cs) * g.ks : 0));
    HIP_CHECK(hipFuncSetAttribute((const void*)grid_tiles_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, int(tile_bytes)));
    launch(ctx, "grid_tiles", grid_tiles_kernel<T>, dim3(ntiles), dim3(GWG), tile_bytes, g, (const Entry*)entries, offsets,
        ostride, cap, d_tab, ktab_lds, d_uv, d_vis_grid, d_sampling_grid);
    // sample_grid_total = 2 * sum(good weights) (gridder.h:415-423)
    canon_reduce_pairs(ctx, wpart, nblk, wscr, wres);
    launch(ctx, "grid_total", scale_total_kernel, dim3(1), dim3(1), 0, (const double*)wres, d_total);
}

template void grid_visibilities<double>(stp_ctx*, const GridderArgs&, size_t, const double*, const double*,
    const double*, cplx<double>*, cplx<double>*, double*, unsigned long long*);

} // namespace stp_hip
