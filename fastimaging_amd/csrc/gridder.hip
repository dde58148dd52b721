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
#include <type_traits>

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
    const int plane = TILE_C * S;
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
    const int r0 = rt * TILE_R, c0 = ct * TILE_C;
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
                if (lc < 0 || lc >= TILE_C)
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
    for (int k = threadIdx.x; k < TILE_C * TILE_R; k += TWG) {
        const int lr = k % TILE_R, lc = k / TILE_R;
        const int r = r0 + lr, c = c0 + lc;
        if (r < g.rows && c < g.N) {
            const size_t o = size_t(r) + size_t(c) * g.rows;
            const int a = lc * S + lr;
            vgrid[o] = { T(Acc::value(acc, plane, 0, a)), T(Acc::value(acc, plane, 1, a)) };
            if (g.beam)
                bgrid[o] = { T(Acc::value(acc, plane, 2, a)), T(0) };
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
    STP_REQUIRE(a.kernel_exact || cs * g.ks <= KV_DOUBLES, "oversampled kernel table too large");
    {
        const bool twofloat = std::is_same<T, float>::value && a.twofloat;
        const int S = TILE_R + (g.ks <= 15 ? g.ks : 1);
        const size_t acc_bytes = twofloat ? 6 * size_t(TILE_C) * S * sizeof(float) : 3 * size_t(TILE_C) * S * sizeof(double);
        const size_t kv_bytes = a.kernel_exact ? sizeof(double) * (KV_DOUBLES + TEB) : sizeof(double) * size_t(cs) * g.ks;
        const size_t dyn_bytes = ((acc_bytes + 7) / 8) * 8 + kv_bytes;
        auto kern = twofloat ? grid_tiles_kernel<T, AccTwoFloat> : grid_tiles_kernel<T, AccF64>;
        HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(dyn_bytes)));
        launch(ctx, "grid_tiles", kern, dim3(ntiles), dim3(TWG), dyn_bytes, g, (const Entry*)entries, offsets, ostride, cap,
            d_tab, d_uv, d_vis_grid, d_sampling_grid);
    }
    // sample_grid_total = 2 * sum(good weights) (gridder.h:415-423)
    canon_reduce_pairs(ctx, wpart, nblk, wscr, wres);
    launch(ctx, "grid_total", scale_total_kernel, dim3(1), dim3(1), 0, (const double*)wres, d_total);
}

template void grid_visibilities<double>(stp_ctx*, const GridderArgs&, size_t, const double*, const double*,
    const double*, cplx<double>*, cplx<double>*, double*, unsigned long long*);
template void grid_visibilities<float>(stp_ctx*, const GridderArgs&, size_t, const double*, const double*,
    const double*, cplx<float>*, cplx<float>*, double*, unsigned long long*);

} // namespace stp_hip
