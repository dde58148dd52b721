// Deterministic reductions shared by the gridder and the source finder.
//
// Canonical summation order (the order the CPU oracle restates): blocks of 4096
// consecutive terms; lane t (0..255) of a block adds terms t, t+256, ...,
// t+15*256 sequentially from 0.0; lane sums combine by the tree
// p[t] = p[t] + p[t+off], off = 128, 64, ..., 1; block results are reduced again
// the same way until one value is left.  Every sum is therefore bit-identical to
// the oracle's, whatever the launch configuration.
#pragma once

#include "common.h"

namespace stp_hip {

constexpr int CANON_LANES = 256;
constexpr int CANON_PER_LANE = 16;
constexpr int CANON_BLOCK = CANON_LANES * CANON_PER_LANE;

// Combines the lane pair (a, b) of all 256 threads of the block in the canonical
// tree; thread 0 returns the block sums.  lds: 512 doubles.
// The levels off = 128 and 64 cross waves (LDS); the levels 32 .. 1 stay inside
// wave 0, where lane t adds lane t + off by a shuffle: the same pairs, so the
// same sums.
__device__ __forceinline__ void canon_tree(double& a, double& b, double* lds)
{
#pragma clang fp contract(off)
    const int t = threadIdx.x;
    lds[t] = a;
    lds[CANON_LANES + t] = b;
    __syncthreads();
    if (t < 128) {
        lds[t] = lds[t] + lds[t + 128];
        lds[CANON_LANES + t] = lds[CANON_LANES + t] + lds[CANON_LANES + t + 128];
    }
    __syncthreads();
    if (t < 64) {
        double x = lds[t] + lds[t + 64];
        double y = lds[CANON_LANES + t] + lds[CANON_LANES + t + 64];
        for (int off = 32; off >= 1; off >>= 1) {
            x = x + __shfl_down(x, off, 64);
            y = y + __shfl_down(y, off, 64);
        }
        if (t == 0) {
            lds[0] = x;
            lds[CANON_LANES] = y;
        }
    }
    __syncthreads();
    a = lds[0];
    b = lds[CANON_LANES];
}

// Last level of the canonical reduction (m <= CANON_BLOCK pairs, one block of
// CANON_LANES threads, the order of the canonical pair kernel); thread 0 gets the sums.
__device__ __forceinline__ void canon_last_level(const double* in, unsigned m, double& a, double& b, double* lds)
{
    // every load is issued before the first sum waits on one: unconditional loads
    // (clamped index; the buffers hold at least one pair)
    double va[CANON_PER_LANE], vb[CANON_PER_LANE];
#pragma unroll
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const unsigned i = min(unsigned(j) * CANON_LANES + threadIdx.x, m > 0 ? m - 1 : 0u);
        va[j] = in[2 * i];
        vb[j] = in[2 * i + 1];
    }
    a = 0.0;
    b = 0.0;
#pragma unroll
    for (int j = 0; j < CANON_PER_LANE; j++) {
        const bool ok = unsigned(j) * CANON_LANES + threadIdx.x < m;
        a = ok ? a + va[j] : a;
        b = ok ? b + vb[j] : b;
    }
    canon_tree(a, b, lds);
}

// Canonical reduction of m pairs (2m doubles at `partials`): runs the levels of
// CANON_BLOCK pairs per block while more than CANON_BLOCK pairs remain and returns
// the remaining pairs (count in *m_out) for a last level done by the caller with one
// block in the same order (canon_last_level).  `scratch` holds
// canon_scratch_doubles(m) doubles.
inline size_t canon_scratch_doubles(size_t m) { return 4 * ((m + CANON_BLOCK - 1) / CANON_BLOCK) + 4; }
const double* canon_reduce_to_block(stp_ctx* ctx, const double* partials, size_t m, double* scratch, size_t* m_out);

// Fill kernels (used instead of hipMemsetAsync so the work stays capturable).
void fill_f64(stp_ctx* ctx, double* p, double v, size_t n);
void fill_u32(stp_ctx* ctx, uint32_t* p, uint32_t v, size_t n);

// Exclusive scan of m uint32 counts into offsets[0..m] and cursor[0..m).
void exclusive_scan_u32(stp_ctx* ctx, const uint32_t* counts, int m, uint32_t* offsets, uint32_t* cursor);

// Device-wide exclusive scan of m uint32 values: out[i] = sum(in[0..i)), out[m] = total.
// Three passes (chunk sums, scan of the sums, chunk rescan) over chunks of 4096.
void scan_u32(stp_ctx* ctx, const uint32_t* in, size_t m, uint32_t* out);
// The same scan over the population counts of m 64-bit masks.
void scan_popc64(stp_ctx* ctx, const unsigned long long* in, size_t m, uint32_t* out);

} // namespace stp_hip
