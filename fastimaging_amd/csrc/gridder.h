// Host entry of the tile-binned gridder (gridder.hip).
#pragma once

#include "common.h"

namespace stp_hip {

struct GridderArgs {
    int image_size;      // grid width N (the padded image size for the imager)
    int support;         // conv support s
    int oversampling;    // as passed by the caller (forced even below)
    bool kernel_exact;
    bool shift_uv;
    bool halfplane;
    bool generate_beam;
    double uv_scale;     // applied to u and v (imager: cell_rad * N; unit gridder: 1)
    stp_kernel_spec kernel; // the kernel_creator functor
    bool twofloat = false;  // FP32 grids: two-float LDS sums instead of FP64 LDS sums
    // W-projection (enabled when wproj.num_wplanes > 0): the uv input carries a w column
    stp_wproj_pars wproj{};
    double cell_rad = 0.0;  // cell size in radians (W-kernel image-domain sampling)
    bool analytic_gcf = false; // image-domain AA kernel of the W-kernels from gcf() (conv_func.h:351-412)
    int tile_cols = 0;      // tile width (16, 32 or 64 grid columns); 0 = default
    int w_phases = 0;       // W-projection tile launches (phases); 0 = about 6 MB of W-kernel cache each
    bool beam_packed = false; // sampling grid as packed real values: rows x N/2 complex, element (v, u') =
                              // (G(v, 2u'), G(v, 2u'+1)) (the kernels are real without W-projection)
    unsigned long long* d_nentries = nullptr; // device u64: receives the gridding entries written (may be NULL)
};

// Grids n visibilities into vis_grid (and sampling_grid), each rows x N complex
// column-major with rows = N/2+1 (half-plane) or N.  uv: n x >=2 column-major
// (u column then v column; a w column may follow).  total: device scalar that
// receives sample_grid_total = 2 * sum of good weights.  ngood (device u64) may
// be NULL.
template <typename T>
void grid_visibilities(stp_ctx* ctx, const GridderArgs& a, size_t n, const double* d_uv, const double* d_vis,
    const double* d_w, cplx<T>* d_vis_grid, cplx<T>* d_sampling_grid, double* d_total, unsigned long long* d_ngood);

// One snapshot of a fused batch (grid_batch): its arguments, inputs, grids and
// the device scalars that receive sample_grid_total and the good count.
struct BatchGridItem {
    GridderArgs a;
    size_t n = 0;
    const double* uv = nullptr;
    const double* vis = nullptr;
    const double* wts = nullptr;
    void* vgrid = nullptr;   // cplx<T> grids; bgrid NULL without a beam
    void* bgrid = nullptr;
    double* total = nullptr;
    unsigned long long* ngood = nullptr;
};

// Whether a snapshot can join a fused batch: no W-projection, oversampled
// kernels of at most 16 taps, N >= 8, at most 24576 tiles and 4096 binning blocks
// (16.7 M visibilities).
bool grid_batch_eligible(const GridderArgs& a, size_t n);

// Grids every item of a batch with one launch per stage (binning counts, column
// scans, scatter, the tile kernel per kernel and tile width class, totals): each stage's
// workgroups span all snapshots.  `extra` (caller tables, e.g. the FFT's per-grid
// pointers) is uploaded with the batch's own tables in one copy; returns its
// device address (valid until the next fused batch of the context).
template <typename T>
const void* grid_batch(stp_ctx* ctx, const std::vector<BatchGridItem>& items, const std::vector<uint8_t>& extra);

// The two halves of grid_batch, for callers that put the tile launches on another
// stream: grid_batch_bin uploads the tables and launches the binning on the
// context stream; grid_batch_tiles launches the tile and total kernels on the
// context stream, which must by then be ordered after the binning.
struct BatchPlan {
    static constexpr int CLASSES = 6;  // kernel width (8 or 16 taps) x tile width (16, 32, 64)
    const void* d_tab = nullptr;       // device tables (snapshot records first)
    const void* d_extra = nullptr;     // the caller's tables
    int items = 0;
    bool twofloat = false;
    size_t pre_t_off = 0;              // byte offset of the tile-class prefixes in d_tab
    size_t tiles[CLASSES] = {};
    size_t tile_lds[CLASSES] = {};
};
template <typename T>
BatchPlan grid_batch_bin(stp_ctx* ctx, const std::vector<BatchGridItem>& items, const std::vector<uint8_t>& extra);
template <typename T>
void grid_batch_tiles(stp_ctx* ctx, const BatchPlan& plan);

// W-projection: each W sub-kernel is cached with this many zero columns on both
// sides (ks rows x (ks + 2 W_KERNEL_PAD_COLS) columns).
constexpr int W_KERNEL_PAD_COLS = 7;

// Entries (40 bytes each) the gridder's binning reserves for n visibilities: the
// most tiles one footprint can touch on this grid, times two passes, times n.
size_t gridder_entry_capacity(int N, bool halfplane, int support, bool exact, int tile_cols, size_t n);
constexpr size_t GRID_ENTRY_BYTES = 40;

} // namespace stp_hip
