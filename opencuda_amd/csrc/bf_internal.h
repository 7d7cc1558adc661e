// bf_internal.h -- shared declarations of the HIP implementation (not a
// public header; the C ABI lives in include/opencuda_bf.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "opencuda_bf.h"

namespace bfimpl {

// Status codes (same values as ErrorCode.h).
constexpr int kOk = 0;
constexpr int kInvalidData = -1;
constexpr int kNullPointer = -2;
constexpr int kOutOfMem = -6;
constexpr int kHipError = -7;
constexpr int kUnmatchImg = -8;

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define BF_HIP(call)                                              \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return ::bfimpl::hip_fail(e_, #call); \
    } while (0)

// One tap of the spatial template as the kernels read it:
//   bits  0..7  dy (int8, row offset), bits 8..15 dx (int8, column offset),
//   bits 16..31 row of the fused weight table W[row][d] = g_row * E[d].
inline uint32_t pack_tap(int dy, int dx, int row)
{
    return (uint32_t)(uint8_t)(int8_t)dy | ((uint32_t)(uint8_t)(int8_t)dx << 8) |
           ((uint32_t)row << 16);
}

// Taps are padded to a multiple of this with "null taps" whose weight row
// is all zeros: t + 0*p == t and s + 0 == s bit for bit, so padding never
// changes a result.
constexpr int kTapPad = 4;

// Device-resident filter tables.
struct Lut {
    int id = 0;                  // unique per table set (LDS reuse key)
    int count = 0;               // template points
    std::vector<int32_t> offsets;    // host copy, 2*count
    std::vector<float> weights;      // host copy, count
    std::vector<float> range;        // host copy, 256
    std::vector<int> row_of_tap;     // count
    std::vector<int> rows_upto;      // rows_upto[n] = distinct weights in taps [0,n)
    int nrows_total = 0;
    int device = -1;
    uint32_t *d_taps = nullptr;      // tap variants (see bf_lut_create)
    float *d_gtap = nullptr;         // spatial weight per entry of d_taps (0 for null taps)
    float *d_rowg = nullptr;         // nrows_total
    bool canonical = false;          // offsets = the GAUSS ring order, one weight per q
    float *d_rowq = nullptr;         // canonical: weight of each distinct q, first-appearance order
    float *d_range = nullptr;        // 256
};

// Job as the exact kernel reads it (one per image in a batch).
struct alignas(16) DevJob {
    const uint8_t *in;
    uint8_t *out;
    int64_t in_pitch;
    int64_t out_pitch;
    int32_t in_w, in_h;      // clamp extent (input ROI size)
    int32_t W, H;            // filtered size
    const uint32_t *taps;
    const float *rowg;
    const float *range;
    const float *gtap;       // per-tap spatial weight (separate-table mode)
    int32_t ntaps_padded;    // multiple of kTapPad; tail taps are null
    int32_t nrows;           // rows of W used by the real taps; row nrows is the zero row
    int32_t mode;
    int32_t lut_key;         // identity of the LDS table contents
    int32_t tiles_x;
    int32_t tile_begin;      // first global tile index of this job
    int32_t sep;             // 1: LDS holds E only, g*E formed per tap (large radii)
    int32_t lut_rows;        // 1 KB rows of the LDS table: nrows + 1, or 1 when sep
    int32_t radius;
};

struct TileConfig {
    int tw;        // tile width in pixels
    int tgy;       // 4-row groups along y; tile height = 4*tgy
    int threads;   // tw / cpl * tgy
    int cpl;       // adjacent columns per lane
};

int tile_config_count();
const TileConfig &tile_config(int id);
int halo_class(int radius);   // compile-time halo (>= radius) used for a radius
int default_config(int radius, int mode);

// LDS bytes one workgroup needs for a table of lut_rows 1 KB rows.
size_t lds_bytes(int config, int halo, int lut_rows);
size_t max_lds_bytes();

// Launch the exact kernel over `ntiles` tiles of the jobs.
//   jobs_dev: device job table (njobs entries), or nullptr when njobs == 1
//             (then job0 is passed by value);
//   counter:  device int for the work queue, or nullptr for a static
//             grid-stride schedule.
int launch_exact(int config, int halo, const DevJob &job0, const DevJob *jobs_dev,
                 int njobs, int ntiles, int max_lut_rows, int *counter, hipStream_t s);
// The same launch for jobs that all use the canonical template at radii
// with unrolled kernels (3, 5, 7, 9, 15; tables from ring_job), tile
// configs 0 and 1, halo <= 16.
bool ring_batch_supported(int config, int halo);
int launch_ring_batch(int config, int halo, const DevJob *jobs_dev, int njobs, int ntiles,
                      int max_lut_rows, int *counter, hipStream_t s);
int launch_reset_counter(int *counter, hipStream_t s);

// Build a DevJob for one view pair (validates; fills tiles_x for config;
// chooses the separate-table mode when the fused table does not fit).
int make_job(const bf_image_view *in, const bf_image_view *out, const Lut *lut,
             int radius, int mode, int config, int halo, int tile_begin, DevJob *job,
             int64_t *pixels);
// Fused-table rows needed by (lut, radius, mode), or -1 if the radius
// exceeds the template.
int fused_rows(const Lut *lut, int radius, int mode);

// Unrolled kernels for the canonical ring template: which radii have one
// (and the tile config they use), their fused-table row count, the launch.
bool ring_available(int radius, int mode, int *config);
int ring_rows(int radius, int mode);
int launch_ring(const DevJob &job, int radius, int mode, int ntiles, hipStream_t s);

Lut *lut_from_handle(bf_lut h);
int current_device();
// True when the byte ranges two non-empty views span intersect.
bool views_overlap(const bf_image_view *a, const bf_image_view *b);

// Surrogate helpers used by bf_sur_filter.
int exact_tiles(const DevJob &job, const int32_t *tile_list_dev, int ntiles_dev_max,
                const int32_t *count_dev, int config, int halo, hipStream_t s);

// Unrolled kernels on 64x64 tiles (the surrogate's check and fallback
// tiles) for the canonical template at these radii.  ring64_job switches a
// job made for config 0 to the canonical per-q table when they apply.
bool ring64_available(int radius);
bool ring64_job(const Lut *lut, int radius, int mode, DevJob *job);
// Switch a job (made for `config` with `halo`) to the canonical per-q
// table of the unrolled kernels; false (job unchanged) when they do not apply.
bool ring_job(const Lut *lut, int radius, int mode, int config, int halo, DevJob *job);
int ring64_tiles(const DevJob &job, int radius, int mode, const int32_t *list,
                 const int32_t *count, int ntiles_max, hipStream_t s);

// Exact sample check of 64x64 tiles (one hashed sample per 8x8 cell).
struct CheckArgs {
    DevJob job;                  // exact job (input view, output view, tables)
    const float *rowg;           // weights of the tap rows (run-time tap list)
    int tiles_x, tiles_y;
    int tol;
    int32_t *list;               // fallback tile ids
    int32_t *count;
    int32_t *accum;              // optional running total of fallback tiles
    int32_t *tiles_accum;        // optional running total of tiles
    uint32_t salt;
    int halo;                    // context rows/columns staged around a tile
};
// ring_radius 0: the run-time tap list; 9 or 15: the unrolled canonical
// template (job from ring64_job).
int launch_check(const CheckArgs &a, int ring_radius, int ntiles, hipStream_t s);

}  // namespace bfimpl
