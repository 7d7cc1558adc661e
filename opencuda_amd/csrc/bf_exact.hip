// bf_exact.hip -- exact FP32 bilateral filter for gfx950 (MI355X).
//
// Computes, per output pixel, exactly what okano_3_0/BilateralFilter.cu
// :71-199 computes (or, in STD mode, the own-window variant), with the
// same rounding sequence: f = g*E, t = t + f*p, s = s + f in tap order, no
// fused multiply-add (built with -ffp-contract=off), IEEE division, u8 by
// truncation.  The host transcription in oracle/ is bit-identical.
//
// Mapping to CDNA4:
//  * Work unit = a tile of tw x 4*tgy output pixels; a workgroup has
//    tw*tgy lanes and each lane owns one column of one 4-row group (the
//    reference's thread shape, BilateralFilter.cu:84-85).  In REF mode one
//    neighbour fetch feeds all four rows (the reference reads the window of
//    the group's first row for every row).
//  * The input tile plus a halo of HALO pixels is staged in LDS once per
//    tile, clamped to the input ROI like the reference's clamp-mode texture
//    (:18, :61-64).  Interior tiles use aligned 128-bit global loads.  Each
//    pixel is stored as a u32 with its byte replicated four times, so that
//    v_sad_u8(p, c, base) = 4*|p - c| + base is directly the LDS byte
//    address of the range weight -- one VALU op per pixel-tap for the
//    whole gather address.
//  * The weights live in LDS as a fused table W[row][d] = g_row * E[d]
//    (one row per distinct spatial weight; 121 rows at r = 15).  The
//    product is formed with one IEEE multiply, exactly as the reference
//    forms factor = g[col] * E[d] (:125), so a table entry holds the same
//    bits as the per-tap product.  One extra all-zero row backs the null
//    taps that pad the tap list to a multiple of 4.
//  * Tap descriptors (dy, dx, row) are wave-uniform: they are read with
//    scalar loads and decoded on the SALU; per pixel-tap the VALU does
//    sad, mul, add, add (REF: plus 1/4 of an address add and a u8->f32
//    convert per fetch).
//  * Workgroups are persistent: each builds the weight table once and then
//    walks tiles (static grid-stride, or an atomic queue for batches),
//    rebuilding the table only when the job's tables change.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "bf_internal.h"
#include "bf_tile.h"

namespace bfimpl {

namespace {

// A lane owns CPL adjacent columns (CPL * lx ..) of one 4-row group; its
// columns' fetches of a tap are adjacent words of the staged tile (one
// ds_read2_b32 for CPL = 2).  Every pixel runs its own taps in tap order,
// so CPL changes no result.
template <int TW, int TGY, int HALO, int CPL, bool SEP>
__device__ __forceinline__ void tile_ref(const DevJob &job, int X0, int Y0, int lx, int gy)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr uint32_t LUT_BASE = kTileOff + G::TILE_BYTES;
    constexpr int ROWB = G::TCOLS * 4;
    const uint32_t lane = kTileOff + (uint32_t)(((4 * gy + HALO) * G::TCOLS + CPL * lx + HALO) * 4);
    uint32_t cen[4][CPL];
    float t[4][CPL], sum[4][CPL];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int j = 0; j < CPL; j++) {
            cen[k][j] = lds_u32(lane + k * ROWB + 4 * j);
            t[k][j] = 0.0f;
            sum[k][j] = 0.0f;
        }
    const BF_CONST uint32_t *taps = as_const(job.taps);
    const BF_CONST float *gtap = as_const(job.gtap);
    const int nt = job.ntaps_padded;
    for (int c = 0; c < nt; c += kTapPad) {
#pragma unroll
        for (int u = 0; u < kTapPad; u++) {
            const uint32_t info = taps[c + u];
            const int dy = (int)(int8_t)(info & 0xFFu);
            const int dx = (int)(int8_t)((info >> 8) & 0xFFu);
            const uint32_t wrow = SEP ? LUT_BASE : LUT_BASE + ((info >> 16) << 10);
            const float g = SEP ? gtap[c + u] : 1.0f;
            const uint32_t a = lane + (uint32_t)((dy * G::TCOLS + dx) * 4);
#pragma unroll
            for (int j = 0; j < CPL; j++) {
                const uint32_t p = lds_u32(a + 4 * j);
                const float pf = (float)(p & 0xFFu);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float f = lds_f32(__builtin_amdgcn_sad_u8(p, cen[k][j], wrow));
                    if (SEP) f = g * f;   // factor = g[col] * E[d] (BilateralFilter.cu:125), one rounding
                    t[k][j] = t[k][j] + f * pf;
                    sum[k][j] = sum[k][j] + f;
                }
            }
        }
    }
    const int y0 = Y0 + 4 * gy;
#pragma unroll
    for (int j = 0; j < CPL; j++) {
        const int x = X0 + CPL * lx + j;
        if (x < job.W) {
            uint8_t *o = job.out + (int64_t)y0 * job.out_pitch + x;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (y0 + k < job.H) o[k * job.out_pitch] = to_u8(t[k][j] / sum[k][j]);
        }
    }
}

template <int TW, int TGY, int HALO, int CPL, bool SEP>
__device__ __forceinline__ void tile_std(const DevJob &job, int X0, int Y0, int lx, int gy)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr uint32_t LUT_BASE = kTileOff + G::TILE_BYTES;
    constexpr int ROWB = G::TCOLS * 4;
    const uint32_t lane = kTileOff + (uint32_t)(((4 * gy + HALO) * G::TCOLS + CPL * lx + HALO) * 4);
    uint32_t cen[4][CPL];
    float t[4][CPL], sum[4][CPL];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int j = 0; j < CPL; j++) {
            cen[k][j] = lds_u32(lane + k * ROWB + 4 * j);
            t[k][j] = 0.0f;
            sum[k][j] = 0.0f;
        }
    const BF_CONST uint32_t *taps = as_const(job.taps);
    const BF_CONST float *gtap = as_const(job.gtap);
    const int nt = job.ntaps_padded;
    for (int c = 0; c < nt; c += kTapPad) {
#pragma unroll
        for (int u = 0; u < kTapPad; u++) {
            const uint32_t info = taps[c + u];
            const int dy = (int)(int8_t)(info & 0xFFu);
            const int dx = (int)(int8_t)((info >> 8) & 0xFFu);
            const uint32_t wrow = SEP ? LUT_BASE : LUT_BASE + ((info >> 16) << 10);
            const float g = SEP ? gtap[c + u] : 1.0f;
            const uint32_t a = lane + (uint32_t)((dy * G::TCOLS + dx) * 4);
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < CPL; j++) {
                    const uint32_t p = lds_u32(a + k * ROWB + 4 * j);
                    float f = lds_f32(__builtin_amdgcn_sad_u8(p, cen[k][j], wrow));
                    if (SEP) f = g * f;
                    t[k][j] = t[k][j] + f * (float)(p & 0xFFu);
                    sum[k][j] = sum[k][j] + f;
                }
        }
    }
    const int y0 = Y0 + 4 * gy;
#pragma unroll
    for (int j = 0; j < CPL; j++) {
        const int x = X0 + CPL * lx + j;
        if (x < job.W) {
            uint8_t *o = job.out + (int64_t)y0 * job.out_pitch + x;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (y0 + k < job.H) o[k * job.out_pitch] = to_u8(t[k][j] / sum[k][j]);
        }
    }
}

template <int TW, int TGY, int HALO, int CPL>
__global__ void __launch_bounds__(TW / CPL * TGY)
bf_exact_kernel(KArgs args)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr int LX = TW / CPL;              // lanes across the tile
    constexpr int NT = LX * TGY;
    const int tid = threadIdx.x;
    const int lx = tid % LX, gy = tid / LX;
    const int ntiles = args.tile_list ? *as_const(args.tile_count) : args.ntiles;
    const BF_CONST DevJob *jobs = as_const(args.jobs);
    int cur_key = -1;
    int iter = 0;
    for (;;) {
        int tile;
        if (args.counter) {
            __syncthreads();                       // everyone done with the slot
            if (tid == 0) lds_store_u32(0, (uint32_t)atomicAdd(args.counter, 1));
            __syncthreads();
            tile = __builtin_amdgcn_readfirstlane((int)lds_u32(0));
        } else {
            tile = (int)blockIdx.x + iter * (int)gridDim.x;
        }
        iter++;
        if (tile >= ntiles) break;
        if (args.tile_list) tile = as_const(args.tile_list)[tile];
        DevJob job;
        if (args.njobs == 1) {
            job = args.job0;
        } else {
            const BF_CONST uint32_t *src =
                (const BF_CONST uint32_t *)(jobs + find_job(jobs, args.njobs, tile));
            uint32_t *dst = reinterpret_cast<uint32_t *>(&job);
#pragma unroll
            for (int i = 0; i < (int)(sizeof(DevJob) / 4); i++) dst[i] = src[i];
        }
        const int t = tile - job.tile_begin;
        const int ty = t / job.tiles_x;
        const int X0 = (t - ty * job.tiles_x) * TW;
        const int Y0 = ty * G::TH;
        __syncthreads();                           // previous tile fully consumed
        if (job.lut_key != cur_key) {
            build_lut<NT>(job, kTileOff + (uint32_t)G::TILE_BYTES, tid);
            cur_key = job.lut_key;
        }
        stage_rect<G::TCOLS, G::TROWS, NT, false>(job, X0 - HALO, Y0 - HALO, kTileOff, tid);
        __syncthreads();
        if (job.sep) {
            if (job.mode == kModeRef)
                tile_ref<TW, TGY, HALO, CPL, true>(job, X0, Y0, lx, gy);
            else
                tile_std<TW, TGY, HALO, CPL, true>(job, X0, Y0, lx, gy);
        } else {
            if (job.mode == kModeRef)
                tile_ref<TW, TGY, HALO, CPL, false>(job, X0, Y0, lx, gy);
            else
                tile_std<TW, TGY, HALO, CPL, false>(job, X0, Y0, lx, gy);
        }
    }
}

// One image, static grid-stride over its tiles (or over the tiles of a
// device list, whose length is read on the device); job.rowg holds the
// per-q weights of the canonical template and job.nrows their count.
template <int R, int MODE, int TW, int TGY, int HALO>
__global__ void __launch_bounds__(TW * TGY)
bf_ring_kernel(DevJob job, int ntiles, const int32_t *list, const int32_t *list_count)
{
    using G = Geo<TW, TGY, HALO>;
    const int tid = threadIdx.x;
    const int lx = tid % TW, gy = tid / TW;
    build_lut<G::NT>(job, kTileOff + (uint32_t)G::TILE_BYTES, tid);
    const int n = list ? *as_const(list_count) : ntiles;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int tile = list ? as_const(list)[i] : i;
        const int ty = tile / job.tiles_x;
        const int X0 = (tile - ty * job.tiles_x) * TW;
        const int Y0 = ty * G::TH;
        __syncthreads();                      // table built / previous tile consumed
        load_tile<TW, TGY, HALO>(job, X0, Y0, tid);
        __syncthreads();
        tile_ring<R, MODE, TW, TGY, HALO>(job, X0, Y0, lx, gy);
    }
}

// ------------------------------------------------------ exact sample check
//
// For each 64x64 tile of the output view, 64 samples (one per 8x8 cell at a
// hashed position) are filtered exactly -- same taps, order and rounding as
// the exact kernels -- and compared with the u8 already in the output (the
// surrogate's); a tile with any sampled |difference| > tol is appended to
// the fallback list.

__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

// Sample of cell `cell` (0..63) of tile `tile`: offsets inside the tile.
__device__ __forceinline__ void check_sample(int tile, int cell, uint32_t salt, int *sx, int *sy)
{
    const uint32_t h = hash32((uint32_t)tile * 64u + (uint32_t)cell + salt);
    *sx = (cell & 7) * 8 + (int)(h & 7);
    *sy = (cell >> 3) * 8 + (int)((h >> 3) & 7);
}

__device__ __forceinline__ void check_verdict(const CheckArgs &a, int tile, bool bad)
{
    if (__any(bad) && (threadIdx.x & 63) == 0) {
        const int slot = atomicAdd(a.count, 1);
        a.list[slot] = tile;
        if (a.accum) atomicAdd(a.accum, 1);
    }
    if ((threadIdx.x & 63) == 0 && a.tiles_accum) atomicAdd(a.tiles_accum, 1);
}

// Any template: one wave per tile, one sample per lane.  The tile, its
// context rows and columns (clamped to the input view, like the exact
// kernels) and the range table are staged in LDS; each lane then runs the
// exact tap loop of the run-time tap list.
constexpr int kCheckHaloMax = 20;     // largest halo class
constexpr int kCheckSpanMax = 64 + 2 * kCheckHaloMax;

__global__ void __launch_bounds__(64) check_kernel(CheckArgs a)
{
    __shared__ uint8_t win[kCheckSpanMax * kCheckSpanMax];
    __shared__ float rng[256];
    const int tile = blockIdx.x;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const DevJob &j = a.job;
    const int halo = a.halo, span = 64 + 2 * halo;
    const int gx0 = tx * 64 - halo, gy0 = ty * 64 - halo;
    for (int i = threadIdx.x; i < 256; i += 64) rng[i] = j.range[i];
    for (int yy = 0; yy < span; yy++) {
        const int gy = clampi(gy0 + yy, 0, j.in_h - 1);
        const unsigned char *src = j.in + (int64_t)gy * j.in_pitch;
        for (int xx = threadIdx.x; xx < span; xx += 64)
            win[yy * span + xx] = src[clampi(gx0 + xx, 0, j.in_w - 1)];
    }
    __syncthreads();
    int sx, sy;
    check_sample(tile, threadIdx.x, a.salt, &sx, &sy);
    const int x = tx * 64 + sx, y = ty * 64 + sy;
    bool bad = false;
    if (x < j.W && y < j.H) {
        const int y0 = j.mode == kModeRef ? y - (y & 3) : y;
        const int base = (y0 - gy0) * span + (x - gx0);
        const uint32_t c = win[(y - gy0) * span + (x - gx0)];
        float t = 0.0f, s = 0.0f;
        for (int k = 0; k < j.ntaps_padded; k++) {
            const uint32_t info = j.taps[k];
            const int dy = (int)(int8_t)(info & 0xFFu), dx = (int)(int8_t)((info >> 8) & 0xFFu);
            const int row = (int)(info >> 16);
            const uint32_t p = win[base + dy * span + dx];
            const uint32_t d = p > c ? p - c : c - p;
            const float g = row < j.nrows ? a.rowg[row] : 0.0f;
            const float f = g * rng[d];
            t = t + f * (float)p;
            s = s + f;
        }
        const int ex = to_u8(t / s);
        const int su = j.out[(int64_t)y * j.out_pitch + x];
        bad = (ex - su > a.tol) || (su - ex > a.tol);
    }
    check_verdict(a, tile, bad);
}

// The canonical ring template, unrolled: a persistent workgroup of four
// waves stages the range table E[d] once, then each wave takes one tile at
// a time, stages its context as bytes in its own LDS window and runs the
// compiled tap list for its 64 samples (tap offsets and weight rows are
// immediates, as in bf_ring_kernel).  The factor is formed per tap as
// g * E[d] -- the product the fused tables hold, with the same single
// rounding -- with g, the same for every lane, read from the job's weight
// row by a scalar load (an LDS copy cost one more LDS read per tap: 0.40 ms
// against 0.37 on c3).  LDS holds 4 windows + 1 KB, so several workgroups
// share a CU (the taps' LDS latency needs the waves).
constexpr int kCheckGroup = 8;            // taps per scheduling window (4 to 32 measured alike)

template <int R, int MODE, int HALO>
__global__ void __launch_bounds__(256) ring_check_kernel(CheckArgs a, int ntiles)
{
    constexpr int SPAN = 64 + 2 * HALO;
    constexpr uint32_t WIN = SPAN * SPAN;
    constexpr uint32_t E_BASE = kTileOff + 4 * WIN;
    constexpr int T = RingTaps<R, MODE>::T;
    constexpr int MINB = ring_min_offset<R, MODE, SPAN>();
    const DevJob &j = a.job;
    const int tid = threadIdx.x, wave = tid >> 6, cell = tid & 63;
    const uint32_t win = kTileOff + (uint32_t)wave * WIN;
    for (int d = tid; d < 256; d += 256) lds_store_u32(E_BASE + 4 * d, __float_as_uint(j.range[d]));
    const int groups = (ntiles + 3) / 4;
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int tile = 4 * g + wave;
        const bool live = tile < ntiles;
        const int X0 = live ? tile % a.tiles_x * 64 : 0, Y0 = live ? tile / a.tiles_x * 64 : 0;
        __syncthreads();                      // tables staged / previous windows consumed
        if (live) stage_rect<SPAN, SPAN, 64, true>(j, X0 - HALO, Y0 - HALO, win, cell);
        __syncthreads();
        if (!live) continue;
        int sx, sy;
        check_sample(tile, cell, a.salt, &sx, &sy);
        const int x = X0 + sx, y = Y0 + sy;
        bool bad = false;
        if (x < j.W && y < j.H) {
            // REF: the taps are centred on the first row of the sample's
            // 4-row group, the range weights on the sample's own pixel
            const int wy = MODE == kModeRef ? sy - (sy & 3) : sy;
            const uint32_t c4 = lds_u8(win + (uint32_t)((sy + HALO) * SPAN + sx + HALO)) * 0x01010101u;
            const uint32_t tbase = win + (uint32_t)((wy + HALO) * SPAN + sx + HALO + MINB);
            float t = 0.0f, s = 0.0f;
            static_for<T>([&](auto ci) {
                constexpr RingTap tp = kRing<R, MODE>.t[decltype(ci)::value];
                constexpr int TAPB = tp.dy * SPAN + tp.dx - MINB;   // >= 0
                const uint32_t p = lds_u8(tbase + TAPB);
                // v_perm_b32 with selector 0 replicates byte 0 (one full-rate op)
                const uint32_t p4 = __builtin_amdgcn_perm(p, p, 0u);
                const float e = lds_f32(__builtin_amdgcn_sad_u8(p4, c4, E_BASE));
                const float f = j.rowg[tp.row] * e;
                t = t + f * (float)(p4 & 0xFFu);
                s = s + f;
                if constexpr (decltype(ci)::value % kCheckGroup == kCheckGroup - 1)
                    asm volatile("" : "+v"(t), "+v"(s)::"memory");
            });
            const int ex = to_u8(t / s);
            const int su = j.out[(int64_t)y * j.out_pitch + x];
            bad = (ex - su > a.tol) || (su - ex > a.tol);
        }
        check_verdict(a, tile, bad);
    }
}

__global__ void reset_counter_kernel(int32_t *counter) { *counter = 0; }

// ---------------------------------------------------------------- configs

// tile width, 4-row groups, threads, columns per lane
constexpr TileConfig kConfigs[] = {
    { 64, 16, 1024, 1 },   // 0: 64 x 64 tile
    { 64, 8, 512, 1 },     // 1: 64 x 32
    { 32, 8, 256, 1 },     // 2: 32 x 32 (the reference's block footprint)
    { 64, 4, 256, 1 },     // 3: 64 x 16
    { 128, 4, 512, 1 },    // 4: 128 x 16
    { 128, 8, 1024, 1 },   // 5: 128 x 32
    { 64, 16, 512, 2 },    // 6: 64 x 64, two columns per lane
    { 128, 8, 512, 2 },    // 7: 128 x 32, two columns per lane
    { 128, 4, 256, 2 },    // 8: 128 x 16, two columns per lane
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);
constexpr int kHalos[] = { 4, 8, 12, 16, 20 };

template <int TW, int TGY, int HALO, int CPL>
int launch_one(const KArgs &a, int max_lut_rows, hipStream_t s)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr int NT = TW / CPL * TGY;
    const size_t lds = kTileOff + (size_t)G::TILE_BYTES + (size_t)max_lut_rows * 1024;
    constexpr auto kern = bf_exact_kernel<TW, TGY, HALO, CPL>;
    BF_HIP(allow_full_lds<kern>());
    const int res = resident_groups(reinterpret_cast<const void *>(kern), NT, lds);
    if (res < 0) return hip_fail(hipGetLastError(), "occupancy query");
    if (res < 1) {
        set_error("exact kernel does not fit on a CU (LDS " + std::to_string(lds) + " bytes)");
        return kInvalidData;
    }
    int grid = res;
    if (!a.tile_list && !a.counter && a.ntiles < grid) grid = a.ntiles;
    if (a.tile_list && a.ntiles < grid) grid = a.ntiles;   // ntiles = max possible
    if (grid < 1) return kOk;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, s, a);
    BF_HIP(hipGetLastError());
    return kOk;
}

template <int TW, int TGY, int CPL = 1>
int launch_halo(int halo, const KArgs &a, int max_lut_rows, hipStream_t s)
{
    switch (halo) {
    case 4: return launch_one<TW, TGY, 4, CPL>(a, max_lut_rows, s);
    case 8: return launch_one<TW, TGY, 8, CPL>(a, max_lut_rows, s);
    case 12: return launch_one<TW, TGY, 12, CPL>(a, max_lut_rows, s);
    case 16: return launch_one<TW, TGY, 16, CPL>(a, max_lut_rows, s);
    case 20: return launch_one<TW, TGY, 20, CPL>(a, max_lut_rows, s);
    default: set_error("bad halo class"); return kInvalidData;
    }
}

}  // namespace

int tile_config_count() { return kNumConfigs; }

const TileConfig &tile_config(int id) { return kConfigs[id]; }

int halo_class(int radius)
{
    for (int h : kHalos)
        if (radius <= h) return h;
    return -1;
}

// Measured on MI355X (8192-wide image): 64x64 tiles win at r >= 13 in both
// modes, 64x32 tiles below (REF r = 3..9, STD r = 7..9).
int default_config(int radius, int mode)
{
    (void)mode;
    return radius >= 13 ? 0 : 1;
}

size_t lds_bytes(int config, int halo, int lut_rows)
{
    const TileConfig &c = kConfigs[config];
    return kTileOff + (size_t)(c.tw + 2 * halo) * (size_t)(4 * c.tgy + 2 * halo) * 4 +
           (size_t)lut_rows * 1024;
}

size_t max_lds_bytes() { return 160 * 1024; }

int launch_exact(int config, int halo, const DevJob &job0, const DevJob *jobs_dev, int njobs,
                 int ntiles, int max_lut_rows, int *counter, hipStream_t s)
{
    KArgs a{};
    a.job0 = job0;
    a.jobs = jobs_dev;
    a.njobs = njobs;
    a.ntiles = ntiles;
    a.tile_list = nullptr;
    a.tile_count = nullptr;
    a.counter = counter;
    switch (config) {
    case 0: return launch_halo<64, 16>(halo, a, max_lut_rows, s);
    case 1: return launch_halo<64, 8>(halo, a, max_lut_rows, s);
    case 2: return launch_halo<32, 8>(halo, a, max_lut_rows, s);
    case 3: return launch_halo<64, 4>(halo, a, max_lut_rows, s);
    case 4: return launch_halo<128, 4>(halo, a, max_lut_rows, s);
    case 5: return launch_halo<128, 8>(halo, a, max_lut_rows, s);
    case 6: return launch_halo<64, 16, 2>(halo, a, max_lut_rows, s);
    case 7: return launch_halo<128, 8, 2>(halo, a, max_lut_rows, s);
    case 8: return launch_halo<128, 4, 2>(halo, a, max_lut_rows, s);
    default: set_error("bad tile config"); return kInvalidData;
    }
}

int exact_tiles(const DevJob &job, const int32_t *tile_list_dev, int ntiles_max,
                const int32_t *count_dev, int config, int halo, hipStream_t s)
{
    KArgs a{};
    a.job0 = job;
    a.jobs = nullptr;
    a.njobs = 1;
    a.ntiles = ntiles_max;
    a.tile_list = tile_list_dev;
    a.tile_count = count_dev;
    a.counter = nullptr;
    switch (config) {
    case 0: return launch_halo<64, 16>(halo, a, job.lut_rows, s);
    case 1: return launch_halo<64, 8>(halo, a, job.lut_rows, s);
    default: set_error("fallback needs a 64-wide tile config"); return kInvalidData;
    }
}

namespace {

template <int R, int MODE, int TW, int TGY, int HALO>
int launch_ring_one(const DevJob &job, int ntiles, const int32_t *list, const int32_t *count,
                    hipStream_t s)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr auto kern = bf_ring_kernel<R, MODE, TW, TGY, HALO>;
    const size_t lds = kTileOff + (size_t)G::TILE_BYTES + (size_t)(job.nrows + 1) * 1024;
    BF_HIP(allow_full_lds<kern>());
    const int res = resident_groups(reinterpret_cast<const void *>(kern), G::NT, lds);
    if (res < 0) return hip_fail(hipGetLastError(), "occupancy query");
    if (res < 1) {
        set_error("ring kernel does not fit on a CU");
        return kInvalidData;
    }
    const int grid = std::min(ntiles, res);   // with a list, ntiles = its maximum length
    if (grid < 1) return kOk;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(G::NT), lds, s, job, ntiles, list, count);
    BF_HIP(hipGetLastError());
    return kOk;
}

template <int R, int MODE, int HALO>
int launch_ring_check_one(const CheckArgs &a, int ntiles, hipStream_t s)
{
    constexpr auto kern = ring_check_kernel<R, MODE, HALO>;
    const size_t lds = kTileOff + 4 * (size_t)(64 + 2 * HALO) * (64 + 2 * HALO) + 1024;
    BF_HIP(allow_full_lds<kern>());
    const int res = resident_groups(reinterpret_cast<const void *>(kern), 256, lds);
    if (res < 0) return hip_fail(hipGetLastError(), "occupancy query");
    if (res < 1) {
        set_error("check kernel does not fit on a CU");
        return kInvalidData;
    }
    const int grid = std::min((ntiles + 3) / 4, res);
    if (grid < 1) return kOk;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, a, ntiles);
    BF_HIP(hipGetLastError());
    return kOk;
}

}  // namespace

bool ring_available(int radius, int mode, int *config)
{
    switch (radius) {
    case 15: *config = 0; return true;
    case 3: case 5: case 7: case 9: *config = 1; return true;
    default: (void)mode; return false;
    }
}

int ring_rows(int radius, int mode)
{
    switch (radius * 2 + (mode == kModeRef ? 0 : 1)) {
    case 30: return RingTaps<15, 0>().nrows;
    case 31: return RingTaps<15, 1>().nrows;
    case 6: return RingTaps<3, 0>().nrows;
    case 7: return RingTaps<3, 1>().nrows;
    case 10: return RingTaps<5, 0>().nrows;
    case 11: return RingTaps<5, 1>().nrows;
    case 14: return RingTaps<7, 0>().nrows;
    case 15: return RingTaps<7, 1>().nrows;
    case 18: return RingTaps<9, 0>().nrows;
    case 19: return RingTaps<9, 1>().nrows;
    default: return -1;
    }
}

int launch_ring(const DevJob &job, int radius, int mode, int ntiles, hipStream_t s)
{
    const bool ref = mode == kModeRef;
    switch (radius) {
    case 15: return ref ? launch_ring_one<15, 0, 64, 16, 16>(job, ntiles, nullptr, nullptr, s)
                        : launch_ring_one<15, 1, 64, 16, 16>(job, ntiles, nullptr, nullptr, s);
    case 9: return ref ? launch_ring_one<9, 0, 64, 8, 12>(job, ntiles, nullptr, nullptr, s)
                       : launch_ring_one<9, 1, 64, 8, 12>(job, ntiles, nullptr, nullptr, s);
    case 7: return ref ? launch_ring_one<7, 0, 64, 8, 8>(job, ntiles, nullptr, nullptr, s)
                       : launch_ring_one<7, 1, 64, 8, 8>(job, ntiles, nullptr, nullptr, s);
    case 5: return ref ? launch_ring_one<5, 0, 64, 8, 8>(job, ntiles, nullptr, nullptr, s)
                       : launch_ring_one<5, 1, 64, 8, 8>(job, ntiles, nullptr, nullptr, s);
    case 3: return ref ? launch_ring_one<3, 0, 64, 8, 4>(job, ntiles, nullptr, nullptr, s)
                       : launch_ring_one<3, 1, 64, 8, 4>(job, ntiles, nullptr, nullptr, s);
    default: set_error("no unrolled kernel for this radius"); return kInvalidData;
    }
}

bool ring64_available(int radius) { return radius == 15 || radius == 9; }

int ring64_tiles(const DevJob &job, int radius, int mode, const int32_t *list,
                 const int32_t *count, int ntiles_max, hipStream_t s)
{
    const bool ref = mode == kModeRef;
    switch (radius) {
    case 15: return ref ? launch_ring_one<15, 0, 64, 16, 16>(job, ntiles_max, list, count, s)
                        : launch_ring_one<15, 1, 64, 16, 16>(job, ntiles_max, list, count, s);
    case 9: return ref ? launch_ring_one<9, 0, 64, 16, 12>(job, ntiles_max, list, count, s)
                       : launch_ring_one<9, 1, 64, 16, 12>(job, ntiles_max, list, count, s);
    default: set_error("no unrolled 64x64 kernel for this radius"); return kInvalidData;
    }
}

int launch_check(const CheckArgs &a, int ring_radius, int ntiles, hipStream_t s)
{
    const bool ref = a.job.mode == kModeRef;
    switch (ring_radius) {
    case 15: return ref ? launch_ring_check_one<15, 0, 16>(a, ntiles, s)
                        : launch_ring_check_one<15, 1, 16>(a, ntiles, s);
    case 9: return ref ? launch_ring_check_one<9, 0, 12>(a, ntiles, s)
                       : launch_ring_check_one<9, 1, 12>(a, ntiles, s);
    case 0:
        if (a.halo < 0 || a.halo > kCheckHaloMax) {
            set_error("radius outside the check kernel's halo");
            return kInvalidData;
        }
        if (ntiles < 1) return kOk;
        hipLaunchKernelGGL(check_kernel, dim3(ntiles), dim3(64), 0, s, a);
        BF_HIP(hipGetLastError());
        return kOk;
    default: set_error("no unrolled check kernel for this radius"); return kInvalidData;
    }
}

int launch_reset_counter(int *counter, hipStream_t s)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, counter);
    BF_HIP(hipGetLastError());
    return kOk;
}

}  // namespace bfimpl
