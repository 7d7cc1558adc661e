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

namespace bfimpl {

namespace {

constexpr int kModeRef = BF_MODE_REF;

struct KArgs {
    DevJob job0;                 // used when njobs == 1
    const DevJob *jobs;          // device job table (njobs > 1)
    int32_t njobs;
    int32_t ntiles;              // static tile count (no tile_list)
    const int32_t *tile_list;    // optional: global tile ids to run
    const int32_t *tile_count;   // device count of tile_list entries
    int32_t *counter;            // optional work-queue counter
    int32_t pad;
};

// Dynamic LDS layout (no static __shared__ objects, so it starts at LDS
// address 0 and every gather address is a plain byte offset):
//   [0, 16)                  work-queue slot
//   [16, 16 + TILE_BYTES)    staged input tile (byte-replicated u32)
//   [16 + TILE_BYTES, ...)   fused weight table, (nrows + 1) x 256 floats
constexpr uint32_t kTileOff = 16;

// LDS is addressed by byte offset through address-space-3 pointers.  The
// kernel declares no static LDS, so the dynamic allocation starts at LDS
// address 0 (checked by the .group_segment_fixed_size of 0 in the code
// object) and a v_sad_u8 result is usable as a ds_read address unchanged.
typedef __attribute__((address_space(3))) uint32_t lds_word;
typedef __attribute__((address_space(3))) float lds_float;

__device__ __forceinline__ uint32_t lds_u32(uint32_t byte_addr)
{
    return *(const lds_word *)(size_t)byte_addr;
}

__device__ __forceinline__ float lds_f32(uint32_t byte_addr)
{
    return *(const lds_float *)(size_t)byte_addr;
}

__device__ __forceinline__ void lds_store_u32(uint32_t byte_addr, uint32_t v)
{
    *(lds_word *)(size_t)byte_addr = v;
}

// Read-only tables (jobs, taps) are read through the constant address
// space: they never change during a launch, and this lets the compiler use
// scalar loads for the wave-uniform tap stream.
#define BF_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const BF_CONST T *as_const(const T *p)
{
    return (const BF_CONST T *)(p);
}

__device__ __forceinline__ uint8_t to_u8(float q)
{
    if (!(q > 0.0f)) return 0;          // NaN (0/0) and non-positive
    if (q >= 255.0f) return 255;
    return (uint8_t)(int)q;             // truncation toward zero
}

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

// Find the job owning global tile `tile` (jobs sorted by tile_begin).
__device__ __forceinline__ int find_job(const BF_CONST DevJob *jobs, int njobs, int tile)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int TW, int TGY, int HALO>
struct Geo {
    static constexpr int TH = 4 * TGY;
    static constexpr int TCOLS = TW + 2 * HALO;
    static constexpr int TROWS = TH + 2 * HALO;
    static constexpr int TILE_BYTES = TCOLS * TROWS * 4;
    static constexpr int NT = TW * TGY;
};

// Stage the clamped input tile (rows Y0-HALO .., cols X0-HALO ..) into LDS
// as byte-replicated u32.
template <int TW, int TGY, int HALO>
__device__ __forceinline__ void load_tile(const DevJob &job, int X0, int Y0, int tid)
{
    using G = Geo<TW, TGY, HALO>;
    const int xs = X0 - HALO;
    const int ys = Y0 - HALO;
    const int wmax = job.in_w - 1, hmax = job.in_h - 1;
    if (xs >= 0 && xs + G::TCOLS <= job.in_w) {
        // Interior columns: aligned 16-byte loads covering [xs, xs+TCOLS).
        // Bytes outside that span are dropped; they lie in the same aligned
        // 16-byte block as a byte of the row, hence in mapped memory.
        constexpr int NCH = (G::TCOLS + 15) / 16 + 1;  // chunks per row
        for (int e = tid; e < G::TROWS * NCH; e += G::NT) {
            const int r = e / NCH, ch = e - r * NCH;
            const int gy = clampi(ys + r, 0, hmax);
            const uint8_t *row = job.in + (int64_t)gy * job.in_pitch;
            const uintptr_t a0 = reinterpret_cast<uintptr_t>(row + xs) & ~(uintptr_t)15;
            const uintptr_t ca = a0 + 16 * (uintptr_t)ch;
            const intptr_t first = (intptr_t)ca - (intptr_t)reinterpret_cast<uintptr_t>(row + xs);
            if (first >= G::TCOLS) continue;
            const uint4 v = *reinterpret_cast<const uint4 *>(ca);
            const uint32_t w[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
            for (int b = 0; b < 16; b++) {
                const int c = (int)first + b;
                if (c >= 0 && c < G::TCOLS) {
                    const uint32_t byte = (w[b >> 2] >> (8 * (b & 3))) & 0xFFu;
                    lds_store_u32(kTileOff + (uint32_t)((r * G::TCOLS + c) * 4), byte * 0x01010101u);
                }
            }
        }
    } else {
        for (int e = tid; e < G::TROWS * G::TCOLS; e += G::NT) {
            const int r = e / G::TCOLS, c = e - r * G::TCOLS;
            const int gy = clampi(ys + r, 0, hmax);
            const int gx = clampi(xs + c, 0, wmax);
            const uint32_t byte = job.in[(int64_t)gy * job.in_pitch + gx];
            lds_store_u32(kTileOff + (uint32_t)(e * 4), byte * 0x01010101u);
        }
    }
}

// Fused weight table: W[row][d] = rowg[row] * range[d]; row `nrows` = 0.
// Separate-table mode: just E[d].
template <int NT>
__device__ __forceinline__ void build_lut(const DevJob &job, uint32_t lut_base, int tid)
{
    if (job.sep) {
        for (int d = tid; d < 256; d += NT)
            lds_store_u32(lut_base + (uint32_t)d * 4, __float_as_uint(job.range[d]));
        return;
    }
    const int n = (job.nrows + 1) * 256;
    for (int e = tid; e < n; e += NT) {
        const int row = e >> 8, d = e & 255;
        float w = 0.0f;
        if (row < job.nrows) w = job.rowg[row] * job.range[d];
        lds_store_u32(lut_base + (uint32_t)e * 4, __float_as_uint(w));
    }
}

template <int TW, int TGY, int HALO, bool SEP>
__device__ __forceinline__ void tile_ref(const DevJob &job, int X0, int Y0, int lx, int gy)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr uint32_t LUT_BASE = kTileOff + G::TILE_BYTES;
    constexpr int ROWB = G::TCOLS * 4;
    const uint32_t lane = kTileOff + (uint32_t)(((4 * gy + HALO) * G::TCOLS + lx + HALO) * 4);
    const uint32_t c0 = lds_u32(lane), c1 = lds_u32(lane + ROWB);
    const uint32_t c2 = lds_u32(lane + 2 * ROWB), c3 = lds_u32(lane + 3 * ROWB);
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
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
            const uint32_t p = lds_u32(lane + (uint32_t)((dy * G::TCOLS + dx) * 4));
            const float pf = (float)(p & 0xFFu);
            float f0 = lds_f32(__builtin_amdgcn_sad_u8(p, c0, wrow));
            float f1 = lds_f32(__builtin_amdgcn_sad_u8(p, c1, wrow));
            float f2 = lds_f32(__builtin_amdgcn_sad_u8(p, c2, wrow));
            float f3 = lds_f32(__builtin_amdgcn_sad_u8(p, c3, wrow));
            if (SEP) {   // factor = g[col] * E[d] (BilateralFilter.cu:125), one rounding
                f0 = g * f0; f1 = g * f1; f2 = g * f2; f3 = g * f3;
            }
            t0 = t0 + f0 * pf; s0 = s0 + f0;
            t1 = t1 + f1 * pf; s1 = s1 + f1;
            t2 = t2 + f2 * pf; s2 = s2 + f2;
            t3 = t3 + f3 * pf; s3 = s3 + f3;
        }
    }
    const int x = X0 + lx, y0 = Y0 + 4 * gy;
    if (x < job.W) {
        uint8_t *o = job.out + (int64_t)y0 * job.out_pitch + x;
        if (y0 < job.H) o[0] = to_u8(t0 / s0);
        if (y0 + 1 < job.H) o[job.out_pitch] = to_u8(t1 / s1);
        if (y0 + 2 < job.H) o[2 * job.out_pitch] = to_u8(t2 / s2);
        if (y0 + 3 < job.H) o[3 * job.out_pitch] = to_u8(t3 / s3);
    }
}

template <int TW, int TGY, int HALO, bool SEP>
__device__ __forceinline__ void tile_std(const DevJob &job, int X0, int Y0, int lx, int gy)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr uint32_t LUT_BASE = kTileOff + G::TILE_BYTES;
    constexpr int ROWB = G::TCOLS * 4;
    const uint32_t lane = kTileOff + (uint32_t)(((4 * gy + HALO) * G::TCOLS + lx + HALO) * 4);
    const uint32_t c0 = lds_u32(lane), c1 = lds_u32(lane + ROWB);
    const uint32_t c2 = lds_u32(lane + 2 * ROWB), c3 = lds_u32(lane + 3 * ROWB);
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
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
            const uint32_t p0 = lds_u32(a), p1 = lds_u32(a + ROWB);
            const uint32_t p2 = lds_u32(a + 2 * ROWB), p3 = lds_u32(a + 3 * ROWB);
            float f0 = lds_f32(__builtin_amdgcn_sad_u8(p0, c0, wrow));
            float f1 = lds_f32(__builtin_amdgcn_sad_u8(p1, c1, wrow));
            float f2 = lds_f32(__builtin_amdgcn_sad_u8(p2, c2, wrow));
            float f3 = lds_f32(__builtin_amdgcn_sad_u8(p3, c3, wrow));
            if (SEP) {
                f0 = g * f0; f1 = g * f1; f2 = g * f2; f3 = g * f3;
            }
            t0 = t0 + f0 * (float)(p0 & 0xFFu); s0 = s0 + f0;
            t1 = t1 + f1 * (float)(p1 & 0xFFu); s1 = s1 + f1;
            t2 = t2 + f2 * (float)(p2 & 0xFFu); s2 = s2 + f2;
            t3 = t3 + f3 * (float)(p3 & 0xFFu); s3 = s3 + f3;
        }
    }
    const int x = X0 + lx, y0 = Y0 + 4 * gy;
    if (x < job.W) {
        uint8_t *o = job.out + (int64_t)y0 * job.out_pitch + x;
        if (y0 < job.H) o[0] = to_u8(t0 / s0);
        if (y0 + 1 < job.H) o[job.out_pitch] = to_u8(t1 / s1);
        if (y0 + 2 < job.H) o[2 * job.out_pitch] = to_u8(t2 / s2);
        if (y0 + 3 < job.H) o[3 * job.out_pitch] = to_u8(t3 / s3);
    }
}

template <int TW, int TGY, int HALO>
__global__ void __launch_bounds__(TW * TGY)
bf_exact_kernel(KArgs args)
{
    using G = Geo<TW, TGY, HALO>;
    const int tid = threadIdx.x;
    const int lx = tid % TW, gy = tid / TW;
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
            build_lut<G::NT>(job, kTileOff + (uint32_t)G::TILE_BYTES, tid);
            cur_key = job.lut_key;
        }
        load_tile<TW, TGY, HALO>(job, X0, Y0, tid);
        __syncthreads();
        if (job.sep) {
            if (job.mode == kModeRef)
                tile_ref<TW, TGY, HALO, true>(job, X0, Y0, lx, gy);
            else
                tile_std<TW, TGY, HALO, true>(job, X0, Y0, lx, gy);
        } else {
            if (job.mode == kModeRef)
                tile_ref<TW, TGY, HALO, false>(job, X0, Y0, lx, gy);
            else
                tile_std<TW, TGY, HALO, false>(job, X0, Y0, lx, gy);
        }
    }
}

// ------------------------------------------------------------------------
// Unrolled variant for the canonical ring template (the one the GAUSS
// factory emits, TemplateFactory.cu:1092-1168).  The tap list of a radius is
// known at compile time, so every tap's tile offset and fused-table row are
// immediates: the per-fetch address add disappears (ds_read offsets), and
// the compiler schedules across taps freely.  The fused table has one row
// per distinct q = dy^2 + dx^2 in order of first appearance; the host checks
// that the template really is the canonical one (same offsets, one weight
// per q) before choosing this kernel.

struct RingTap {
    int dy, dx, row;
};

template <int R, int MODE>
struct RingTaps {
    static constexpr int T = (2 * R + 1) * (2 * R + 1) + (MODE == kModeRef ? 1 : 0);
    RingTap t[T];
    int nrows;
    constexpr RingTaps() : t{}, nrows(0)
    {
        int qs[T] = {};
        int n = 0;
        for (int i = 0; i <= R + 1 && n < T; i++) {
            for (int j = (i == 0 ? 0 : -i + 1); j <= i && n < T; j++) {
                const int pts[4][2] = { { i, j }, { -i, j }, { j, i }, { j, -i } };
                for (int k = 0; k < (i == 0 ? 1 : 4) && n < T; k++) {
                    const int dy = pts[k][0], dx = pts[k][1], q = dy * dy + dx * dx;
                    int row = -1;
                    for (int m = 0; m < nrows; m++)
                        if (qs[m] == q) row = m;
                    if (row < 0) {
                        qs[nrows] = q;
                        row = nrows++;
                    }
                    t[n++] = RingTap{ dy, dx, row };
                }
            }
        }
    }
};

template <int R, int MODE>
inline constexpr RingTaps<R, MODE> kRing{};

// Smallest tap offset (in tile elements) of a radius: the unrolled kernels
// address every fetch from one base placed there, so all ds_read offsets
// are non-negative immediates and no per-tap base register is needed.
template <int R, int MODE, int TCOLS>
constexpr int ring_min_offset()
{
    int m = 0;
    for (int c = 0; c < RingTaps<R, MODE>::T; c++) {
        const int o = kRing<R, MODE>.t[c].dy * TCOLS + kRing<R, MODE>.t[c].dx;
        m = o < m ? o : m;
    }
    return m;
}

// Taps scheduled together in the unrolled kernels (STD has 4 fetches per
// tap, so it keeps fewer taps in flight).
template <int MODE>
constexpr int ring_group() { return MODE == kModeRef ? 4 : 2; }

template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

template <int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int R, int MODE, int TW, int TGY, int HALO>
__device__ __forceinline__ void tile_ring(const DevJob &job, int X0, int Y0, int lx, int gy)
{
    using G = Geo<TW, TGY, HALO>;
    constexpr uint32_t LUT_BASE = kTileOff + G::TILE_BYTES;
    constexpr int ROWB = G::TCOLS * 4;
    constexpr int T = RingTaps<R, MODE>::T;
    const uint32_t lane = kTileOff + (uint32_t)(((4 * gy + HALO) * G::TCOLS + lx + HALO) * 4);
    const uint32_t c0 = lds_u32(lane), c1 = lds_u32(lane + ROWB);
    const uint32_t c2 = lds_u32(lane + 2 * ROWB), c3 = lds_u32(lane + 3 * ROWB);
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    constexpr int MINB = ring_min_offset<R, MODE, G::TCOLS>() * 4;
    const uint32_t tbase = lane + (uint32_t)MINB;   // >= 0: the halo covers every tap
    // Gathers address W[row][d] as (4d + LUT_BASE) + row*1024; rows past 63
    // go through a second base so every offset fits the 16-bit field.
    static_for<T>([&](auto ci) {
        constexpr RingTap tp = kRing<R, MODE>.t[decltype(ci)::value];
        constexpr int TAPB = (tp.dy * G::TCOLS + tp.dx) * 4 - MINB;   // >= 0
        constexpr uint32_t WB = LUT_BASE + (tp.row >= 64 ? 65536u : 0u);
        constexpr uint32_t WOFF = (uint32_t)(tp.row & 63) * 1024u;
        if constexpr (MODE == kModeRef) {
            const uint32_t p = lds_u32(tbase + TAPB);
            const float pf = (float)(p & 0xFFu);
            const float f0 = lds_f32(__builtin_amdgcn_sad_u8(p, c0, WB) + WOFF);
            const float f1 = lds_f32(__builtin_amdgcn_sad_u8(p, c1, WB) + WOFF);
            const float f2 = lds_f32(__builtin_amdgcn_sad_u8(p, c2, WB) + WOFF);
            const float f3 = lds_f32(__builtin_amdgcn_sad_u8(p, c3, WB) + WOFF);
            t0 = t0 + f0 * pf; s0 = s0 + f0;
            t1 = t1 + f1 * pf; s1 = s1 + f1;
            t2 = t2 + f2 * pf; s2 = s2 + f2;
            t3 = t3 + f3 * pf; s3 = s3 + f3;
        } else {
            const uint32_t p0 = lds_u32(tbase + TAPB), p1 = lds_u32(tbase + TAPB + ROWB);
            const uint32_t p2 = lds_u32(tbase + TAPB + 2 * ROWB), p3 = lds_u32(tbase + TAPB + 3 * ROWB);
            const float f0 = lds_f32(__builtin_amdgcn_sad_u8(p0, c0, WB) + WOFF);
            const float f1 = lds_f32(__builtin_amdgcn_sad_u8(p1, c1, WB) + WOFF);
            const float f2 = lds_f32(__builtin_amdgcn_sad_u8(p2, c2, WB) + WOFF);
            const float f3 = lds_f32(__builtin_amdgcn_sad_u8(p3, c3, WB) + WOFF);
            t0 = t0 + f0 * (float)(p0 & 0xFFu); s0 = s0 + f0;
            t1 = t1 + f1 * (float)(p1 & 0xFFu); s1 = s1 + f1;
            t2 = t2 + f2 * (float)(p2 & 0xFFu); s2 = s2 + f2;
            t3 = t3 + f3 * (float)(p3 & 0xFFu); s3 = s3 + f3;
        }
        // Keep the schedule local (a group of taps in flight), as the loop
        // version does: without a fence the compiler issues every gather of
        // the window first and spills them.  The empty asm reads and writes
        // the accumulators and clobbers memory, so neither the adds nor the
        // LDS reads move across it; it emits no instruction.
        if constexpr (decltype(ci)::value % ring_group<MODE>() == ring_group<MODE>() - 1)
            asm volatile("" : "+v"(t0), "+v"(t1), "+v"(t2), "+v"(t3), "+v"(s0), "+v"(s1),
                         "+v"(s2), "+v"(s3)::"memory");
    });
    const int x = X0 + lx, y0 = Y0 + 4 * gy;
    if (x < job.W) {
        uint8_t *o = job.out + (int64_t)y0 * job.out_pitch + x;
        if (y0 < job.H) o[0] = to_u8(t0 / s0);
        if (y0 + 1 < job.H) o[job.out_pitch] = to_u8(t1 / s1);
        if (y0 + 2 < job.H) o[2 * job.out_pitch] = to_u8(t2 / s2);
        if (y0 + 3 < job.H) o[3 * job.out_pitch] = to_u8(t3 / s3);
    }
}

// One image, static grid-stride over its tiles; job.rowg holds the
// per-q weights of the canonical template and job.nrows their count.
template <int R, int MODE, int TW, int TGY, int HALO>
__global__ void __launch_bounds__(TW * TGY) bf_ring_kernel(DevJob job, int ntiles)
{
    using G = Geo<TW, TGY, HALO>;
    const int tid = threadIdx.x;
    const int lx = tid % TW, gy = tid / TW;
    build_lut<G::NT>(job, kTileOff + (uint32_t)G::TILE_BYTES, tid);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int ty = tile / job.tiles_x;
        const int X0 = (tile - ty * job.tiles_x) * TW;
        const int Y0 = ty * G::TH;
        __syncthreads();                      // table built / previous tile consumed
        load_tile<TW, TGY, HALO>(job, X0, Y0, tid);
        __syncthreads();
        tile_ring<R, MODE, TW, TGY, HALO>(job, X0, Y0, lx, gy);
    }
}

__global__ void reset_counter_kernel(int32_t *counter) { *counter = 0; }

// ---------------------------------------------------------------- configs

constexpr TileConfig kConfigs[] = {
    { 64, 16, 1024 },   // 0: 64 x 64 tile
    { 64, 8, 512 },     // 1: 64 x 32
    { 32, 8, 256 },     // 2: 32 x 32 (the reference's block footprint)
    { 64, 4, 256 },     // 3: 64 x 16
    { 128, 4, 512 },    // 4: 128 x 16
    { 128, 8, 1024 },   // 5: 128 x 32
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);
constexpr int kHalos[] = { 4, 8, 12, 16, 20 };

template <int TW, int TGY, int HALO>
int launch_one(const KArgs &a, int max_lut_rows, hipStream_t s)
{
    using G = Geo<TW, TGY, HALO>;
    const size_t lds = kTileOff + (size_t)G::TILE_BYTES + (size_t)max_lut_rows * 1024;
    auto kern = bf_exact_kernel<TW, TGY, HALO>;
    static int configured = 0;   // per instantiation
    if (!configured) {
        BF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        configured = 1;
    }
    int dev = 0, cus = 0, per_cu = 0;
    BF_HIP(hipGetDevice(&dev));
    BF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, G::NT, lds));
    if (per_cu < 1) {
        set_error("exact kernel does not fit on a CU (LDS " + std::to_string(lds) + " bytes)");
        return kInvalidData;
    }
    int grid = cus * per_cu;
    if (!a.tile_list && !a.counter && a.ntiles < grid) grid = a.ntiles;
    if (a.tile_list && a.ntiles < grid) grid = a.ntiles;   // ntiles = max possible
    if (grid < 1) return kOk;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(G::NT), lds, s, a);
    BF_HIP(hipGetLastError());
    return kOk;
}

template <int TW, int TGY>
int launch_halo(int halo, const KArgs &a, int max_lut_rows, hipStream_t s)
{
    switch (halo) {
    case 4: return launch_one<TW, TGY, 4>(a, max_lut_rows, s);
    case 8: return launch_one<TW, TGY, 8>(a, max_lut_rows, s);
    case 12: return launch_one<TW, TGY, 12>(a, max_lut_rows, s);
    case 16: return launch_one<TW, TGY, 16>(a, max_lut_rows, s);
    case 20: return launch_one<TW, TGY, 20>(a, max_lut_rows, s);
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
int launch_ring_one(const DevJob &job, int ntiles, hipStream_t s)
{
    using G = Geo<TW, TGY, HALO>;
    const size_t lds = kTileOff + (size_t)G::TILE_BYTES + (size_t)(job.nrows + 1) * 1024;
    auto kern = bf_ring_kernel<R, MODE, TW, TGY, HALO>;
    static int configured = 0;
    if (!configured) {
        BF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        configured = 1;
    }
    int dev = 0, cus = 0, per_cu = 0;
    BF_HIP(hipGetDevice(&dev));
    BF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, G::NT, lds));
    if (per_cu < 1) {
        set_error("ring kernel does not fit on a CU");
        return kInvalidData;
    }
    const int grid = std::min(ntiles, cus * per_cu);
    if (grid < 1) return kOk;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(G::NT), lds, s, job, ntiles);
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
    case 15: return ref ? launch_ring_one<15, 0, 64, 16, 16>(job, ntiles, s)
                        : launch_ring_one<15, 1, 64, 16, 16>(job, ntiles, s);
    case 9: return ref ? launch_ring_one<9, 0, 64, 8, 12>(job, ntiles, s)
                       : launch_ring_one<9, 1, 64, 8, 12>(job, ntiles, s);
    case 7: return ref ? launch_ring_one<7, 0, 64, 8, 8>(job, ntiles, s)
                       : launch_ring_one<7, 1, 64, 8, 8>(job, ntiles, s);
    case 5: return ref ? launch_ring_one<5, 0, 64, 8, 8>(job, ntiles, s)
                       : launch_ring_one<5, 1, 64, 8, 8>(job, ntiles, s);
    case 3: return ref ? launch_ring_one<3, 0, 64, 8, 4>(job, ntiles, s)
                       : launch_ring_one<3, 1, 64, 8, 4>(job, ntiles, s);
    default: set_error("no unrolled kernel for this radius"); return kInvalidData;
    }
}

int launch_reset_counter(int *counter, hipStream_t s)
{
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, counter);
    BF_HIP(hipGetLastError());
    return kOk;
}

}  // namespace bfimpl
