// bf_tile.h -- device building blocks shared by the exact-filter kernels
// (bf_exact.hip: single image, tile lists, sample check; bf_ring_batch.hip:
// batches of canonical-template jobs).  Not a public header.
//
// LDS is addressed by byte offset through address-space-3 pointers.  The
// kernels that use these helpers declare no static LDS, so the dynamic
// allocation starts at LDS address 0 and a v_sad_u8 result is usable as a
// ds_read address unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "bf_internal.h"

namespace bfimpl {

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

typedef __attribute__((address_space(3))) uint8_t lds_byte;

__device__ __forceinline__ uint32_t lds_u8(uint32_t byte_addr)
{
    return *(const lds_byte *)(size_t)byte_addr;
}

// Stage the clamped input rectangle (rows ys .. ys+ROWS-1, columns xs ..
// xs+COLS-1, clamped to the input view) into LDS at `base`, row-major with
// COLS elements per row: byte-replicated u32 elements (the exact kernels'
// v_sad_u8 form) or, with BYTES, plain bytes.  NT threads (ids 0..NT-1)
// share the work.
template <int COLS, int ROWS, int NT, bool BYTES>
__device__ __forceinline__ void stage_rect(const DevJob &job, int xs, int ys, uint32_t base,
                                           int tid)
{
    constexpr uint32_t ESZ = BYTES ? 1u : 4u;
    const int wmax = job.in_w - 1, hmax = job.in_h - 1;
    auto put = [&](int e, uint32_t byte) {
        if (BYTES)
            *(lds_byte *)(size_t)(base + (uint32_t)e) = (uint8_t)byte;
        else
            lds_store_u32(base + (uint32_t)e * ESZ, byte * 0x01010101u);
    };
    if (xs >= 0 && xs + COLS <= job.in_w) {
        // Interior columns: aligned 16-byte loads covering [xs, xs+COLS).
        // Bytes outside that span are dropped; they lie in the same aligned
        // 16-byte block as a byte of the row, hence in mapped memory.
        constexpr int NCH = (COLS + 15) / 16 + 1;  // chunks per row
        for (int e = tid; e < ROWS * NCH; e += NT) {
            const int r = e / NCH, ch = e - r * NCH;
            const int gy = clampi(ys + r, 0, hmax);
            const uint8_t *row = job.in + (int64_t)gy * job.in_pitch;
            const uintptr_t a0 = reinterpret_cast<uintptr_t>(row + xs) & ~(uintptr_t)15;
            const uintptr_t ca = a0 + 16 * (uintptr_t)ch;
            const intptr_t first = (intptr_t)ca - (intptr_t)reinterpret_cast<uintptr_t>(row + xs);
            if (first >= COLS) continue;
            const uint4 v = *reinterpret_cast<const uint4 *>(ca);
            const uint32_t w[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
            for (int b = 0; b < 16; b++) {
                const int c = (int)first + b;
                if (c >= 0 && c < COLS) put(r * COLS + c, (w[b >> 2] >> (8 * (b & 3))) & 0xFFu);
            }
        }
    } else {
        for (int e = tid; e < ROWS * COLS; e += NT) {
            const int r = e / COLS, c = e - r * COLS;
            const int gy = clampi(ys + r, 0, hmax);
            const int gx = clampi(xs + c, 0, wmax);
            put(e, job.in[(int64_t)gy * job.in_pitch + gx]);
        }
    }
}

// Stage the clamped input tile (rows Y0-HALO .., cols X0-HALO ..) into LDS
// as byte-replicated u32.
template <int TW, int TGY, int HALO>
__device__ __forceinline__ void load_tile(const DevJob &job, int X0, int Y0, int tid)
{
    using G = Geo<TW, TGY, HALO>;
    stage_rect<G::TCOLS, G::TROWS, G::NT, false>(job, X0 - HALO, Y0 - HALO, kTileOff, tid);
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
    if constexpr (NT % 256 == 0) {
        // Each thread keeps one column d (its range value loaded once); the
        // row a wave writes is the same for all its lanes, so its spatial
        // weight is a scalar load.  (A loop over flat indices waited for
        // two vector loads per entry, ~30 round trips before the first
        // tile of a launch or of a job with other tables.)
        const int d = tid & 255;
        const float rd = job.range[d];
        const int r0 = __builtin_amdgcn_readfirstlane(tid >> 8);
        for (int row = r0; row <= job.nrows; row += NT / 256) {
            const float w = row < job.nrows ? job.rowg[row] * rd : 0.0f;
            lds_store_u32(lut_base + (uint32_t)(row * 256 + d) * 4, __float_as_uint(w));
        }
    } else {
        const int n = (job.nrows + 1) * 256;
        for (int e = tid; e < n; e += NT) {
            const int row = e >> 8, d = e & 255;
            float w = 0.0f;
            if (row < job.nrows) w = job.rowg[row] * job.range[d];
            lds_store_u32(lut_base + (uint32_t)e * 4, __float_as_uint(w));
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
    int side;      // 0: the ring's top/bottom rows; 1: right column; 2: left column
    bool first;    // first tap of its side column in this ring (j = -i + 1)
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
                    const int side = k < 2 ? 0 : k - 1;
                    t[n++] = RingTap{ dy, dx, row, side, side != 0 && j == -i + 1 };
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

// Taps scheduled together between fences in the unrolled kernels.
// (Measured on MI355X, c3 image, alternating builds: REF is flat from 4 to
// 8 taps and slower at 2; STD gains 8-14% going from 2 taps to 12 at r = 15
// and 24 at r = 3..9, and at r = 15 runs 8.90 ms at 20 taps against 9.14
// at 12, 9.02 at 16, 9.23 at 18, 9.20 at 22 and 9.13 at 24 or 32.)
template <int MODE, int R>
constexpr int ring_group() { return MODE == kModeRef ? 4 : (R >= 13 ? 20 : 24); }

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
    uint32_t tbase = lane + (uint32_t)MINB;         // >= 0: the halo covers every tap
    // Kept opaque: folded into `lane`, the taps above and left of the pixel
    // would need negative offsets, which ds_read's unsigned offset field
    // cannot hold (each such fetch then costs a v_add_u32).
    asm volatile("" : "+v"(tbase));
    uint32_t wp[2][4] = {};                         // STD: side-column windows (right, left)
    float wf[2][4] = {};
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
            // Own window per row: row k reads tile row k + dy.  Along a side
            // column of a ring (dx = +-i, dy = j = -i+1 .. i, taps four apart
            // in ring order) row k at dy = j reads the pixel row k+1 read at
            // dy = j - 1: a window of the four pixels rows j .. j+3 of that
            // column (bytes and their float values) slides down one row per
            // tap, so such a tap fetches and converts one pixel, not four.
            uint32_t p0, p1, p2, p3;
            float q0, q1, q2, q3;
            if constexpr (tp.side != 0) {
                constexpr int sd = tp.side - 1;
                if constexpr (tp.first) {
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        wp[sd][m] = lds_u32(tbase + TAPB + m * ROWB);
                        wf[sd][m] = (float)(wp[sd][m] & 0xFFu);
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < 3; m++) {
                        wp[sd][m] = wp[sd][m + 1];
                        wf[sd][m] = wf[sd][m + 1];
                    }
                    wp[sd][3] = lds_u32(tbase + TAPB + 3 * ROWB);
                    wf[sd][3] = (float)(wp[sd][3] & 0xFFu);
                }
                p0 = wp[sd][0]; p1 = wp[sd][1]; p2 = wp[sd][2]; p3 = wp[sd][3];
                q0 = wf[sd][0]; q1 = wf[sd][1]; q2 = wf[sd][2]; q3 = wf[sd][3];
            } else {
                p0 = lds_u32(tbase + TAPB); p1 = lds_u32(tbase + TAPB + ROWB);
                p2 = lds_u32(tbase + TAPB + 2 * ROWB); p3 = lds_u32(tbase + TAPB + 3 * ROWB);
                q0 = (float)(p0 & 0xFFu); q1 = (float)(p1 & 0xFFu);
                q2 = (float)(p2 & 0xFFu); q3 = (float)(p3 & 0xFFu);
            }
            const float f0 = lds_f32(__builtin_amdgcn_sad_u8(p0, c0, WB) + WOFF);
            const float f1 = lds_f32(__builtin_amdgcn_sad_u8(p1, c1, WB) + WOFF);
            const float f2 = lds_f32(__builtin_amdgcn_sad_u8(p2, c2, WB) + WOFF);
            const float f3 = lds_f32(__builtin_amdgcn_sad_u8(p3, c3, WB) + WOFF);
            t0 = t0 + f0 * q0; s0 = s0 + f0;
            t1 = t1 + f1 * q1; s1 = s1 + f1;
            t2 = t2 + f2 * q2; s2 = s2 + f2;
            t3 = t3 + f3 * q3; s3 = s3 + f3;
        }
        // Keep the schedule local (a group of taps in flight), as the loop
        // version does: without a fence the compiler issues every gather of
        // the window first and spills them.  The empty asm reads and writes
        // the accumulators and clobbers memory, so neither the adds nor the
        // LDS reads move across it; it emits no instruction.
        if constexpr (decltype(ci)::value % ring_group<MODE, R>() == ring_group<MODE, R>() - 1)
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

// hipFuncSetAttribute once per kernel (thread-safe static initialisation):
// allow the dynamic LDS of a kernel to reach the CU's 160 KB.
template <auto K>
hipError_t allow_full_lds()
{
    static const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(K),
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    160 * 1024);
    return e;
}

// Workgroups of `kern` (NT threads, lds bytes) that run at once on the
// device; 0 when one does not fit.
inline int resident_groups(const void *kern, int nt, size_t lds)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nt, lds) != hipSuccess)
        return -1;
    return cus * per_cu;
}

}  // namespace bfimpl
