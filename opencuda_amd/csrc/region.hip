// region.hip -- the stages around the bilateral filter in
// okano_3_0/SimpleRegionDetect.cu:287-312: LocalCluster before it and
// ConnectRegion after it, for gfx950.
//
// LocalCluster (okano_3_0/LocalCluster.cu:45-153): one lane per pixel runs
// the eight directional walks the reference spreads over eight threads and
// a shared array, with the same FP32 operations in the same order
// (-ffp-contract=off, IEEE division), so the u8 result is the reference's.
// The reference addresses every read as data + y * pitch + x from the ROI
// origin and its walks reach 2 (pntRange - 1) pixels from the centre while
// the guard band is pntRange + 2 wide; reads here use the same linear
// address, clamped into the allocation the caller declares.  A workgroup
// stages its 64 x 16 tile plus the walks' reach in LDS by that rule.
//
// ConnectRegion (okano_3_0/ConnectRegion.cu:789-1043): 8-connected
// components of pixels whose gray values lie in [minGray, maxGray], each
// labelled by its minimum raster index.  The reference gets there with
// shared-memory labelling per block and rounds of border merging; here each
// 64 x 64 tile is labelled in LDS, then the runs that touch across tile
// edges are joined by a lock-free union-find pass in global memory
// (atomicMin on roots, parents always smaller than children, so every loop
// ends), then path compression -- the result does not depend on the order
// the unions run in.  Areas, the area filter, the raster-order
// renumbering and the u8 output follow the reference, including its area
// launch covering only rows below 256 * ceil(floor((H + 4) / 32) / 8)
// (ConnectRegion.cu:45, :957-967) and never dropping a root whose counted
// area is 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "bf_internal.h"

namespace bfimpl {
namespace region {

// ------------------------------------------------------------ LocalCluster

struct LcArgs {
    const uint8_t *in;        // ROI origin
    int64_t pitch;
    int64_t lo, hi;           // readable byte offsets from `in`: [lo, hi)
    uint8_t *out;
    int64_t out_pitch;
    int W, H;
    int pnt_range, pnt_count;
    float gap, diffe;
    int pro_black, pro_white;
};

__device__ __forceinline__ uint32_t lc_read(const LcArgs &a, int y, int x)
{
    int64_t off = (int64_t)y * a.pitch + x;
    off = off < a.lo ? a.lo : (off >= a.hi ? a.hi - 1 : off);
    return a.in[off];
}

__device__ __forceinline__ uint8_t lc_sat_u8(float q)
{
    if (!(q > 0.0f)) return 0;
    if (q >= 255.0f) return 255;
    return (uint8_t)(int)q;
}

// The filter of one pixel (x, y) whose value is cv; rd(dx, dy) returns the
// byte at (x + dx, y + dy) by the linear-address rule of lc_read.
template <typename Read>
__device__ __forceinline__ uint8_t lc_pixel(const LcArgs &a, int x, int y, uint32_t cv, Read rd)
{
    const int sx = a.pnt_range + 2, ex = a.W - a.pnt_range - 2;
    const int sy = a.pnt_range + 2, ey = a.H - a.pnt_range - 2;
    if (x < sx || x > ex || y < sy || y > ey || (int)cv <= a.pro_black || (int)cv >= a.pro_white)
        return (uint8_t)cv;
    // directions {dx, dy} of LocalCluster.cu:14-19
    const int ddx_t[8] = { 1, 1, 0, -1, -1, -1, 0, 1 };
    const int ddy_t[8] = { 0, -1, -1, -1, 0, 1, 1, 1 };
    const float fcv = (float)cv;
    float temp[8];
#pragma unroll
    for (int z = 0; z < 8; z++) {
        const int ddx = ddx_t[z], ddy = ddy_t[z];
        uint32_t cuv = rd(ddx, ddy);
        uint32_t pre1 = rd(-ddx, -ddy);
        uint32_t pre2 = cv;
        uint32_t latt1 = rd(2 * ddx, 2 * ddy);
        float sum = fcv;
        int i, k = 0;
        for (i = 1; i < a.pnt_range; i++) {
            if (fabsf(fcv - (float)cuv) > a.gap) break;
            k += 2;
            const uint32_t latt2 = rd(k * ddx, k * ddy);
            const float side1 = (float)(pre1 + pre2);
            const float side2 = (float)(latt1 + latt2);
            if (fabsf(side1 - side2) > a.diffe) break;
            pre1 = pre2;
            pre2 = cuv;
            cuv = latt1;
            latt1 = latt2;
            sum += (float)cuv;
        }
        temp[z] = sum / (float)i;
    }
    int mark = 0;
    for (int j = 8 - a.pnt_count; j > 0; j--) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (fabsf(temp[i] - fcv) >= fabsf(temp[mark] - fcv)) mark = i;
#pragma unroll
        for (int i = 0; i < 8; i++) temp[i] = i == mark ? fcv : temp[i];
    }
    float sumag = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; j++) sumag += temp[j];
    sumag -= (float)((8 - a.pnt_count) * (int)cv);
    const float q = sumag / (float)a.pnt_count;
    return q > 255.0f ? 255 : lc_sat_u8(q);
}

// Every read straight from global memory (any pntRange).
__global__ void __launch_bounds__(256) local_cluster_kernel(LcArgs a)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.W || y >= a.H) return;
    a.out[(int64_t)y * a.out_pitch + x] =
        lc_pixel(a, x, y, lc_read(a, y, x),
                 [&](int dx, int dy) { return lc_read(a, y + dy, x + dx); });
}

// A 64 x 16 tile with a halo of `halo` = 2 (pntRange - 1) pixels (the walks'
// reach) staged in LDS through lc_read, so every read the walks make sees
// the byte the global-read kernel would.  Each thread filters 4 pixels of a
// column (rows ty, ty + 4, ty + 8, ty + 12).
constexpr int kLcTw = 64, kLcTh = 16;
constexpr int64_t kLcMaxLds = 48 * 1024;   // larger halos read global memory

__global__ void __launch_bounds__(256) local_cluster_lds_kernel(LcArgs a, int halo)
{
    extern __shared__ uint8_t s_img[];
    const int sw = kLcTw + 2 * halo, sh = kLcTh + 2 * halo;
    const int x0 = blockIdx.x * kLcTw, y0 = blockIdx.y * kLcTh;
    // row r of the tile is the linear range from (y0 - halo + r, x0 - halo),
    // clamped byte by byte into the allocation as in lc_read
    for (int r = threadIdx.x >> 6; r < sh; r += 4) {
        const int64_t row = (int64_t)(y0 - halo + r) * a.pitch + (x0 - halo);
        for (int c = threadIdx.x & 63; c < sw; c += 64) {
            int64_t off = row + c;
            off = off < a.lo ? a.lo : (off >= a.hi ? a.hi - 1 : off);
            s_img[r * sw + c] = a.in[off];
        }
    }
    __syncthreads();
    const int x = x0 + (threadIdx.x & 63);
    if (x >= a.W) return;
#pragma unroll 1
    for (int k = 0; k < kLcTh / 4; k++) {
        const int y = y0 + (threadIdx.x >> 6) + 4 * k;
        if (y >= a.H) break;
        const int base = (y - y0 + halo) * sw + (x - x0 + halo);
        a.out[(int64_t)y * a.out_pitch + x] =
            lc_pixel(a, x, y, s_img[base],
                     [&](int dx, int dy) { return (uint32_t)s_img[base + dy * sw + dx]; });
    }
}

// ----------------------------------------------------------- ConnectRegion

__device__ __forceinline__ int cr_load(const int *L, int i)
{
    return __hip_atomic_load(L + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x: parents are never larger than their children, so the walk ends.
__device__ __forceinline__ int cr_find(const int *L, int x)
{
    int p = cr_load(L, x);
    while (p != x) {
        x = p;
        p = cr_load(L, x);
    }
    return x;
}

// Root of x with path halving: each visited node is re-pointed to its
// grandparent, an ancestor it keeps (trees only ever merge), so parents stay
// smaller than children.  A hook another lane makes meanwhile on a node
// this walk re-points could be overwritten, but such a node was no longer a
// root, so that lane's union loop does not stop there and merges the trees
// through the node's former parent.
__device__ __forceinline__ int cr_find_halving(int *L, int x)
{
    while (true) {
        const int p = cr_load(L, x);
        if (p == x) return x;
        const int gp = cr_load(L, p);
        if (gp == p) return p;
        __hip_atomic_store(L + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

// Lock-free union (the larger root is hooked under the smaller one with
// atomicMin; a failed hook means that root was hooked meanwhile, and the
// loop continues from its new parent, a strictly smaller index).
__device__ void cr_unite(int *L, int a, int b)
{
    while (true) {
        a = cr_find_halving(L, a);
        b = cr_find_halving(L, b);
        if (a == b) return;
        if (a < b) {
            const int old = atomicMin(L + b, a);
            if (old == b) return;
            b = old;
        } else {
            const int old = atomicMin(L + a, b);
            if (old == a) return;
            a = old;
        }
    }
}

struct CrArgs {
    const uint8_t *in;
    int64_t pitch;
    uint8_t *out;
    int64_t out_pitch;
    int W, H;
    int min_gray, max_gray;
    int min_area, max_area;
    int covered;              // rows the area count covers
    int *label, *area, *seq, *rowcount, *rowoff;
};

__device__ __forceinline__ bool cr_fg(const CrArgs &a, uint32_t v)
{
    return (int)v >= a.min_gray && (int)v <= a.max_gray;
}

// Areas start at zero (every label is written by the tile pass).
__global__ void cr_init_kernel(CrArgs a)
{
    const int64_t n = (int64_t)a.W * a.H;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a.area[i] = 0;
}

// Tile pass: each 64 x 64 tile labels its own pixels in LDS with the same
// lock-free union (LDS atomics), then writes every pixel's label as the
// global index of its tile-local root.  Row-major order inside a tile is
// the global order restricted to it, so a local root is the component's
// smallest global index within the tile, never larger than its children.
// (64 x 64 measured fastest on MI355X, 4096^2 chain image: 64 x 16 2.88 ms,
// 64 x 32 2.69, 64 x 64 2.48, 64 x 128 2.98, 128 x 64 3.47 -- taller tiles
// halve the border unions and the depth of the global trees, while wider
// ones lengthen the LDS union chains.)
constexpr int kCrTw = 64, kCrTh = 64;

__device__ __forceinline__ int lds_load(const int *L, int i)
{
    return __hip_atomic_load(L + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int l_find(const int *L, int x)
{
    int p = lds_load(L, x);
    while (p != x) {
        x = p;
        p = lds_load(L, x);
    }
    return x;
}

__device__ void l_unite(int *L, int a, int b)
{
    while (true) {
        a = l_find(L, a);
        b = l_find(L, b);
        if (a == b) return;
        if (a < b) {
            const int old = atomicMin(L + b, a);
            if (old == b) return;
            b = old;
        } else {
            const int old = atomicMin(L + a, b);
            if (old == a) return;
            a = old;
        }
    }
}

constexpr int kCrLocalThreads = 256;    // 512 and 1024 measured no faster

// A tile row is one wave (kCrTw = 64): its foreground runs are linked to
// their first pixel with a ballot (no atomics), and a run is united with
// the row above once per run it touches there: through N at the first of
// its pixels under each above segment, through NW / NE at its ends when N
// is background there (a diagonal neighbour next to a foreground N pixel is
// in that pixel's run already).
static_assert(kCrTw == 64, "one wave per tile row");

__global__ void __launch_bounds__(kCrLocalThreads) cr_local_kernel(CrArgs a)
{
    constexpr int kWaves = kCrLocalThreads / 64;
    __shared__ int loc[kCrTw * kCrTh];
    __shared__ uint64_t fgrow[kCrTh];            // foreground bits of each tile row
    const int x0 = blockIdx.x * kCrTw, y0 = blockIdx.y * kCrTh;
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int gx = x0 + lane;
    const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;   // lanes 0..lane
    for (int ly = wave; ly < kCrTh; ly += kWaves) {
        const int gy = y0 + ly;
        const bool f = gx < a.W && gy < a.H && cr_fg(a, a.in[(int64_t)gy * a.pitch + gx]);
        const uint64_t m = __ballot(f);
        const uint64_t starts = m & ~(m << 1);   // foreground with background on the left
        const int s = f ? 63 - __clzll((long long)(starts & upto)) : lane;
        loc[ly * kCrTw + lane] = ly * kCrTw + s;
        if (lane == 0) fgrow[ly] = m;
    }
    __syncthreads();
    for (int ly = 1 + wave; ly < kCrTh; ly += kWaves) {
        const uint64_t m = fgrow[ly], u = fgrow[ly - 1];
        if (!((m >> lane) & 1)) continue;
        const int i = ly * kCrTw + lane;
        const bool start = lane == 0 || !((m >> (lane - 1)) & 1);
        const bool end = lane == 63 || !((m >> (lane + 1)) & 1);
        const bool n = (u >> lane) & 1;
        const bool nw = lane > 0 && ((u >> (lane - 1)) & 1);
        const bool ne = lane < 63 && ((u >> (lane + 1)) & 1);
        if (n && (start || !nw)) l_unite(loc, i, i - kCrTw);
        if (!n && start && nw) l_unite(loc, i, i - kCrTw - 1);
        if (!n && end && ne) l_unite(loc, i, i - kCrTw + 1);
    }
    __syncthreads();
    for (int ly = wave; ly < kCrTh; ly += kWaves) {
        const int gy = y0 + ly;
        if (gx >= a.W || gy >= a.H) continue;
        const int root = l_find(loc, ly * kCrTw + lane);
        a.label[gy * a.W + gx] = (y0 + root / kCrTw) * a.W + x0 + (root & (kCrTw - 1));
    }
}

// Border pass: the pairs of 8-neighbours in different tiles.  One wave per
// tile edge: the top edge (64 pixels of the tile's first row) against the
// row above, the left edge (64 pixels of its first column) against the
// column to the left, each line extended by one pixel past both ends (the
// corner diagonals).  Pixels along an edge are joined where they touch
// inside their own tile, so the run rule of the tile pass applies with the
// edge as the line: one union per pair of touching runs.  A right edge is
// the next tile's left edge (its NE pairs are that edge's SW pairs).
__global__ void __launch_bounds__(256) cr_union_kernel(CrArgs a, int tiles_x, int ntiles)
{
    const int lane = (int)threadIdx.x & 63;
    const int e = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);   // 2 * tile + (0 top, 1 left)
    const int tile = e >> 1;
    if (tile >= ntiles) return;                                  // whole waves
    const int x0 = tile % tiles_x * kCrTw, y0 = tile / tiles_x * kCrTh;
    const bool top = (e & 1) == 0;
    if (top ? y0 == 0 : x0 == 0) return;
    // this pixel (x, y) and its neighbour across the edge (xo, yo); the
    // line's step from one lane to the next is (dx, dy)
    const int x = top ? x0 + lane : x0, y = top ? y0 : y0 + lane;
    const int dx = top ? 1 : 0, dy = top ? 0 : 1;
    const int xo = top ? x : x - 1, yo = top ? y - 1 : y;
    auto fg_at = [&](int px, int py) {
        return px >= 0 && px < a.W && py >= 0 && py < a.H &&
               cr_fg(a, a.in[(int64_t)py * a.pitch + px]);
    };
    const bool f = fg_at(x, y), o = fg_at(xo, yo);
    const uint64_t fm = __ballot(f), om = __ballot(o);
    // the other line's pixels before and after this one (past the ends: read)
    const bool ob = lane > 0 ? ((om >> (lane - 1)) & 1) : fg_at(xo - dx, yo - dy);
    const bool oa = lane < 63 ? ((om >> (lane + 1)) & 1) : fg_at(xo + dx, yo + dy);
    if (!f) return;
    const bool start = lane == 0 || !((fm >> (lane - 1)) & 1);
    const bool end = lane == 63 || !((fm >> (lane + 1)) & 1);
    const int idx = y * a.W + x, io = yo * a.W + xo, step = dy * a.W + dx;
    if (o && (start || !ob)) cr_unite(a.label, idx, io);
    if (!o && start && ob) cr_unite(a.label, idx, io - step);
    if (!o && end && oa) cr_unite(a.label, idx, io + step);
}

// Path compression and areas in one pass: one wave per row walks it in
// 64-pixel chunks, replaces every label by its root (the unions are all
// done, so a root is the same whichever lane finds it; a walk through a
// label another wave has already replaced reaches the same root), and for
// the rows below a.covered adds the areas per root.  Inside a chunk the
// lane that starts a run of equal roots adds the run; the run that reaches
// the chunk's end is carried, with a wave-uniform label and count, into the
// next chunk.  A row's first
// long run (>= 32 pixels) names its "hot" label, whose later runs are
// summed in the wave and added once at the row's end: in a large region
// broken by many small ones (here 1.26M runs on a 4096^2 image, half of them
// one region's) every run would otherwise be an atomic on the same
// counter, and those serialise.
__global__ void __launch_bounds__(256) cr_area_kernel(CrArgs a)
{
    const int lane = threadIdx.x & 63;
    const int y = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (y >= a.H) return;                           // whole waves
    int *row = a.label + (int64_t)y * a.W;
    const bool counted = y < a.covered;             // wave-uniform
    int cur = -1, cnt = 0;                          // carried run
    int hot = -1, hot_cnt = 0;                      // the row's hot label and its sum
    // a finished run of `lab` (wave-uniform): to the hot sum or one atomic
    auto flush = [&](int lab, int c) {
        if (lab < 0 || c <= 0) return;
        if (lab == hot) {
            hot_cnt += c;
        } else if (hot < 0 && c >= 32) {
            hot = lab;
            hot_cnt = c;
        } else if (lane == 0) {
            atomicAdd(a.area + lab, c);
        }
    };
    for (int x0 = 0; x0 < a.W; x0 += 64) {
        const int n = min(64, a.W - x0);            // pixels in this chunk
        // path compression: the root, written back for the later passes
        const int l = lane < n ? cr_find(a.label, y * a.W + x0 + lane) : -1;
        if (lane < n)
            __hip_atomic_store(row + x0 + lane, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!counted) continue;
        int prev = __shfl_up(l, 1);
        if (lane == 0) prev = cur;
        const bool start = lane < n && l != prev;
        const uint64_t starts = __ballot(start);
        // runs that start and end inside the chunk: the hot label's are
        // summed across the wave, the others added by their start lane
        int len = 0;
        if (start) {
            const uint64_t later = lane == 63 ? 0ull : (starts >> (lane + 1)) << (lane + 1);
            if (later) len = __ffsll((unsigned long long)later) - 1 - lane;
        }
        const bool to_hot = len > 0 && l == hot;
        if (len > 0 && !to_hot) atomicAdd(a.area + l, len);
        if (__ballot(to_hot)) {
            int sum = to_hot ? len : 0;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
            hot_cnt += sum;
        }
        const int first = starts ? __ffsll((unsigned long long)starts) - 1 : n;
        cnt += first;                               // the carried run goes on to `first`
        if (starts) {
            flush(cur, cnt);
            const int last = 63 - __clzll((long long)starts);
            cur = __shfl(l, last);
            cnt = n - last;
        }
    }
    if (!counted) return;
    flush(cur, cnt);
    if (lane == 0 && hot >= 0) atomicAdd(a.area + hot, hot_cnt);
}

__device__ __forceinline__ bool cr_kept_root(const CrArgs &a, int i)
{
    if (a.label[i] != i) return false;
    const int ar = a.area[i];
    return !(ar != 0 && (ar < a.min_area || ar > a.max_area));
}

// One wave per row: kept roots in the row.
__global__ void __launch_bounds__(64) cr_rowcount_kernel(CrArgs a)
{
    const int y = blockIdx.x;
    int count = 0;
    for (int x0 = 0; x0 < a.W; x0 += 64) {
        const int x = x0 + (int)threadIdx.x;
        const bool k = x < a.W && cr_kept_root(a, y * a.W + x);
        count += __popcll(__ballot(k));
    }
    if (threadIdx.x == 0) a.rowcount[y] = count;
}

// Exclusive scan of the row counts (one workgroup).
__global__ void __launch_bounds__(1024) cr_scan_kernel(CrArgs a)
{
    __shared__ int part[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < a.H; base += 1024) {
        const int i = base + (int)threadIdx.x;
        const int v = i < a.H ? a.rowcount[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < a.H) a.rowoff[i] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
}

// One wave per row: raster-order numbers of the kept roots (-1 for the
// other roots and for numbers >= 256).
__global__ void __launch_bounds__(64) cr_number_kernel(CrArgs a)
{
    const int y = blockIdx.x;
    const uint64_t below = (1ull << threadIdx.x) - 1ull;
    int next = a.rowoff[y];
    for (int x0 = 0; x0 < a.W; x0 += 64) {
        const int x = x0 + (int)threadIdx.x;
        const int i = y * a.W + x;
        const bool in = x < a.W;
        const bool root = in && a.label[i] == i;
        const bool k = in && cr_kept_root(a, i);
        const uint64_t m = __ballot(k);
        if (root) {
            const int s = next + __popcll(m & below);
            a.seq[i] = (k && s < 256) ? s : -1;
        }
        next += __popcll(m);
    }
}

__global__ void __launch_bounds__(256) cr_output_kernel(CrArgs a)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= a.W || y >= a.H) return;
    const int s = a.seq[a.label[y * a.W + x]];
    a.out[(int64_t)y * a.out_pitch + x] = (uint8_t)(s + 1);
}

}  // namespace region
}  // namespace bfimpl

using namespace bfimpl;
using namespace bfimpl::region;

namespace {

bool view_ok(const bf_image_view *v)
{
    return v && v->data && v->width > 0 && v->height > 0 && v->pitch >= v->width;
}

}  // namespace

extern "C" {

int bf_local_cluster(const bf_image_view *in, int64_t alloc_lo, int64_t alloc_hi,
                     const bf_image_view *out, const bf_local_cluster_params *p, void *stream)
{
    if (!in || !out || !p || !in->data || !out->data) {
        set_error("null argument");
        return kNullPointer;
    }
    if (!view_ok(in) || !view_ok(out)) {
        set_error("empty or malformed view");
        return kInvalidData;
    }
    if (alloc_lo == 0 && alloc_hi == 0) {          // the ROI itself
        alloc_hi = (int64_t)(in->height - 1) * in->pitch + in->width;
    }
    if (alloc_lo > 0 || alloc_hi < (int64_t)(in->height - 1) * in->pitch + in->width) {
        set_error("declared allocation does not contain the ROI");
        return kInvalidData;
    }
    LcArgs a{};
    a.in = in->data;
    a.pitch = in->pitch;
    a.lo = alloc_lo;
    a.hi = alloc_hi;
    a.out = out->data;
    a.out_pitch = out->pitch;
    a.W = std::min(in->width, out->width);
    a.H = std::min(in->height, out->height);
    a.pnt_range = p->pnt_range;
    a.pnt_count = p->pnt_count;
    a.gap = (float)p->gap_thred;
    a.diffe = (float)p->diffe_thred;
    a.pro_black = p->pro_black;
    a.pro_white = p->pro_white;
    // the walks reach 2 (pntRange - 1) pixels from the centre, and 2 at least
    const int64_t halo = std::max<int64_t>(2, 2 * ((int64_t)a.pnt_range - 1));
    const int64_t lds = (kLcTw + 2 * halo) * (kLcTh + 2 * halo);
    if (lds <= kLcMaxLds)
        hipLaunchKernelGGL(local_cluster_lds_kernel, dim3((a.W + kLcTw - 1) / kLcTw, (a.H + kLcTh - 1) / kLcTh),
                           dim3(256), (size_t)lds, (hipStream_t)stream, a, (int)halo);
    else
        hipLaunchKernelGGL(local_cluster_kernel, dim3((a.W + 63) / 64, (a.H + 3) / 4), dim3(256), 0,
                           (hipStream_t)stream, a);
    BF_HIP(hipGetLastError());
    return kOk;
}

int bf_connect_region(const bf_image_view *in, const bf_image_view *out, int32_t min_area,
                      int32_t max_area, int32_t min_gray, int32_t max_gray, void *stream)
{
    if (!in || !out || !in->data || !out->data) {
        set_error("null argument");
        return kNullPointer;
    }
    if (!view_ok(in) || !view_ok(out)) {
        set_error("empty or malformed view");
        return kInvalidData;
    }
    const int W = std::min(in->width, out->width), H = std::min(in->height, out->height);
    if ((int64_t)W * H > 0x7FFFFFFF) {
        set_error("image too large for 32-bit labels");
        return kInvalidData;
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)W * H;
    const size_t bytes = (3 * n + 2 * (size_t)H) * sizeof(int);
    int *work = nullptr;
    BF_HIP(hipMallocAsync(reinterpret_cast<void **>(&work), bytes, s));
    CrArgs a{};
    a.in = in->data;
    a.pitch = in->pitch;
    a.out = out->data;
    a.out_pitch = out->pitch;
    a.W = W;
    a.H = H;
    a.min_gray = min_gray;
    a.max_gray = max_gray;
    a.min_area = min_area;
    a.max_area = max_area;
    a.covered = std::min(H, 256 * ((((H + 4) / 32) + 7) / 8));
    a.label = work;
    a.area = work + n;
    a.seq = work + 2 * n;
    a.rowcount = work + 3 * n;
    a.rowoff = work + 3 * n + H;
    const int flat = (int)std::min<size_t>((n + 255) / 256, 65536);
    const dim3 g2((W + 63) / 64, (H + 3) / 4);
    hipLaunchKernelGGL(cr_init_kernel, dim3(flat), dim3(256), 0, s, a);
    hipLaunchKernelGGL(cr_local_kernel, dim3((W + kCrTw - 1) / kCrTw, (H + kCrTh - 1) / kCrTh),
                       dim3(kCrLocalThreads), 0, s, a);
    {
        const int tiles_x = (W + kCrTw - 1) / kCrTw, tiles_y = (H + kCrTh - 1) / kCrTh;
        const int edges = 2 * tiles_x * tiles_y;
        hipLaunchKernelGGL(cr_union_kernel, dim3((edges + 3) / 4), dim3(256), 0, s, a, tiles_x,
                           tiles_x * tiles_y);
    }
    hipLaunchKernelGGL(cr_area_kernel, dim3((H + 3) / 4), dim3(256), 0, s, a);
    hipLaunchKernelGGL(cr_rowcount_kernel, dim3(H), dim3(64), 0, s, a);
    hipLaunchKernelGGL(cr_scan_kernel, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(cr_number_kernel, dim3(H), dim3(64), 0, s, a);
    hipLaunchKernelGGL(cr_output_kernel, g2, dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    hipError_t f = hipFreeAsync(work, s);
    if (e != hipSuccess) return hip_fail(e, "bf_connect_region");
    if (f != hipSuccess) return hip_fail(f, "bf_connect_region");
    return kOk;
}

}  // extern "C"
