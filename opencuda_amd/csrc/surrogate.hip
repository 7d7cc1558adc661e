// surrogate.hip -- learned BF16 surrogate of the exact filter (STD or REF
// target), trained and run with hand-written MFMA kernels on gfx950.
//
// No reference counterpart.  Network: six 3x3 convolutions with zero
// padding, 32 channels, ReLU after layers 1-5, inputs (image/255,
// sigma_range/100), output = image/255 + conv6(...).  Weights and
// activations are BF16, every product accumulates in FP32 on
// v_mfma_f32_16x16x32_bf16; master weights and Adam moments are FP32.
//
// Parameter layout (FP32, 37,889 values): for layer l = 1..6,
// W_l[t][ci][co] then b_l[co], with t = (dy+1)*3 + (dx+1) the tap and
// (CIN, COUT) = (2,32), (32,32) x4, (32,1).
//
// MFMA operand layout (16x16x32, checked on MI355X): lane l holds
//   A[i = l%16][k = 8*(l/16) .. +7], B[k = 8*(l/16) .. +7][j = l%16],
//   D[i = 4*(l/16) + r][j = l%16], r = 0..3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "bf_internal.h"

namespace bfimpl {
namespace sur {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int C = 32;          // hidden channels
constexpr int NLAYERS = 6;
constexpr int TX = 64;         // conv output tile: 64 x 4 pixels per workgroup
constexpr int TY = 4;
constexpr int PITCH = 40;      // staged NHWC pixel: 32 channels + 8 pad = 80 bytes
constexpr int MARGIN = 6;      // crop border excluded from the loss (receptive field)

constexpr int cin_of(int l) { return l == 0 ? 2 : C; }
constexpr int cout_of(int l) { return l == NLAYERS - 1 ? 1 : C; }
constexpr int wcount(int l) { return 9 * cin_of(l) * cout_of(l); }
constexpr int pcount(int l) { return wcount(l) + cout_of(l); }
constexpr int poff(int l) { return l == 0 ? 0 : poff(l - 1) + pcount(l - 1); }
constexpr int kParams = poff(NLAYERS);
static_assert(kParams == 37889, "parameter count");

// Packed BF16 weights: forward B operand FW_l[nt][16][KF_l] and
// data-gradient B operand BW_l[nt][16][KB_l] (K contiguous).
constexpr int kf_of(int l) { return l == 0 ? 32 : 9 * C; }           // K of the forward GEMM
constexpr int nt_of(int l) { return l == NLAYERS - 1 ? 1 : 2; }       // 16-wide co tiles
constexpr int kb_of(int l) { return l == NLAYERS - 1 ? 32 : 9 * C; }  // K of the dgrad GEMM
constexpr int fw_size(int l) { return nt_of(l) * 16 * kf_of(l); }
constexpr int bw_size(int l) { return 2 * 16 * kb_of(l); }

enum Kind { K_FIRST = 0, K_HIDDEN = 1, K_LAST = 2, K_DG_HIDDEN = 3, K_DG_LAST = 4 };

struct ConvArgs {
    const __bf16 *in;       // HIDDEN/LAST/DG_HIDDEN: NHWC, 32 ch; DG_LAST: g6 [N][H][W]
    const uint8_t *img;     // FIRST, LAST: u8 image(s), row pitch img_pitch, image stride img_stride
    int64_t img_pitch;
    int64_t img_stride;
    const float *sigma;     // FIRST: per-image sigma_range (device), or null:
    float sigma_const;      //        then sigma_const for every image
    const __bf16 *wpk;      // packed weights [nt][16][K]
    const float *bias;      // FIRST/HIDDEN/LAST
    const __bf16 *mask;     // DG_*: activation whose > 0 is the ReLU mask (NHWC 32)
    __bf16 *out;            // FIRST/HIDDEN/DG_*: NHWC 32
    float *pred;            // LAST: f32 [N][H][W] (may be null)
    uint8_t *out_u8;        // LAST: optional u8 output, pitch u8_pitch
    int64_t u8_pitch;
    int N, H, W;
};

__device__ __forceinline__ __bf16 to_bf(float v) { return (__bf16)v; }
__device__ __forceinline__ float from_bf(__bf16 v) { return (float)v; }

// Activation of a pre-activation sum (bias included): rounded to BF16, then
// the ReLU on the BF16 bits -- a BF16 value is > 0 exactly when its bits,
// read as int16, are > 0, so max(bits, 0) is the ReLU (-0 becomes +0).
__device__ __forceinline__ __bf16 relu_bf(float v)
{
    const short b = __builtin_bit_cast(short, to_bf(v));
    return __builtin_bit_cast(__bf16, (short)(b > 0 ? b : 0));
}

// Two activations packed in a dword (lo = first), with one packed max.
__device__ __forceinline__ uint32_t act_pack2(float lo, float hi)
{
    typedef short i16x2_t __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    // one v_cvt_pk_bf16_f32 for both (round to nearest even, as to_bf)
    const bf16x2_t b = __builtin_convertvector((f32x2_t){ lo, hi }, bf16x2_t);
    const i16x2_t v = __builtin_elementwise_max(__builtin_bit_cast(i16x2_t, b), (i16x2_t){ 0, 0 });
    return __builtin_bit_cast(uint32_t, v);
}

// Eight BF16 values from two runs of four (the last layer's channel order
// in a k-slot group: channels 4kq..4kq+3, then 16+4kq..16+4kq+3).
__device__ __forceinline__ bf16x8 cat4(const __bf16 *p0, const __bf16 *p1)
{
    typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
    const bf16x4_t a = *reinterpret_cast<const bf16x4_t *>(p0);
    const bf16x4_t b = *reinterpret_cast<const bf16x4_t *>(p1);
    return (bf16x8){ a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3] };
}

template <int KIND>
struct KTraits {
    static constexpr int KS = (KIND == K_FIRST || KIND == K_DG_LAST) ? 1 : 9;   // 32-wide k-steps
    static constexpr int NT = (KIND == K_LAST) ? 1 : 2;
    static constexpr int KTOT = KS * 32;
    static constexpr bool STAGE32 = (KIND == K_HIDDEN || KIND == K_LAST || KIND == K_DG_HIDDEN);
};

// Two BF16 gradients with the ReLU mask of the matching activations: a half
// is kept where that activation is > 0, else it becomes +0.
__device__ __forceinline__ uint32_t relu_keep(uint32_t v, uint32_t act)
{
    const float lo = __uint_as_float(act << 16), hi = __uint_as_float(act & 0xFFFF0000u);
    return v & ((lo > 0.0f ? 0x0000FFFFu : 0u) | (hi > 0.0f ? 0xFFFF0000u : 0u));
}

// Implicit-GEMM 3x3 convolution: out[p][co] = sum_{t,ci} in[p+off(t)][ci] * W[t][ci][co].
// Workgroup = 4 waves = 64 x 4 output pixels; wave w owns row y0 + w as 4
// M-tiles of 16 pixels.
template <int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) conv_kernel(ConvArgs a)
{
    using T = KTraits<KIND>;
    // Staged input tile; after the MFMAs it also holds the output tile (the
    // epilogue's [wave][64 pixels][PITCH] rows) for 16-byte global stores.
    __shared__ __align__(16) __bf16 s_in[T::STAGE32 ? (TY + 2) * (TX + 2) * PITCH
                                                   : (KIND == K_LAST ? 8 : TY * TX * PITCH)];
    __shared__ float s_first[(KIND == K_FIRST || KIND == K_DG_LAST) ? (TY + 2) * (TX + 2) : 1];
    __shared__ float s_last[KIND == K_LAST ? TY * TX : 1];   // the last layer's row sums

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int H = a.H, W = a.W;
    const int64_t npix = (int64_t)H * W;
    const int tiles_x = (W + TX - 1) / TX, tiles_y = (H + TY - 1) / TY;
    const int64_t ntiles = (int64_t)tiles_x * tiles_y * a.N;

    // ---- B operand (weights) into registers, once per persistent workgroup
    // (The last layer's k-slots of lane group kq hold channels 4kq..4kq+3 and
    // 16+4kq..16+4kq+3: the order in which the fused kernel's layer-4 lanes
    // hold their outputs.)
    bf16x8 bw[T::KS][T::NT];
#pragma unroll
    for (int s = 0; s < T::KS; s++)
#pragma unroll
        for (int nt = 0; nt < T::NT; nt++) {
            const __bf16 *w = a.wpk + ((size_t)(nt * 16 + (lane & 15)) * T::KTOT + s * 32);
            bw[s][nt] = KIND == K_LAST ? cat4(w + 4 * (lane >> 4), w + 16 + 4 * (lane >> 4))
                                       : *reinterpret_cast<const bf16x8 *>(w + 8 * (lane >> 4));
        }

    // 32-channel inputs: the tile's (TY+2) x (TX+2) pixels as 16-byte
    // chunks, SE per thread, loaded into registers one tile ahead so the
    // loads are in flight during the previous tile's MFMAs (a loop that
    // stored each chunk as it arrived waited for every load in turn).
    constexpr int SN = (TY + 2) * (TX + 2) * 4;
    constexpr int SE = T::STAGE32 ? (SN + 255) / 256 : 1;
    uint4 pre[SE];
    auto load_stage = [&](int64_t t) {
        const int tn = (int)(t / ((int64_t)tiles_x * tiles_y));
        const int tr = (int)(t - (int64_t)tn * tiles_x * tiles_y);
        const int ty0 = (tr / tiles_x) * TY, tx0 = (tr % tiles_x) * TX;
        const __bf16 *src = a.in + (size_t)tn * npix * C;
#pragma unroll
        for (int k = 0; k < SE; k++) {
            const int e = tid + 256 * k;
            const int chunk = e & 3, pix = e >> 2;
            const int r = pix / (TX + 2), c = pix - r * (TX + 2);
            const int gy = ty0 - 1 + r, gx = tx0 - 1 + c;
            pre[k] = make_uint4(0, 0, 0, 0);
            if (e < SN && gy >= 0 && gy < H && gx >= 0 && gx < W)
                pre[k] = *reinterpret_cast<const uint4 *>(src + ((size_t)gy * W + gx) * C + chunk * 8);
        }
    };
    if (T::STAGE32 && (int64_t)blockIdx.x < ntiles) load_stage(blockIdx.x);

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int n = (int)(tile / ((int64_t)tiles_x * tiles_y));
    const int trem = (int)(tile - (int64_t)n * tiles_x * tiles_y);
    const int y0 = (trem / tiles_x) * TY, x0 = (trem % tiles_x) * TX;
    __syncthreads();   // the previous tile's reads of the staging buffer are done

    // ---- stage the input tile (rows y0-1 .. y0+4, cols x0-1 .. x0+64), zero outside
    if (T::STAGE32) {
#pragma unroll
        for (int k = 0; k < SE; k++) {
            const int e = tid + 256 * k;
            if (e < SN)
                *reinterpret_cast<uint4 *>(&s_in[(e >> 2) * PITCH + (e & 3) * 8]) = pre[k];
        }
    } else if (KIND == K_FIRST) {
        const uint8_t *im = a.img + (size_t)n * a.img_stride;
        for (int e = tid; e < (TY + 2) * (TX + 2); e += 256) {
            const int r = e / (TX + 2), c = e - r * (TX + 2);
            const int gy = y0 - 1 + r, gx = x0 - 1 + c;
            float v = -1.0f;   // marks padding
            if (gy >= 0 && gy < H && gx >= 0 && gx < W) v = (float)im[(int64_t)gy * a.img_pitch + gx];
            s_first[e] = v;
        }
    } else {   // K_DG_LAST: 1-channel gradient g6
        const __bf16 *src = a.in + (size_t)n * npix;
        for (int e = tid; e < (TY + 2) * (TX + 2); e += 256) {
            const int r = e / (TX + 2), c = e - r * (TX + 2);
            const int gy = y0 - 1 + r, gx = x0 - 1 + c;
            float v = 0.0f;
            if (gy >= 0 && gy < H && gx >= 0 && gx < W) v = from_bf(src[(size_t)gy * W + gx]);
            s_first[e] = v;
        }
    }
    __syncthreads();
    if (T::STAGE32 && tile + gridDim.x < ntiles) load_stage(tile + gridDim.x);
    // data-gradient kinds: the ReLU mask of this wave's output row, loaded
    // now so it arrives during the MFMAs (pixels 16 j + lane / 4, 16-byte
    // chunk lane % 4, as the epilogue stores them)
    constexpr bool DG = KIND == K_DG_HIDDEN || KIND == K_DG_LAST;
    uint4 mask[DG ? 4 : 1];
    if (DG) {
        const int my = y0 + wave;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int mx = x0 + 16 * j + (lane >> 2);
            mask[j] = make_uint4(0, 0, 0, 0);
            if (my < H && mx < W)
                mask[j] = *reinterpret_cast<const uint4 *>(
                    a.mask + ((size_t)n * npix + (size_t)my * W + mx) * C + (lane & 3) * 8);
        }
    }

    const int i16 = lane & 15, kq = lane >> 4;
    // The forward layers start from their bias (the accumulator's initial
    // value); the data-gradient kinds from zero.
    f32x4 acc[4][T::NT];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int nt = 0; nt < T::NT; nt++) {
            const float b = (KIND == K_FIRST || KIND == K_HIDDEN) ? a.bias[16 * nt + i16] : 0.0f;
            acc[m][nt] = (f32x4){ b, b, b, b };
        }

    if (KIND == K_LAST) {
        // Per tap s the product of the 32 channels with W6[s] (zero
        // accumulator), summed over the taps in order in FP32.
#pragma unroll
        for (int s = 0; s < 9; s++) {
            const int dy = s / 3 - 1, dx = s % 3 - 1;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int pix = (wave + 1 + dy) * (TX + 2) + (16 * m + i16 + 1 + dx);
                const bf16x8 av = cat4(&s_in[pix * PITCH + 4 * kq], &s_in[pix * PITCH + 16 + 4 * kq]);
                const f32x4 p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[s][0], (f32x4){ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
                acc[m][0] = s == 0 ? p : acc[m][0] + p;
            }
            // Finish each tap's sums before the next tap's loads: without
            // this fence the compiler issued all 36 operand loads and
            // products first (480 registers, one wave per SIMD).  The empty
            // asm takes the sums and clobbers memory; it emits nothing.
            asm volatile("" : "+v"(acc[0][0]), "+v"(acc[1][0]), "+v"(acc[2][0]), "+v"(acc[3][0])::"memory");
        }
    } else if (T::STAGE32) {
#pragma unroll
        for (int s = 0; s < 9; s++) {
            const int dy = s / 3 - 1, dx = s % 3 - 1;
            // Data-gradient convs run the flipped taps: the packed BW weights
            // already hold W[8 - t], so the input offset is the forward one.
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const int pix = (wave + 1 + dy) * (TX + 2) + (16 * m + i16 + 1 + dx);
                const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&s_in[pix * PITCH + 8 * kq]);
#pragma unroll
                for (int nt = 0; nt < T::NT; nt++)
                    acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[s][nt], acc[m][nt], 0, 0, 0);
            }
        }
    } else {
        const float sr = (KIND == K_FIRST) ? (a.sigma ? a.sigma[n] : a.sigma_const) / 100.0f : 0.0f;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            bf16x8 av;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = 8 * kq + q;
                float v = 0.0f;
                if (KIND == K_FIRST) {
                    const int t = k >> 1, ci = k & 1;
                    if (t < 9) {
                        const int dy = t / 3 - 1, dx = t % 3 - 1;
                        const float raw = s_first[(wave + 1 + dy) * (TX + 2) + 16 * m + i16 + 1 + dx];
                        if (raw >= 0.0f) v = ci == 0 ? raw / 255.0f : sr;
                    }
                } else {
                    if (k < 9) {
                        const int dy = k / 3 - 1, dx = k % 3 - 1;
                        v = s_first[(wave + 1 + dy) * (TX + 2) + 16 * m + i16 + 1 + dx];
                    }
                }
                av[q] = to_bf(v);
            }
#pragma unroll
            for (int nt = 0; nt < T::NT; nt++)
                acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[0][nt], acc[m][nt], 0, 0, 0);
        }
    }

    // ---- epilogue
    const int y = y0 + wave;
    if constexpr (KIND == K_LAST) {
        // Column 0 of the products (lanes i16 == 0) holds the row's 64 sums;
        // through LDS each lane takes one pixel, so the image loads and the
        // stores are one coalesced access per row instead of sixteen
        // four-lane ones, each waiting for the last.
        if (i16 == 0) {
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int r = 0; r < 4; r++) s_last[wave * TX + 16 * m + 4 * kq + r] = acc[m][0][r];
        }
        __syncthreads();
        const int x = x0 + lane;
        if (y < H && x < W) {
            const size_t p = (size_t)n * npix + (size_t)y * W + x;
            const uint8_t u = a.img[(size_t)n * a.img_stride + (int64_t)y * a.img_pitch + x];
            const float v = (float)u / 255.0f + (s_last[wave * TX + lane] + a.bias[0]);
            if (a.pred) a.pred[p] = v;
            if (a.out_u8) {
                const float q = rintf(v * 255.0f);
                a.out_u8[(int64_t)y * a.u8_pitch + x] =
                    (uint8_t)(q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q));
            }
        }
    } else {
        // NHWC-32 output: each wave writes its 64 x 32 tile to LDS (the
        // staged input is dead once every wave has done its MFMAs), then
        // stores it as 16 contiguous bytes per lane, the data-gradient kinds
        // applying the ReLU mask loaded before the MFMAs.
        __syncthreads();
        __bf16 *s_out = s_in + wave * TX * PITCH;
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int nt = 0; nt < T::NT; nt++) {
                    const int co = 16 * nt + i16;
                    const float v = acc[m][nt][r];
                    s_out[(16 * m + 4 * kq + r) * PITCH + co] =
                        (KIND == K_FIRST || KIND == K_HIDDEN) ? relu_bf(v) : to_bf(v);
                }
        __syncthreads();
        if (y < H) {
            const int chunk = lane & 3;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int px = 16 * j + (lane >> 2), x = x0 + px;
                if (x >= W) continue;
                const size_t p = (size_t)n * npix + (size_t)y * W + x;
                uint4 v = *reinterpret_cast<const uint4 *>(&s_out[px * PITCH + chunk * 8]);
                if (DG) {
                    v.x = relu_keep(v.x, mask[j].x);
                    v.y = relu_keep(v.y, mask[j].y);
                    v.z = relu_keep(v.z, mask[j].z);
                    v.w = relu_keep(v.w, mask[j].w);
                }
                *reinterpret_cast<uint4 *>(a.out + p * C + chunk * 8) = v;
            }
        }
    }
    }   // tile loop
}

// ------------------------------------------------------------ weight grads

struct WgradArgs {
    const __bf16 *act;       // HIDDEN/LAST: layer input activation NHWC 32
    const uint8_t *img;      // FIRST: u8 crops [N][H][W]
    const float *sigma;      // FIRST
    const __bf16 *dpre;      // FIRST/HIDDEN: NHWC 32 gradient; LAST: g6 [N][H][W]
    float *partial;          // [gridDim.x][pcount]
    int N, H, W;
};

// Weight gradients of the first layer (inputs: the image and sigma_range):
// dW[t][ci][co] = sum_p in[p + off(t)][ci] * d[p][co]; db[co] = sum_p d[p][co].
// GEMM with M = (t, ci) (18 rows, two M-tiles), N = co, K = pixels: a
// workgroup walks segments of 64 pixels of one row; the input rows are
// staged transposed ([ci][x]) in three copies shifted by dx = -1, 0, 1 so
// every operand is an aligned 16-byte LDS read.  Each thread's global loads
// of a segment (one image byte, one 16-byte gradient chunk) are issued one
// segment ahead, during the previous segment's MFMAs.  (The 32-channel
// inputs go through wgrad32_kernel.)
template <int KIND>
__global__ void __launch_bounds__(256) wgrad_kernel(WgradArgs a)
{
    static_assert(KIND == K_FIRST, "first layer only");
    constexpr int CIN = 2, COUT = C;
    constexpr int M = 9 * CIN;
    constexpr int MT = (M + 15) / 16;
    constexpr int NT = COUT / 16;
    constexpr int PSZ = 9 * CIN * COUT + COUT;
    static_assert(MT * NT == 4, "one (m, n) tile pair per wave");
    static_assert(3 * (TX + 2) <= 256 && TX * 4 == 256, "one staging element per thread");
    __shared__ __align__(16) __bf16 s_in[3][3][CIN][TX];
    __shared__ __align__(16) __bf16 s_d[NT * 16][TX];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i16 = lane & 15, kq = lane >> 4;
    const int segs_x = (a.W + TX - 1) / TX;
    const int64_t nseg = (int64_t)a.N * a.H * segs_x;
    const int64_t npix = (int64_t)a.H * a.W;

    // staging tasks: context pixel (row ir, column ic of the 3 x (TX + 2)
    // window) for threads < 3 (TX + 2); gradient chunk dch of pixel dxx
    const bool img_task = tid < 3 * (TX + 2);
    const int ir = tid / (TX + 2), ic = tid - ir * (TX + 2);
    const int dch = tid & 3, dxx = tid >> 2;
    float pimg = -1.0f, psr = 0.0f;     // pimg < 0: outside the crop
    uint4 pd = make_uint4(0, 0, 0, 0);
    auto load_seg = [&](int64_t seg) {
        const int n = (int)(seg / ((int64_t)a.H * segs_x));
        const int rem = (int)(seg - (int64_t)n * a.H * segs_x);
        const int y = rem / segs_x, x0 = (rem - y * segs_x) * TX;
        psr = a.sigma[n] / 100.0f;
        pimg = -1.0f;
        if (img_task) {
            const int gy = y - 1 + ir, gx = x0 - 1 + ic;
            if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
                pimg = (float)a.img[(size_t)n * npix + (size_t)gy * a.W + gx];
        }
        pd = make_uint4(0, 0, 0, 0);
        if (x0 + dxx < a.W)
            pd = *reinterpret_cast<const uint4 *>(a.dpre + ((size_t)n * npix + (size_t)y * a.W + x0 + dxx) * C + dch * 8);
    };

    f32x4 acc = (f32x4){ 0.f, 0.f, 0.f, 0.f };
    float dbias = 0.0f;
    const int mt = wave / NT, nt = wave - mt * NT;
    const int mrow = 16 * mt + i16;          // this lane's A row
    const int t = mrow / CIN, ci = mrow - t * CIN;
    const int dy = t / 3 - 1, dx = t % 3 - 1;

    if ((int64_t)blockIdx.x < nseg) load_seg(blockIdx.x);
    for (int64_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        __syncthreads();   // the previous segment's reads of s_in / s_d are done
        if (img_task) {
            const bool in = pimg >= 0.0f;
            const float v0 = in ? pimg / 255.0f : 0.0f;
            const float v1 = in ? psr : 0.0f;
#pragma unroll
            for (int sx = -1; sx <= 1; sx++) {
                const int xx = ic - 1 - sx;
                if (xx >= 0 && xx < TX) {
                    s_in[sx + 1][ir][0][xx] = to_bf(v0);
                    s_in[sx + 1][ir][1][xx] = to_bf(v1);
                }
            }
        }
        {
            const __bf16 *vb = reinterpret_cast<const __bf16 *>(&pd);
#pragma unroll
            for (int q = 0; q < 8; q++) s_d[dch * 8 + q][dxx] = vb[q];
        }
        __syncthreads();
        if (seg + gridDim.x < nseg) load_seg(seg + gridDim.x);
        if (tid < COUT) {
            float s = 0.0f;
            for (int xx = 0; xx < TX; xx++) s += from_bf(s_d[tid][xx]);
            dbias += s;
        }
#pragma unroll
        for (int ks = 0; ks < TX / 32; ks++) {
            bf16x8 av = (bf16x8){ 0, 0, 0, 0, 0, 0, 0, 0 };
            if (mrow < M) av = *reinterpret_cast<const bf16x8 *>(&s_in[dx + 1][dy + 1][ci][32 * ks + 8 * kq]);
            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(&s_d[16 * nt + i16][32 * ks + 8 * kq]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
        }
    }
    // write this workgroup's partial sums
    float *part = a.partial + (size_t)blockIdx.x * PSZ;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = 16 * mt + 4 * kq + r, co = 16 * nt + i16;
        if (m < M) part[m * COUT + co] = acc[r];
    }
    if (tid < COUT) part[9 * CIN * COUT + tid] = dbias;
}

// Transposes 8 pixels x 8 channels of BF16 (v[j] = 8 channels of pixel j)
// into out[q] = channel q of the 8 pixels, 16 bytes each.
__device__ __forceinline__ void transpose8x8(const uint4 (&v)[8], uint4 (&out)[8])
{
    const uint32_t *w[8];
#pragma unroll
    for (int j = 0; j < 8; j++) w[j] = reinterpret_cast<const uint32_t *>(&v[j]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t lo[4], hi[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t a = w[2 * i][k], b = w[2 * i + 1][k];
            lo[i] = (a & 0xFFFFu) | (b << 16);          // channel 2k of pixels 2i, 2i+1
            hi[i] = (a >> 16) | (b & 0xFFFF0000u);      // channel 2k+1
        }
        out[2 * k] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
        out[2 * k + 1] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    }
}

// Weight gradients of the 32-input-channel layers (hidden, last), same
// result layout as wgrad_kernel.  The column shift of a tap is moved onto the
// gradient: sum_x in[y+dy][x+dx] d[y][x] = sum_x' in[y+dy][x'] d[y][x'-dx],
// with both factors zero outside the crop, so a segment of 64 input columns
// x' needs its three input rows without a column halo and three copies of
// the one gradient row shifted by dx = -1, 0, 1.  Both are transposed in
// registers (8 pixels x 8 channels per thread) and written to LDS as 16-byte
// rows, K = pixels contiguous for the MFMA operands.
constexpr int WG_PITCH = TX + 8;   // LDS row of 64 pixels + 8 pad (bank spread)

// Register budget for 3 waves per SIMD (3 resident workgroups per CU; with
// no bound the compiler's operand prefetch left 2, a budget for 4 spills).
template <int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) wgrad32_kernel(WgradArgs a)
{
    static_assert(KIND == K_HIDDEN || KIND == K_LAST, "32-channel inputs only");
    constexpr int CIN = C;
    constexpr int COUT = (KIND == K_LAST) ? 1 : C;
    constexpr int M = 9 * CIN;
    constexpr int MT = M / 16;                    // one tap per M-tile
    constexpr int NT = (COUT + 15) / 16;
    constexpr int PAIRS = MT * NT;
    constexpr int PPW = (PAIRS + 3) / 4;
    constexpr int PSZ = 9 * CIN * COUT + COUT;
    constexpr int DROWS = NT * 16;
    __shared__ __align__(16) __bf16 s_in[3][CIN][WG_PITCH];
    __shared__ __align__(16) __bf16 s_d[3][DROWS][WG_PITCH];
    __shared__ float s_bias[64];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i16 = lane & 15, kq = lane >> 4;
    const int segs_x = (a.W + TX - 1) / TX;
    const int64_t nseg = (int64_t)a.N * a.H * segs_x;
    const int64_t npix = (int64_t)a.H * a.W;

    // Staging tasks.  Threads 0..95: input row r = y-1+(tid>>5), channel
    // chunk (tid>>3)&3, pixels 8*(tid&7)..+7.  Hidden, threads 96..191:
    // gradient copy dxi = (u>>5), same chunk/pixel split.  Last layer,
    // threads 0..191: one pixel of one gradient copy (dxi = tid / 64).
    const bool in_task = tid < 96;
    const bool d_task = (KIND == K_HIDDEN) ? (tid >= 96 && tid < 192) : (tid < 192);
    const int u = (KIND == K_HIDDEN) ? tid - 96 : tid;
    const int t_row = tid >> 5, t_chunk = (tid >> 3) & 3, t_xg = tid & 7;
    const int d_dxi = (KIND == K_HIDDEN) ? (u >> 5) : (u >> 6);
    const int d_chunk = (u >> 3) & 3, d_xg = u & 7, d_xx = u & 63;

    uint4 v[8];   // the thread's 8 pixels x 8 channels (input or gradient task)
    __bf16 vd1 = to_bf(0.0f);
    auto load_seg = [&](int64_t seg) {
        const int n = (int)(seg / ((int64_t)a.H * segs_x));
        const int rem = (int)(seg - (int64_t)n * a.H * segs_x);
        const int y = rem / segs_x, x0 = (rem - y * segs_x) * TX;
        if (in_task) {
            const int gy = y - 1 + t_row;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int gx = x0 + 8 * t_xg + j;
                v[j] = make_uint4(0, 0, 0, 0);
                if (gy >= 0 && gy < a.H && gx < a.W)
                    v[j] = *reinterpret_cast<const uint4 *>(a.act + ((size_t)n * npix + (size_t)gy * a.W + gx) * C + t_chunk * 8);
            }
        }
        if (KIND == K_HIDDEN && d_task) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int gx = x0 + 8 * d_xg + j - (d_dxi - 1);
                v[j] = make_uint4(0, 0, 0, 0);
                if (gx >= 0 && gx < a.W)
                    v[j] = *reinterpret_cast<const uint4 *>(a.dpre + ((size_t)n * npix + (size_t)y * a.W + gx) * C + d_chunk * 8);
            }
        }
        if (KIND == K_LAST && d_task) {
            const int gx = x0 + d_xx - (d_dxi - 1);
            vd1 = to_bf(0.0f);
            if (gx >= 0 && gx < a.W) vd1 = a.dpre[(size_t)n * npix + (size_t)y * a.W + gx];
        }
    };

    if (KIND == K_LAST) {   // gradient rows 1..15 of the single output channel stay zero
        for (int e = tid; e < 3 * 15 * TX; e += 256) {
            const int dxi = e / (15 * TX), rem = e - dxi * 15 * TX;
            s_d[dxi][1 + rem / TX][rem % TX] = to_bf(0.0f);
        }
    }

    f32x4 acc[PPW];
#pragma unroll
    for (int q = 0; q < PPW; q++) acc[q] = (f32x4){ 0.f, 0.f, 0.f, 0.f };
    // bias gradient, accumulated by the unshifted gradient copy's tasks:
    // hidden, 8 channels of 8 pixels per thread; last, one pixel per thread
    float bsum[8] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };

    if ((int64_t)blockIdx.x < nseg) load_seg(blockIdx.x);
    for (int64_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        __syncthreads();   // the previous segment's MFMAs are done with the LDS rows
        if (in_task) {
            uint4 t[8];
            transpose8x8(v, t);
#pragma unroll
            for (int q = 0; q < 8; q++)
                *reinterpret_cast<uint4 *>(&s_in[t_row][t_chunk * 8 + q][8 * t_xg]) = t[q];
        }
        if (KIND == K_HIDDEN && d_task) {
            uint4 t[8];
            transpose8x8(v, t);
#pragma unroll
            for (int q = 0; q < 8; q++)
                *reinterpret_cast<uint4 *>(&s_d[d_dxi][d_chunk * 8 + q][8 * d_xg]) = t[q];
            if (d_dxi == 1) {
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const uint32_t *w = reinterpret_cast<const uint32_t *>(&t[q]);
                    float sq = 0.0f;
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        sq += __uint_as_float(w[k] << 16) + __uint_as_float(w[k] & 0xFFFF0000u);
                    bsum[q] += sq;
                }
            }
        }
        if (KIND == K_LAST && d_task) {
            s_d[d_dxi][0][d_xx] = vd1;
            if (d_dxi == 1) bsum[0] += from_bf(vd1);
        }
        __syncthreads();
        // the next segment's loads are in flight during this segment's MFMAs
        if (seg + gridDim.x < nseg) load_seg(seg + gridDim.x);
#pragma unroll
        for (int q = 0; q < PPW; q++) {
            const int pair = wave * PPW + q;
            if (pair >= PAIRS) break;
            const int mt = pair / NT, nt = pair - mt * NT;
            const int t = (16 * mt) / CIN, ci = (16 * mt) % CIN + i16;
            const int dy = t / 3 - 1, dx = t % 3 - 1;
#pragma unroll
            for (int ks = 0; ks < TX / 32; ks++) {
                const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&s_in[dy + 1][ci][32 * ks + 8 * kq]);
                const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(&s_d[dx + 1][16 * nt + i16][32 * ks + 8 * kq]);
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[q], 0, 0, 0);
            }
        }
    }
    float *part = a.partial + (size_t)blockIdx.x * PSZ;
#pragma unroll
    for (int q = 0; q < PPW; q++) {
        const int pair = wave * PPW + q;
        if (pair >= PAIRS) break;
        const int mt = pair / NT, nt = pair - mt * NT;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = 16 * mt + 4 * kq + r, co = 16 * nt + i16;
            if (co < COUT) part[m * COUT + co] = acc[q][r];
        }
    }
    // bias: sum the unshifted copy's per-thread sums in a fixed order
    if (KIND == K_HIDDEN) {
        if (d_task && d_dxi == 1) {
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (d_xg == 0) s_bias[d_chunk * 8 + q] = 0.0f;
        }
        __syncthreads();
        for (int xg = 0; xg < 8; xg++) {
            if (d_task && d_dxi == 1 && d_xg == xg) {
#pragma unroll
                for (int q = 0; q < 8; q++) s_bias[d_chunk * 8 + q] += bsum[q];
            }
            __syncthreads();
        }
        if (tid < COUT) part[9 * CIN * COUT + tid] = s_bias[tid];
    } else {
        if (d_task && d_dxi == 1) s_bias[d_xx] = bsum[0];
        __syncthreads();
        if (tid == 0) {
            float b = 0.0f;
            for (int xx = 0; xx < 64; xx++) b += s_bias[xx];
            part[9 * CIN * COUT] = b;
        }
    }
}

// The per-block partial gradients are summed in two passes with a fixed
// order (training stays bit-reproducible): pass 1 (grid.y = run) sums each
// run of kReduceRun partials into the run's first partial, pass 2 sums the
// runs' totals.  One thread per (gradient entry, run) keeps 32x more loads in
// flight than one serial sum over all partials.
constexpr int kReduceRun = 32;

__global__ void reduce_partial_runs(float *partial, int nparts, int psz)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= psz) return;
    const int b0 = blockIdx.y * kReduceRun;
    const int b1 = min(b0 + kReduceRun, nparts);
    float s = 0.0f;
#pragma unroll 8
    for (int b = b0; b < b1; b++) s += partial[(size_t)b * psz + k];
    partial[(size_t)b0 * psz + k] = s;
}

__global__ void reduce_partials(const float *partial, int nparts, int psz, float *grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= psz) return;
    float s = 0.0f;
#pragma unroll 8
    for (int b = 0; b < nparts; b += kReduceRun) s += partial[(size_t)b * psz + k];
    grad[k] = s;
}

// ------------------------------------------------------------- loss, Adam

// pred vs target/255 on pixels at least MARGIN from the crop border, except
// along the sides that lie on the image's own edge (edges[n]: bit 0 top,
// 1 bottom, 2 left, 3 right; NULL = none): there the target is the exact
// filter's clamp-to-edge result, which the network can see (its inputs are
// zero past the edge), so those pixels are trained too.
// g6 = 2 (pred - y) / count (bf16), loss += (pred - y)^2 / count, with
// count = N (H - 2 MARGIN) (W - 2 MARGIN) whatever the edges (the step size
// does not depend on a batch's mix of edge crops).
// With tail_w > 0 each pixel adds tail_w * max(|d| - tail_t, 0)^2 to the
// loss (d = pred - y): the errors past tail_t, which are what the sampled
// check rejects tiles for, weigh more.  tail_w = 0 leaves the MSE bits as
// they are.
__global__ void loss_kernel(const float *pred, const uint8_t *target, const uint8_t *edges,
                            int N, int H, int W, float tail_w, float tail_t, __bf16 *g6,
                            float *loss)
{
    __shared__ float red[256];
    const int64_t rows = (int64_t)N * H;
    const int64_t count = (int64_t)N * (H - 2 * MARGIN) * (W - 2 * MARGIN);
    const float inv = 1.0f / (float)count;
    float acc = 0.0f;
    // a workgroup per image row at a time (its crop, row and margins set
    // once per row), its threads along the row
    for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const int n = (int)(row / H), y = (int)(row - (int64_t)n * H);
        const int e = edges ? edges[n] : 0;
        const int mt = (e & 1) ? 0 : MARGIN, mb = (e & 2) ? 0 : MARGIN;
        const int ml = (e & 4) ? 0 : MARGIN, mr = (e & 8) ? 0 : MARGIN;
        const bool yin = y >= mt && y < H - mb;
        for (int x = threadIdx.x; x < W; x += blockDim.x) {
            const int64_t i = row * W + x;
            float g = 0.0f;
            if (yin && x >= ml && x < W - mr) {
                const float d = pred[i] - (float)target[i] / 255.0f;
                float q = d * d, gd = 2.0f * d;
                if (tail_w > 0.0f) {
                    const float a = fabsf(d) - tail_t;
                    if (a > 0.0f) {
                        q += tail_w * a * a;
                        gd += 2.0f * tail_w * copysignf(a, d);
                    }
                }
                acc += q * inv;
                g = gd * inv;
            }
            g6[i] = to_bf(g);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

__global__ void zero_floats(float *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void adam_kernel(float *w, float *m, float *v, const float *g, int n, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
}

struct PackTables {
    __bf16 *fw[NLAYERS];
    __bf16 *bw[NLAYERS];
};

// BF16 operand tables from the FP32 master weights.
//   FW_l[nt][j][k]: k = t*CIN + ci < 9*CIN, co = 16 nt + j < COUT: W[t][ci][co]
//   BW_l[nt][j][k]: k = t'*COUT + co < 9*COUT, ci = 16 nt + j:     W[8-t'][ci][co]
__global__ void pack_kernel(const float *params, PackTables pk)
{
    const int l = blockIdx.y;
    const int cin = cin_of(l), cout = cout_of(l);
    const float *W = params + poff(l);
    const int nf = fw_size(l), nb = l == 0 ? 0 : bw_size(l);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nf + nb; e += gridDim.x * blockDim.x) {
        if (e < nf) {
            const int kf = kf_of(l);
            const int row = e / kf, k = e - row * kf;
            const int co = row;   // row = nt*16 + j
            float v = 0.0f;
            if (k < 9 * cin && co < cout) {
                const int t = k / cin, ci = k - t * cin;
                v = W[(t * cin + ci) * cout + co];
            }
            pk.fw[l][e] = to_bf(v);
        } else {
            const int f = e - nf, kb = kb_of(l);
            const int row = f / kb, k = f - row * kb;
            const int ci = row;
            float v = 0.0f;
            if (k < 9 * cout && ci < cin) {
                const int tp = k / cout, co = k - tp * cout;
                v = W[((8 - tp) * cin + ci) * cout + co];
            }
            pk.bw[l][f] = to_bf(v);
        }
    }
}

// ------------------------------------------------------- fused inference
//
// All six layers of the forward pass in one kernel, for inference: no
// activation leaves the chip.  A workgroup owns a strip of FS_W output
// columns and a band of rows and streams down the band two rows per step.
// At step k, hidden layer l (1..4) computes rows Y0 + 2k + 20 - 4l and the
// row below, the first layer (0) rows Y0 + 2k + 19 and + 20, and the output
// rows are Y0 + 2k and + 1 (64 columns starting 5-l left of the strip for
// layer l: the context the layers above need).  A hidden layer writes its
// first row at the end of the step and its second row during the next step
// (its epilogue overlaps that step's first MFMAs), so the layer above, four
// rows behind, reads a row two steps after it was computed; each activation
// ring holds FS_RING = 4 rows, the last layer's product ring FS_DRING = 8.
//
// Eight waves, two per SIMD.  Wave w runs hidden layer 1 + w/2 on M-tiles
// 2(w%2) and 2(w%2)+1 (16 pixels each), both rows, both 16-channel N-tiles:
// 72 MFMAs per step.  The fragments of a layer's input rows r+1 and r+2 are
// the input rows r'-1 and r' of the next step, so they stay in registers:
// each step reads only two new input rows from LDS, and starts on the
// fragments it already holds (and on the previous step's deferred
// epilogue) while the new ones arrive.  Waves 0..3 also
// run M-tile w of the first layer (its input is a ring of image dwords
// bf16(u/255) | bf16(sigma_r/100) << 16, zero outside the image); waves 6,7
// (layer 4) feed their own BF16 outputs straight from registers into the
// last layer, computed per pixel as D[t][x] = sum_ci W6[t][ci] * act[x][ci]
// for the nine taps t (one MFMA with the taps as rows), written to a ring of
// FP32 products; waves 4,5 sum each output pixel's nine products in tap
// order, add the residual and store.  Waves 4..7 stream the image: each
// step they write two image rows loaded into registers the step before and
// load the next two, so a load has a whole step to arrive.
//
// The MFMAs compute the transposed product (weights as the A operand,
// activations as B), so a lane ends with four consecutive channels of one
// pixel and stores them with one 8-byte write.  Bias is the accumulator's
// initial value; the epilogue rounds to BF16 and then applies the ReLU to
// the BF16 bits.  Activations are 64-byte NHWC pixels in 64-pixel ring
// rows; the 16-byte chunks of a pixel are permuted by bits 1..2 of its
// column (fs_col), so the ds_read_b128 and ds_write_b64 of a step run at
// the conflict-free rate.  Each wave lowers its priority through the step
// (the ladder in `step`) so that a SIMD's two waves stay in step.
//
// Operands, their order and every rounding are those of conv_kernel, so the
// output equals the per-layer path's bit for bit.

constexpr int FS_W = 54;                  // output columns per strip
constexpr int FS_WAVES = 8;
constexpr int FS_RING = 4;                // rows per activation ring
constexpr int FS_DRING = 8;               // rows of the last layer's product ring
constexpr int FS_ROWB = 64 * 64;          // bytes per activation ring row
constexpr int FS_D_TAPS = 16;             // product rows per ring row (9 taps used)
constexpr int FS_D_PITCH = 68;            // words per product row (64 columns + pad)
constexpr int FS_D_ROWB = FS_D_TAPS * FS_D_PITCH * 4;
constexpr int FS_D_OFF = 4 * FS_RING * FS_ROWB;                        // rings of layers 0..3 first
constexpr int FS_IMG_ROWS = 64;           // image ring rows
constexpr int FS_IMG_DUP = 3;             // ring rows 0..2 repeated after row 63
constexpr int FS_IMG_PITCH = 80;          // image ring columns
constexpr int FS_IMG_W = 66;              // image columns a strip reads
constexpr int FS_IMGW_OFF = FS_D_OFF + FS_DRING * FS_D_ROWB;           // dword ring (first layer)
constexpr int FS_IMGU_OFF = FS_IMGW_OFF + (FS_IMG_ROWS + FS_IMG_DUP) * FS_IMG_PITCH * 4;  // u8 ring
constexpr int FS_LUT_OFF = FS_IMGU_OFF + FS_IMG_ROWS * FS_IMG_PITCH;   // image dword of u = 0..255
constexpr int FS_CB_OFF = FS_LUT_OFF + 256 * 4;                       // biases of layers 0..4 (32 floats each)
constexpr int FS_W0_OFF = FS_CB_OFF + 5 * 32 * 4;                      // first layer's A operand (64 lanes x 2 x 16 B)
constexpr int FS_W6_OFF = FS_W0_OFF + 2 * 64 * 16;                     // last layer's A operand (64 lanes x 16 B)
constexpr int FS_RES_OFF = FS_W6_OFF + 64 * 16;                        // residual u / 255 (FP32), u = 0..255
constexpr int FS_LDS = FS_RES_OFF + 256 * 4;
constexpr int FS_PRE = 4;                 // image rows written before the first step
constexpr int FS_KBEGIN = -12;            // first step: layer 0 computes rows Y0-5, Y0-4
static_assert(FS_IMG_W <= 128, "an image row fits one byte per lane of two waves");
static_assert(FS_LDS <= 160 * 1024, "LDS");
static_assert(FS_IMG_PITCH >= FS_IMG_W, "image ring rows");

struct FusedArgs {
    const uint8_t *img;
    int64_t img_pitch;
    uint8_t *out;
    int64_t out_pitch;
    int H, W;
    int band;                     // rows per band
    int strips;                   // strips per band row
    float sr;                     // sigma_range / 100
    const __bf16 *fw[NLAYERS];    // packed forward weight tables
    const float *bias[NLAYERS];
};

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) bf16x8 fs_lds_b128;
typedef __attribute__((address_space(3))) u32x2 fs_lds_u32x2;
typedef __attribute__((address_space(3))) uint32_t fs_lds_u32;
typedef __attribute__((address_space(3))) float fs_lds_f32;
typedef __attribute__((address_space(3))) unsigned char fs_lds_u8;
typedef __attribute__((address_space(3))) __bf16 fs_lds_bf;

__device__ __forceinline__ bf16x8 fs_read_frag(uint32_t addr) { return *(const fs_lds_b128 *)(size_t)addr; }
__device__ __forceinline__ void fs_write_b64(uint32_t addr, u32x2 v) { *(fs_lds_u32x2 *)(size_t)addr = v; }
__device__ __forceinline__ uint32_t fs_read_u32(uint32_t addr) { return *(const fs_lds_u32 *)(size_t)addr; }
__device__ __forceinline__ void fs_write_u32(uint32_t addr, uint32_t v) { *(fs_lds_u32 *)(size_t)addr = v; }
__device__ __forceinline__ float fs_read_f32(uint32_t addr) { return *(const fs_lds_f32 *)(size_t)addr; }
__device__ __forceinline__ f32x4 fs_read_f4(uint32_t addr) { return *(const __attribute__((address_space(3))) f32x4 *)(size_t)addr; }
__device__ __forceinline__ void fs_write_f32(uint32_t addr, float v) { *(fs_lds_f32 *)(size_t)addr = v; }
__device__ __forceinline__ uint8_t fs_read_u8(uint32_t addr) { return *(const fs_lds_u8 *)(size_t)addr; }
__device__ __forceinline__ void fs_write_u8(uint32_t addr, uint8_t v) { *(fs_lds_u8 *)(size_t)addr = v; }
__device__ __forceinline__ void fs_write_bf(uint32_t addr, __bf16 v) { *(fs_lds_bf *)(size_t)addr = v; }
__device__ __forceinline__ __bf16 fs_read_bf(uint32_t addr) { return *(const fs_lds_bf *)(size_t)addr; }

// Four consecutive channels of one pixel after a transposed MFMA (bias
// included): BF16, ReLU, as two dwords.
__device__ __forceinline__ u32x2 fs_act4(f32x4 acc)
{
    return (u32x2){ act_pack2(acc[0], acc[1]), act_pack2(acc[2], acc[3]) };
}

// Byte offset, within a ring row, of 16-byte chunk `chunk` of column x.
// The chunk is XORed with bits 1..2 of x.  Measured on gfx950 (256
// workgroups of 8 waves): the fragment read (ds_read_b128 of chunk kq of 16
// consecutive columns starting at any column) and the epilogue's ds_write_b64
// then run at the conflict-free rate; XORing with (b1^b2, b2^b3) of x, or
// with no swizzle, made the reads starting at an odd column (or all writes)
// two to four times slower.
__device__ __forceinline__ uint32_t fs_col(int x, int chunk)
{
    const int sw = (x >> 1) & 3;
    return (uint32_t)(x * 64 + ((chunk ^ sw) << 4));
}

// LDS writes visible to the whole workgroup, without waiting for the
// register prefetch of image rows still in flight.
__device__ __forceinline__ void fs_step_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ __forceinline__ uint8_t fs_load_img(const FusedArgs &a, int row, int col)
{
    row = row < 0 ? 0 : (row >= a.H ? a.H - 1 : row);
    col = col < 0 ? 0 : (col >= a.W ? a.W - 1 : col);
    return a.img[(int64_t)row * a.img_pitch + col];
}

// Slot of a row in the product ring.
__device__ __forceinline__ int fs_dslot(int row, int Y0) { return (row - Y0) & (FS_DRING - 1); }

template <int I> using fs_ic = std::integral_constant<int, I>;

__global__ void __launch_bounds__(FS_WAVES * 64) fused_infer_kernel(FusedArgs a)
{
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int i16 = lane & 15, kq = lane >> 4;
    const int strip = blockIdx.x % a.strips, bandi = blockIdx.x / a.strips;
    const int x0 = strip * FS_W;
    const int Y0 = bandi * a.band;
    const int Y1 = min(Y0 + a.band, a.H);
    const int KE = (Y1 - Y0 + 1) / 2;         // steps with output rows
    const int H = a.H, W = a.W;
    const int img_base = Y0 - 6;              // image row of ring row 0
    const int img_x = x0 - 6;                 // image column of ring column 0

    auto img_slot = [&](int row) { return (row - img_base) & (FS_IMG_ROWS - 1); };

    // ---- hidden layer L on M-tiles m0, m0 + 1: weights (A), bias (C)
    const int L = 1 + (wave >> 1);
    const int m0 = 2 * (wave & 1);
    const __bf16 *fw = L == 1 ? a.fw[1] : (L == 2 ? a.fw[2] : (L == 3 ? a.fw[3] : a.fw[4]));
    bf16x8 wA[9][2];
#pragma unroll
    for (int s = 0; s < 9; s++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
            wA[s][nt] = *reinterpret_cast<const bf16x8 *>(fw + ((size_t)(nt * 16 + i16) * 288 + s * 32 + 8 * kq));
    // reads of layer L-1: column c + 1 + dx for c = 16 m0 + i16 (+ 16 for the second M-tile)
    uint32_t ha[3];
#pragma unroll
    for (int d = 0; d < 3; d++) ha[d] = fs_col(16 * m0 + i16 + d, kq);
    // writes of layer L: channels 16 nt + 4 kq .. +3 of column 16 m0 + i16
    uint32_t hw[2];
#pragma unroll
    for (int nt = 0; nt < 2; nt++) hw[nt] = fs_col(16 * m0 + i16, 2 * nt + (kq >> 1)) + (kq & 1) * 8;
    // whether this lane's column of layer L's two M-tiles lies in the image
    bool hcol[2];
#pragma unroll
    for (int mm = 0; mm < 2; mm++) {
        const int c0 = x0 - (5 - L) + 16 * (m0 + mm);
        hcol[mm] = c0 + i16 >= 0 && c0 + i16 < W;
    }
    const uint32_t h_in = (uint32_t)((L - 1) * FS_RING * FS_ROWB);
    const uint32_t h_out = (uint32_t)(L * FS_RING * FS_ROWB);     // unused by layer 4
    // per-lane ring addresses without the row slot: the slot of every row a
    // step touches is a compile-time constant of the step's parity (RP), so
    // it goes into the ds_read/ds_write immediate
    uint32_t hb[3], hwb[2];
#pragma unroll
    for (int d = 0; d < 3; d++) hb[d] = h_in + ha[d];
#pragma unroll
    for (int nt = 0; nt < 2; nt++) hwb[nt] = h_out + hw[nt];

    // ---- role data
    const bool first_w = wave < 4;            // first layer, M-tile `wave`
    // this lane's first-layer taps 4 kq + q; taps >= 9 are K padding whose
    // weights are zero: they read the centre tap's (finite, >= 0) dword, so
    // their products are +0, as the per-layer kernel's zero operands give
    uint32_t ioff[4] = {};
    bool fcol = false;
    // first layer's ring addresses of this lane's channels, without the slot
    const uint32_t fw1a = fs_col(16 * wave + i16, kq >> 1) + (kq & 1) * 8;
    const uint32_t fw1b = fs_col(16 * wave + i16, 2 + (kq >> 1)) + (kq & 1) * 8;
    if (first_w) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int t = 4 * kq + q < 9 ? 4 * kq + q : 4;
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            ioff[q] = (uint32_t)(((dy + 1) * FS_IMG_PITCH + 16 * wave + i16 + 1 + dx) * 4);
        }
        const int c0 = x0 - 5 + 16 * wave;
        fcol = c0 + i16 >= 0 && c0 + i16 < W;
    }
    const float b6 = a.bias[NLAYERS - 1][0];

    // ---- tables, per-layer constants and the first FS_PRE image rows
    if (tid < 256) {
        // the first layer's input dword of a pixel u inside the image
        const uint32_t sr_hi = (uint32_t)__builtin_bit_cast(uint16_t, to_bf(a.sr)) << 16;
        fs_write_u32(FS_LUT_OFF + 4 * tid, (uint32_t)__builtin_bit_cast(uint16_t, to_bf((float)tid / 255.0f)) | sr_hi);
        fs_write_f32(FS_RES_OFF + 4 * tid, (float)tid / 255.0f);
    }
    if (tid < 5 * 32) fs_write_f32(FS_CB_OFF + 4 * tid, a.bias[tid >> 5][tid & 31]);
    if (tid < 128) {
        // first layer A operand of lane (kq, i16), N-tile nt: W0[co = 16 nt + i16][k = 8 kq ..]
        const int nt = tid >> 6, l = tid & 63;
        const u32x4 v = *reinterpret_cast<const u32x4 *>(a.fw[0] + ((size_t)(nt * 16 + (l & 15)) * 32 + 8 * (l >> 4)));
        *(__attribute__((address_space(3))) u32x4 *)(size_t)(FS_W0_OFF + 16 * tid) = v;
    } else if (tid < 192) {
        // last layer A operand: A[t][k], tap t = i16 (< 9), k-slot (kq, q) = channel
        // q < 4 ? 4 kq + q : 16 + 4 kq + q - 4
        const int l = tid - 128, t = l & 15, q4 = 4 * (l >> 4);
        bf16x8 v = {};
        if (t < 9) v = cat4(a.fw[NLAYERS - 1] + t * 32 + q4, a.fw[NLAYERS - 1] + t * 32 + 16 + q4);
        *(__attribute__((address_space(3))) bf16x8 *)(size_t)(FS_W6_OFF + 16 * l) = v;
    }
    fs_step_barrier();
    // image byte u of ring column cc (image column img_x + cc, inside the
    // image when col_in) into both rings
    auto put_img_in = [&](int row, int cc, uint8_t u, bool col_in) {
        const int sl = img_slot(row);
        const bool in = row >= 0 && row < H && col_in;
        const uint32_t wv = in ? fs_read_u32(FS_LUT_OFF + 4 * u) : 0u;
        fs_write_u8(FS_IMGU_OFF + sl * FS_IMG_PITCH + cc, u);
        fs_write_u32(FS_IMGW_OFF + (sl * FS_IMG_PITCH + cc) * 4, wv);
        if (sl < FS_IMG_DUP) fs_write_u32(FS_IMGW_OFF + ((sl + FS_IMG_ROWS) * FS_IMG_PITCH + cc) * 4, wv);
    };
    auto put_img = [&](int row, int cc, uint8_t u) {
        const int gx = img_x + cc;
        put_img_in(row, cc, u, gx >= 0 && gx < W);
    };
    for (int e = tid; e < FS_PRE * FS_IMG_W; e += FS_WAVES * 64) {
        const int rr = e / FS_IMG_W, cc = e - rr * FS_IMG_W;
        const int row = img_base + rr;
        put_img(row, cc, fs_load_img(a, row, img_x + cc));
    }
    fs_step_barrier();

    // hidden layer bias, the accumulators' initial value: read from LDS
    // before each use (two reads a step) rather than held in registers
    const uint32_t bias_off = (uint32_t)(FS_CB_OFF + 128 * L + 16 * kq);
    auto read_bias = [&](f32x4 (&b)[2]) __attribute__((always_inline)) {
#pragma unroll
        for (int nt = 0; nt < 2; nt++) b[nt] = fs_read_f4(bias_off + 64 * nt);
    };

    bf16x8 fr[3][3][2];                       // [row slot][dx][M-tile] fragments of layer L-1
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int mm = 0; mm < 2; mm++) fr[i][d][mm] = (bf16x8){};
    // second-row accumulators, finished during the next step (the first
    // step finishes these initial values into a row no needed row reads)
    f32x4 acc1[2][2];
    {
        f32x4 b[2];
        read_bias(b);
#pragma unroll
        for (int mm = 0; mm < 2; mm++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++) acc1[mm][nt] = b[nt];
    }
    // Image streaming (waves 4..7): at step s = k - FS_KBEGIN the first layer
    // reads image rows img_base + 2s .. + 2s + 3.  Step s writes rows
    // img_base + 2s + 4 and + 5 into the ring (visible from step s + 1) from
    // the bytes loaded at step s - 1, then loads rows + 6 and + 7: waves 4, 5
    // the first of each pair, waves 6, 7 the second, one byte per lane
    // (lanes past the row's FS_IMG_W bytes load a valid byte and drop it).
    const int sq = tid - 256;                 // waves 4..7
    const int prow = sq >> 7;                 // waves 4, 5: the first row; 6, 7: the second
    const int pcc = min(sq & 127, FS_IMG_W - 1);
    const bool pact = (sq & 127) < FS_IMG_W;
    const int pcol = min(max(img_x + pcc, 0), W - 1);
    const bool pin = img_x + pcc >= 0 && img_x + pcc < W;
    uint8_t pb = 0;
    if (tid >= 256) pb = fs_load_img(a, img_base + FS_PRE + prow, img_x + pcc);

    // Every layer computes its two rows at every step, needed or not: a row
    // outside a layer's needed range may hold garbage, but a needed row reads
    // only needed rows of the layer below (its range is one row wider on each
    // side), and those were written at earlier steps, so garbage never
    // reaches an output.  Only the output stores are conditional.  The
    // fragments a step carries in are the ones the previous step loaded.
    //
    // Role R_FIRST (waves 0..3): the first layer's M-tile `wave`; R_GATHER
    // (waves 4, 5, layer 3): output rows and image prefetch; R_LAST (waves 6,
    // 7, layer 4): the last layer's products.  MASKED: zero every activation
    // outside the image (workgroups at an image border).
    constexpr int R_FIRST = 0, R_GATHER = 1, R_LAST = 2;
    auto step = [&](auto ROLE, auto MASKED, auto SA, auto SB, auto SC, auto RPH, int k) __attribute__((always_inline)) {
        constexpr int RL = decltype(ROLE)::value;
        constexpr int RP = decltype(RPH)::value;   // k & 1: row r lies in ring slot 2 RP
        constexpr bool MK = decltype(MASKED)::value;
        constexpr int A = decltype(SA)::value, B = decltype(SB)::value, Cs = decltype(SC)::value;
        const int r = Y0 + 2 * k + 20 - 4 * L;         // this layer's rows r, r + 1
        auto rd = [&](auto OFF, int d, int mm) {         // row r + OFF
            constexpr uint32_t sl = (uint32_t)((2 * RP + decltype(OFF)::value) & (FS_RING - 1));
            return fs_read_frag(hb[d] + sl * FS_ROWB + (uint32_t)(mm * 1024));
        };

        const int r0 = Y0 + 2 * k + 19;                // first layer's rows (R_FIRST)
        const int yg = Y0 + 2 * k + (wave - 4);        // output row (R_GATHER)
        bf16x8 w6;
        if constexpr (RL == R_LAST) w6 = fs_read_frag(FS_W6_OFF + 16 * lane);

        // epilogue of one row: BF16 + ReLU (zero outside the image), then the
        // ring write, or for layer 4 the last layer's products
        auto finish = [&](const f32x4 (&acc)[2][2], int row, auto SLOT) __attribute__((always_inline)) {
            constexpr uint32_t sl = (uint32_t)decltype(SLOT)::value;
            const bool row_in = row >= 0 && row < H;
#pragma unroll
            for (int mm = 0; mm < 2; mm++) {
                u32x2 p[2];
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    p[nt] = fs_act4(acc[mm][nt]);
                    if constexpr (MK) {
                        const bool in = hcol[mm] && row_in;
                        p[nt] = in ? p[nt] : (u32x2){ 0u, 0u };
                    }
                }
                if constexpr (RL != R_LAST) {
                    fs_write_b64(hwb[0] + sl * FS_ROWB + (uint32_t)(mm * 1024), p[0]);
                    fs_write_b64(hwb[1] + sl * FS_ROWB + (uint32_t)(mm * 1024), p[1]);
                } else {
                    // D[t][x] = sum over the 32 channels of W6[t] * act[x]
                    const bf16x8 bv = __builtin_bit_cast(bf16x8, (u32x4){ p[0][0], p[0][1], p[1][0], p[1][1] });
                    const f32x4 dsum = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w6, bv, (f32x4){ 0.f, 0.f, 0.f, 0.f }, 0, 0, 0);
                    const uint32_t db = FS_D_OFF + (uint32_t)(fs_dslot(row, Y0) * FS_D_ROWB) +
                                        (uint32_t)(((4 * kq) * FS_D_PITCH + 16 * (m0 + mm) + i16) * 4);
#pragma unroll
                    for (int q = 0; q < 4; q++) fs_write_f32(db + q * FS_D_PITCH * 4, dsum[q]);
                }
            }
        };

        // ---- the MFMAs.  Each accumulator takes its taps in order 0..8 from
        // the bias, as the per-layer kernel does.  Order: row r's taps on row
        // r-1 (slot A) with the previous step's row r-1 epilogue beside them
        // (it reads acc1 before acc1 is reused); row r's taps on row r (slot
        // B) while row r+1 loads into slot C; row r+1's taps on row r; row
        // r+2 loads into slot A (row r-1 is consumed) during the role's
        // work; row r's taps on row r+1; row r+1's on rows r+1 and r+2, with
        // row r's epilogue beside the last 24 MFMAs.  Loading each row only
        // when a slot frees keeps two row slots live, not three.
        f32x4 bias[2];
        auto mf = [&](f32x4 (&acc)[2][2], int s, const bf16x8 (&f)[3][2], int d) __attribute__((always_inline)) {
#pragma unroll
            for (int mm = 0; mm < 2; mm++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
                    acc[mm][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        wA[s][nt], f[d][mm], s == 0 ? bias[nt] : acc[mm][nt], 0, 0, 0);
        };
        // Wave priority falls 3, 2, 1, 0 through the step's four phases, so
        // of a SIMD's two waves the one behind issues first and the two
        // reach the barrier together (the arbiter otherwise favours the
        // older wave, and the younger one ends the step alone).
        __builtin_amdgcn_s_setprio(3);
        f32x4 acc0[2][2];
        read_bias(bias);
#pragma unroll
        for (int d = 0; d < 3; d++) mf(acc0, d, fr[A], d);
        finish(acc1, r - 1, fs_ic<(2 * RP + 3) & (FS_RING - 1)>());
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(2);
        // row r+1 into slot C
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int mm = 0; mm < 2; mm++) fr[Cs][d][mm] = rd(fs_ic<1>(), d, mm);
#pragma unroll
        for (int d = 0; d < 3; d++) mf(acc0, 3 + d, fr[B], d);
        // the role's inputs
        u32x4 fdw[2];
        bf16x8 fw0[2];
        f32x4 fb0[2];
        float gv[9];
        uint8_t px = 0;
        if constexpr (RL == R_FIRST) {
#pragma unroll
            for (int rw = 0; rw < 2; rw++) {
                const uint32_t base = FS_IMGW_OFF + (uint32_t)(img_slot(r0 + rw - 1) * FS_IMG_PITCH * 4);
#pragma unroll
                for (int q = 0; q < 4; q++) fdw[rw][q] = fs_read_u32(base + ioff[q]);
            }
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                fw0[nt] = fs_read_frag(FS_W0_OFF + 1024 * nt + 16 * lane);
                fb0[nt] = fs_read_f4(FS_CB_OFF + 64 * nt + 16 * kq);
            }
        } else if constexpr (RL == R_GATHER) {
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int dy = t / 3 - 1, dx = t % 3 - 1;
                gv[t] = fs_read_f32(FS_D_OFF + (uint32_t)(fs_dslot(yg + dy, Y0) * FS_D_ROWB) +
                                    (uint32_t)((t * FS_D_PITCH + lane + 1 + dx) * 4));
            }
            px = fs_read_u8(FS_IMGU_OFF + img_slot(yg) * FS_IMG_PITCH + lane + 6);
        }

        read_bias(bias);
#pragma unroll
        for (int d = 0; d < 3; d++) mf(acc1, d, fr[B], d);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(1);
        // row r+2 into slot A (row r-1 is consumed)
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int mm = 0; mm < 2; mm++) fr[A][d][mm] = rd(fs_ic<2>(), d, mm);

        // ---- the role's own work (independent of this step's MFMAs)
        if constexpr (RL == R_FIRST) {
#pragma unroll
            for (int rw = 0; rw < 2; rw++) {
                const int row = r0 + rw;
                const bf16x8 av = __builtin_bit_cast(bf16x8, fdw[rw]);
                u32x2 p[2];
#pragma unroll
                for (int nt = 0; nt < 2; nt++) {
                    p[nt] = fs_act4(__builtin_amdgcn_mfma_f32_16x16x32_bf16(fw0[nt], av, fb0[nt], 0, 0, 0));
                    if constexpr (MK) {
                        const bool in = fcol && row >= 0 && row < H;
                        p[nt] = in ? p[nt] : (u32x2){ 0u, 0u };
                    }
                }
                const uint32_t wb = (uint32_t)(((2 * RP + 3 + rw) & (FS_RING - 1)) * FS_ROWB);
                fs_write_b64(fw1a + wb, p[0]);
                fs_write_b64(fw1b + wb, p[1]);
            }
        } else if constexpr (RL == R_LAST) {
            const int rs = img_base + FS_PRE + 2 * (k - FS_KBEGIN);
            if (pact) put_img_in(rs + prow, pcc, pb, pin);
            pb = (a.img + (int64_t)min(rs + 2 + prow, H - 1) * a.img_pitch)[pcol];
        } else if constexpr (RL == R_GATHER) {
            const int rs = img_base + FS_PRE + 2 * (k - FS_KBEGIN);
            if (pact) put_img_in(rs + prow, pcc, pb, pin);
            pb = (a.img + (int64_t)min(rs + 2 + prow, H - 1) * a.img_pitch)[pcol];
            // output row: residual + the products summed in tap order
            float gsum = gv[0];
#pragma unroll
            for (int t = 1; t < 9; t++) gsum = gsum + gv[t];
            const int gx = x0 + lane;
            if (k >= 0 && yg < Y1 && lane < FS_W && gx < W) {
                const float v = fs_read_f32(FS_RES_OFF + 4 * px) + (gsum + b6);
                const float q = rintf(v * 255.0f);
                a.out[(int64_t)yg * a.out_pitch + gx] = (uint8_t)(q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q));
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_setprio(0);

#pragma unroll
        for (int d = 0; d < 3; d++) mf(acc0, 6 + d, fr[Cs], d);
#pragma unroll
        for (int d = 0; d < 3; d++) mf(acc1, 3 + d, fr[Cs], d);
        finish(acc0, r, fs_ic<(2 * RP) & (FS_RING - 1)>());
#pragma unroll
        for (int d = 0; d < 3; d++) mf(acc1, 6 + d, fr[A], d);
        fs_step_barrier();
    };

    // Steps k0 .. k1 - 1 (k0 on the 6-step grid from FS_KBEGIN).
    auto run = [&](auto ROLE, auto MASKED, int k0, int k1) __attribute__((always_inline)) {
        static_assert((FS_KBEGIN & 1) == 0, "the unrolled steps start on an even k");
        for (int k = k0; k < k1; k += 6) {
            step(ROLE, MASKED, fs_ic<0>(), fs_ic<1>(), fs_ic<2>(), fs_ic<0>(), k);
            if (k + 1 >= k1) break;
            step(ROLE, MASKED, fs_ic<2>(), fs_ic<0>(), fs_ic<1>(), fs_ic<1>(), k + 1);
            if (k + 2 >= k1) break;
            step(ROLE, MASKED, fs_ic<1>(), fs_ic<2>(), fs_ic<0>(), fs_ic<0>(), k + 2);
            if (k + 3 >= k1) break;
            step(ROLE, MASKED, fs_ic<0>(), fs_ic<1>(), fs_ic<2>(), fs_ic<1>(), k + 3);
            if (k + 4 >= k1) break;
            step(ROLE, MASKED, fs_ic<2>(), fs_ic<0>(), fs_ic<1>(), fs_ic<0>(), k + 4);
            if (k + 5 >= k1) break;
            step(ROLE, MASKED, fs_ic<1>(), fs_ic<2>(), fs_ic<0>(), fs_ic<1>(), k + 5);
        }
    };
    auto run_all = [&](auto ROLE) __attribute__((always_inline)) {
        // Masking matters only where a step writes a row or column outside
        // the image: every column of a strip at the image's side, and at the
        // top and bottom of the image the steps whose written rows (Y0 + 2k
        // + 3 .. Y0 + 2k + 20) leave [0, H).  The other steps run the
        // unmasked loop (same results: nothing there is masked), so a band
        // at the image's top or bottom spends almost all its time in the
        // same code as an interior band.  Boundaries lie on the 6-step grid.
        const bool cols_in = x0 - 5 >= 0 && x0 + 58 < W;
        if (!cols_in) {
            run(ROLE, std::true_type(), FS_KBEGIN, KE);
            return;
        }
        auto grid_up = [&](int k) { return k <= FS_KBEGIN ? FS_KBEGIN : FS_KBEGIN + 6 * ((k - FS_KBEGIN + 5) / 6); };
        auto grid_down = [&](int k) { return k <= FS_KBEGIN ? FS_KBEGIN : FS_KBEGIN + 6 * ((k - FS_KBEGIN) / 6); };
        // first step with Y0 + 2k + 3 >= 0; first step with Y0 + 2k + 20 >= H
        const int kt = (-3 - Y0 + 1) >> 1;
        const int kb = (H - 20 - Y0 + 1) >> 1;
        const int k1 = min(grid_up(kt), KE);
        const int k2 = max(min(grid_down(kb), KE), k1);
        if (k1 > FS_KBEGIN) run(ROLE, std::true_type(), FS_KBEGIN, k1);
        if (k2 > k1) run(ROLE, std::false_type(), k1, k2);
        if (KE > k2) run(ROLE, std::true_type(), k2, KE);
    };
    const int role = wave < 4 ? R_FIRST : (wave < 6 ? R_GATHER : R_LAST);
    if (role == R_FIRST) run_all(fs_ic<R_FIRST>());
    else if (role == R_GATHER) run_all(fs_ic<R_GATHER>());
    else run_all(fs_ic<R_LAST>());
}

__global__ void zero_ints(int32_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace sur
}  // namespace bfimpl

// ================================================================== host

namespace bfimpl {
namespace sur {

struct Net {
    int device = -1;
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    __bf16 *fw[NLAYERS] = {}, *bw[NLAYERS] = {};
    int step = 0;
    float tail_w = 0.0f, tail_t = 0.0f;   // loss tail term (loss_kernel)
    // activation workspace
    size_t ws_pix = 0;
    __bf16 *act[5] = {};         // layer 1..5 outputs, NHWC 32
    __bf16 *dbuf[2] = {};        // gradient ping-pong, NHWC 32
    __bf16 *g6 = nullptr;        // last-layer gradient, 1 ch
    float *pred = nullptr;
    float *partial = nullptr;
    int partial_blocks = 0;
    // fallback workspace
    int32_t *list = nullptr, *count = nullptr;
    int fb_tiles = 0;
};

Net *net_of(bf_sur h) { return reinterpret_cast<Net *>(h); }

int ensure_ws(Net *net, size_t pix)
{
    if (pix <= net->ws_pix) return kOk;
    for (auto &p : net->act) { (void)hipFree(p); p = nullptr; }
    for (auto &p : net->dbuf) { (void)hipFree(p); p = nullptr; }
    (void)hipFree(net->g6);
    (void)hipFree(net->pred);
    net->g6 = nullptr;
    net->pred = nullptr;
    net->ws_pix = 0;
    for (auto &p : net->act) BF_HIP(hipMalloc(&p, pix * C * sizeof(__bf16)));
    for (auto &p : net->dbuf) BF_HIP(hipMalloc(&p, pix * C * sizeof(__bf16)));
    BF_HIP(hipMalloc(&net->g6, pix * sizeof(__bf16)));
    BF_HIP(hipMalloc(&net->pred, pix * sizeof(float)));
    net->ws_pix = pix;
    return kOk;
}

int pack(Net *net, hipStream_t s)
{
    PackTables pk;
    for (int l = 0; l < NLAYERS; l++) {
        pk.fw[l] = net->fw[l];
        pk.bw[l] = net->bw[l];
    }
    hipLaunchKernelGGL(pack_kernel, dim3(16, NLAYERS), dim3(256), 0, s, net->params, pk);
    BF_HIP(hipGetLastError());
    return kOk;
}

// Workgroups of 256 threads of `kern` resident on the whole device at once.
int resident_blocks(const void *kern)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess)
        return -1;
    return cus * per_cu;
}

template <int KIND>
int launch_conv(const ConvArgs &a, hipStream_t s)
{
    // Persistent: as many workgroups as are resident at once walk the 64 x 4
    // tiles (a workgroup beyond that would start only when one has finished).
    static const int resident = resident_blocks(reinterpret_cast<const void *>(&conv_kernel<KIND>));
    if (resident <= 0) {
        set_error("conv_kernel: occupancy query failed");
        return kHipError;
    }
    const int64_t ntiles = (int64_t)((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY) * a.N;
    const int grid = (int)std::min<int64_t>(ntiles, resident);
    hipLaunchKernelGGL(conv_kernel<KIND>, dim3(grid), dim3(256), 0, s, a);
    BF_HIP(hipGetLastError());
    return kOk;
}

// Band height of the fused kernel.  A band of B rows costs ceil(B / 2) steps
// plus -FS_KBEGIN warm-up steps (and a few steps' worth of set-up), and the
// grid runs in rounds of one workgroup per CU, so the launch takes about
// ceil(workgroups / CUs) x that.  Images whose 256-row bands fill the
// device get the band height that minimises this; smaller ones keep
// 256-row bands (more workgroups for a device they do not fill, and less
// warm-up work when several images share the device).
int fused_band(int H, int strips)
{
    const int base = std::min(H, 256);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return base;
    auto makespan = [&](int B) {
        const int64_t bands = (H + B - 1) / B;
        const int64_t rounds = ((int64_t)strips * bands + cus - 1) / cus;
        return rounds * ((B + 1) / 2 - FS_KBEGIN + 3);
    };
    if ((int64_t)strips * ((H + 255) / 256) < cus) return base;
    int best = base;
    int64_t best_cost = makespan(base);
    for (int B = 258; B <= H; B += 2) {
        const int64_t c = makespan(B);
        if (c < best_cost) { best_cost = c; best = B; }
    }
    return best;
}

// Fused inference of one image (u8 in, u8 out): strips of FS_W columns x
// bands of fused_band rows, one workgroup of FS_WAVES waves each.
int launch_fused(Net *net, const uint8_t *img, int64_t img_pitch, uint8_t *out, int64_t out_pitch,
                 int H, int W, float sigma_range, hipStream_t s)
{
    static const hipError_t configured = hipFuncSetAttribute(
        reinterpret_cast<const void *>(fused_infer_kernel), hipFuncAttributeMaxDynamicSharedMemorySize, FS_LDS);
    BF_HIP(configured);
    FusedArgs a{};
    a.img = img;
    a.img_pitch = img_pitch;
    a.out = out;
    a.out_pitch = out_pitch;
    a.H = H;
    a.W = W;
    a.strips = (W + FS_W - 1) / FS_W;
    a.band = fused_band(H, a.strips);
    a.sr = sigma_range / 100.0f;
    for (int l = 0; l < NLAYERS; l++) {
        a.fw[l] = net->fw[l];
        a.bias[l] = net->params + poff(l) + wcount(l);
    }
    const int64_t grid = (int64_t)a.strips * ((H + a.band - 1) / a.band);
    if (grid > 0x7FFFFFFF) {
        set_error("image too large");
        return kInvalidData;
    }
    hipLaunchKernelGGL(fused_infer_kernel, dim3((unsigned)grid), dim3(FS_WAVES * 64), FS_LDS, s, a);
    BF_HIP(hipGetLastError());
    return kOk;
}

// Forward pass over N images of H x W: img (u8, pitch img_pitch, image
// stride img_stride), sigma (device, N).  Activations into net->act; the
// prediction into pred (f32 dense) and/or out_u8.
int forward(Net *net, const uint8_t *img, int64_t img_pitch, int64_t img_stride,
            const float *sigma, float sigma_const, int N, int H, int W, float *pred,
            uint8_t *out_u8, int64_t u8_pitch, hipStream_t s)
{
    ConvArgs a{};
    a.sigma_const = sigma_const;
    a.N = N;
    a.H = H;
    a.W = W;
    a.img = img;
    a.img_pitch = img_pitch;
    a.img_stride = img_stride;
    a.sigma = sigma;
    a.wpk = net->fw[0];
    a.bias = net->params + poff(0) + wcount(0);
    a.out = net->act[0];
    int rc = launch_conv<K_FIRST>(a, s);
    for (int l = 1; l < NLAYERS - 1 && rc == kOk; l++) {
        a.in = net->act[l - 1];
        a.wpk = net->fw[l];
        a.bias = net->params + poff(l) + wcount(l);
        a.out = net->act[l];
        rc = launch_conv<K_HIDDEN>(a, s);
    }
    if (rc) return rc;
    a.in = net->act[NLAYERS - 2];
    a.wpk = net->fw[NLAYERS - 1];
    a.bias = net->params + poff(NLAYERS - 1) + wcount(NLAYERS - 1);
    a.out = nullptr;
    a.pred = pred;
    a.out_u8 = out_u8;
    a.u8_pitch = u8_pitch;
    return launch_conv<K_LAST>(a, s);
}

template <int KIND>
int launch_wgrad(Net *net, const WgradArgs &w0, int layer, hipStream_t s)
{
    WgradArgs w = w0;
    w.partial = net->partial;
    const int64_t nseg = (int64_t)w.N * w.H * ((w.W + TX - 1) / TX);
    static const int resident = resident_blocks(
        KIND == K_FIRST ? reinterpret_cast<const void *>(&wgrad_kernel<K_FIRST>)
                        : reinterpret_cast<const void *>(&wgrad32_kernel<KIND == K_FIRST ? K_HIDDEN : KIND>));
    if (resident <= 0) {
        set_error("wgrad kernel: occupancy query failed");
        return kHipError;
    }
    const int blocks = (int)std::min<int64_t>(std::min(net->partial_blocks, resident), nseg);
    if constexpr (KIND == K_FIRST)
        hipLaunchKernelGGL(wgrad_kernel<KIND>, dim3(blocks), dim3(256), 0, s, w);
    else
        hipLaunchKernelGGL(wgrad32_kernel<KIND>, dim3(blocks), dim3(256), 0, s, w);
    BF_HIP(hipGetLastError());
    const int psz = pcount(layer);
    const int runs = (blocks + kReduceRun - 1) / kReduceRun;
    hipLaunchKernelGGL(reduce_partial_runs, dim3((psz + 255) / 256, runs), dim3(256), 0, s,
                       net->partial, blocks, psz);
    BF_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_partials, dim3((psz + 255) / 256), dim3(256), 0, s, net->partial,
                       blocks, psz, net->grad + poff(layer));
    BF_HIP(hipGetLastError());
    return kOk;
}

// Backward pass after forward(): gradients of the masked MSE (loss_kernel)
// into net->grad.  loss (device float) is accumulated (caller zeroes it).
int backward(Net *net, const uint8_t *img, const uint8_t *target, const uint8_t *edges,
             const float *sigma, int N, int H, int W, float *loss, hipStream_t s)
{
    // 512 workgroups: each adds its partial sum to the one loss float, and
    // those atomics serialise
    const int lb = (int)std::min<int64_t>(512, (int64_t)N * H);
    hipLaunchKernelGGL(loss_kernel, dim3(lb), dim3(256), 0, s, net->pred, target, edges, N, H, W,
                       net->tail_w, net->tail_t, net->g6, loss);
    BF_HIP(hipGetLastError());
    int rc;
    // layer 6
    WgradArgs w{};
    w.N = N;
    w.H = H;
    w.W = W;
    w.act = net->act[4];
    w.dpre = net->g6;
    if ((rc = launch_wgrad<K_LAST>(net, w, 5, s))) return rc;
    ConvArgs a{};
    a.N = N;
    a.H = H;
    a.W = W;
    a.in = net->g6;
    a.wpk = net->bw[5];
    a.mask = net->act[4];
    a.out = net->dbuf[0];
    if ((rc = launch_conv<K_DG_LAST>(a, s))) return rc;
    // layers 5..2: dbuf[cur] holds d(pre-activation) of layer l
    int cur = 0;
    for (int l = NLAYERS - 2; l >= 1; l--) {
        w.act = net->act[l - 1];
        w.dpre = net->dbuf[cur];
        if ((rc = launch_wgrad<K_HIDDEN>(net, w, l, s))) return rc;
        a.in = net->dbuf[cur];
        a.wpk = net->bw[l];
        a.mask = net->act[l - 1];
        a.out = net->dbuf[cur ^ 1];
        if ((rc = launch_conv<K_DG_HIDDEN>(a, s))) return rc;
        cur ^= 1;
    }
    // layer 1
    w.act = nullptr;
    w.img = img;
    w.sigma = sigma;
    w.dpre = net->dbuf[cur];
    return launch_wgrad<K_FIRST>(net, w, 0, s);
}

void free_net(Net *net)
{
    (void)hipFree(net->params);
    (void)hipFree(net->m);
    (void)hipFree(net->v);
    (void)hipFree(net->grad);
    for (auto p : net->fw) (void)hipFree(p);
    for (auto p : net->bw) (void)hipFree(p);
    for (auto p : net->act) (void)hipFree(p);
    for (auto p : net->dbuf) (void)hipFree(p);
    (void)hipFree(net->g6);
    (void)hipFree(net->pred);
    (void)hipFree(net->partial);
    (void)hipFree(net->list);
    (void)hipFree(net->count);
    delete net;
}

}  // namespace sur
}  // namespace bfimpl

using namespace bfimpl;
using namespace bfimpl::sur;

extern "C" {

int64_t bf_sur_param_count(void) { return kParams; }

int bf_sur_create(uint64_t seed, bf_sur *out)
{
    if (!out) return kNullPointer;
    Net *net = new (std::nothrow) Net();
    if (!net) return kOutOfMem;
    net->device = current_device();
    // He-normal weights from splitmix64 + Box-Muller; zero biases; the last
    // layer starts at 1/10 scale so the initial output is close to the input.
    std::vector<float> h(kParams, 0.0f);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        const double u1 = ((double)(next() >> 11) + 1.0) * 0x1.0p-53;
        const double u2 = (double)(next() >> 11) * 0x1.0p-53;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int l = 0; l < NLAYERS; l++) {
        const double sd = std::sqrt(2.0 / (9.0 * cin_of(l))) * (l == NLAYERS - 1 ? 0.1 : 1.0);
        for (int i = 0; i < wcount(l); i++) h[poff(l) + i] = (float)(sd * normal());
    }
    hipError_t e = hipSuccess;
    auto alloc = [&](auto **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    };
    alloc(&net->params, kParams * 4);
    alloc(&net->m, kParams * 4);
    alloc(&net->v, kParams * 4);
    alloc(&net->grad, kParams * 4);
    for (int l = 0; l < NLAYERS; l++) {
        alloc(&net->fw[l], fw_size(l) * 2);
        alloc(&net->bw[l], bw_size(l) * 2);
    }
    net->partial_blocks = 1024;
    alloc(&net->partial, (size_t)net->partial_blocks * pcount(1) * 4);
    if (e == hipSuccess) e = hipMemcpy(net->params, h.data(), kParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(net->m, 0, kParams * 4);
    if (e == hipSuccess) e = hipMemset(net->v, 0, kParams * 4);
    if (e == hipSuccess) e = hipMemset(net->grad, 0, kParams * 4);
    if (e != hipSuccess) {
        free_net(net);
        return hip_fail(e, "bf_sur_create");
    }
    int rc = pack(net, nullptr);
    if (rc == kOk && hipDeviceSynchronize() != hipSuccess) rc = kHipError;
    if (rc) {
        free_net(net);
        return rc;
    }
    *out = reinterpret_cast<bf_sur>(net);
    return kOk;
}

int bf_sur_destroy(bf_sur h)
{
    if (!h) return kNullPointer;
    free_net(net_of(h));
    return kOk;
}

int bf_sur_get_params(bf_sur h, float *params)
{
    if (!h || !params) return kNullPointer;
    BF_HIP(hipMemcpy(params, net_of(h)->params, kParams * 4, hipMemcpyDeviceToHost));
    return kOk;
}

int bf_sur_set_params(bf_sur h, const float *params)
{
    if (!h || !params) return kNullPointer;
    Net *net = net_of(h);
    BF_HIP(hipMemcpy(net->params, params, kParams * 4, hipMemcpyHostToDevice));
    int rc = pack(net, nullptr);
    if (rc) return rc;
    BF_HIP(hipDeviceSynchronize());
    return kOk;
}

int bf_sur_set_loss_tail(bf_sur h, float weight, float start_lsb)
{
    if (!h) return kNullPointer;
    if (!(weight >= 0.0f && weight < 1e30f) || !(start_lsb >= 0.0f && start_lsb < 256.0f)) {
        set_error("tail weight must be finite and >= 0, tail start in [0, 256) LSB");
        return kInvalidData;
    }
    Net *net = net_of(h);
    net->tail_w = weight;
    net->tail_t = start_lsb / 255.0f;
    return kOk;
}

static int check_batch(const void *x, const float *sigma, int n, int h, int w)
{
    if (!x || !sigma) {
        set_error("null argument");
        return kNullPointer;
    }
    if (n <= 0 || h <= 2 * MARGIN || w <= 2 * MARGIN) {
        set_error("crops must be larger than 12 x 12");
        return kInvalidData;
    }
    return kOk;
}

int bf_sur_forward_backward(bf_sur hnd, const unsigned char *x, const unsigned char *y,
                            const float *sigma, const unsigned char *edges, int32_t n, int32_t h,
                            int32_t w, float *pred, float *grads, float *loss_out, void *stream)
{
    if (!hnd) return kNullPointer;
    int rc = check_batch(x, sigma, n, h, w);
    if (rc) return rc;
    if (grads && !y) {
        set_error("gradients need a target");
        return kNullPointer;
    }
    Net *net = net_of(hnd);
    hipStream_t s = (hipStream_t)stream;
    const size_t pix = (size_t)n * h * w;
    if ((rc = ensure_ws(net, pix))) return rc;
    if ((rc = forward(net, x, w, (int64_t)h * w, sigma, 0.0f, n, h, w, net->pred, nullptr, 0, s)))
        return rc;
    if (pred) BF_HIP(hipMemcpyAsync(pred, net->pred, pix * 4, hipMemcpyDeviceToDevice, s));
    if (loss_out) {
        hipLaunchKernelGGL(zero_floats, dim3(1), dim3(64), 0, s, loss_out, 1);
        BF_HIP(hipGetLastError());
    }
    if (y) {
        if ((rc = backward(net, x, y, edges, sigma, n, h, w, loss_out, s))) return rc;
        if (grads) BF_HIP(hipMemcpyAsync(grads, net->grad, kParams * 4, hipMemcpyDeviceToDevice, s));
    }
    return kOk;
}

int bf_sur_train_step(bf_sur hnd, const unsigned char *x, const unsigned char *y, const float *sigma,
                      const unsigned char *edges, int32_t n, int32_t h, int32_t w, float lr,
                      float beta1, float beta2, float eps, float *loss_out, void *stream)
{
    if (!hnd || !y) return kNullPointer;
    int rc = check_batch(x, sigma, n, h, w);
    if (rc) return rc;
    Net *net = net_of(hnd);
    hipStream_t s = (hipStream_t)stream;
    if ((rc = ensure_ws(net, (size_t)n * h * w))) return rc;
    if ((rc = forward(net, x, w, (int64_t)h * w, sigma, 0.0f, n, h, w, net->pred, nullptr, 0, s)))
        return rc;
    if (loss_out) {
        hipLaunchKernelGGL(zero_floats, dim3(1), dim3(64), 0, s, loss_out, 1);
        BF_HIP(hipGetLastError());
    }
    if ((rc = backward(net, x, y, edges, sigma, n, h, w, loss_out, s))) return rc;
    net->step++;
    const float c1 = 1.0f - std::pow(beta1, (float)net->step);
    const float c2 = 1.0f - std::pow(beta2, (float)net->step);
    hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, net->params,
                       net->m, net->v, net->grad, kParams, lr, beta1, beta2, eps, c1, c2);
    BF_HIP(hipGetLastError());
    return pack(net, s);
}

static int infer_view(Net *net, const bf_image_view *in, const bf_image_view *out, float sigma_range,
                      hipStream_t s, int *Wo, int *Ho)
{
    if (!in || !out || !in->data || !out->data) {
        set_error("null view");
        return kNullPointer;
    }
    const int W = std::min(in->width, out->width), H = std::min(in->height, out->height);
    if (W <= 0 || H <= 0 || in->pitch < in->width || out->pitch < out->width) {
        set_error("empty view or pitch smaller than width");
        return kInvalidData;
    }
    // The fused kernel reads a halo around every strip that neighbouring
    // workgroups write, and the check and the fallback read the input again
    // after inference: the input must survive the whole call.
    if (views_overlap(in, out)) {
        set_error("input and output views overlap");
        return kInvalidData;
    }
    *Wo = W;
    *Ho = H;
    return launch_fused(net, in->data, in->pitch, out->data, out->pitch, H, W, sigma_range, s);
}

int bf_sur_infer(bf_sur hnd, const bf_image_view *in, const bf_image_view *out, float sigma_range,
                 void *stream)
{
    if (!hnd) return kNullPointer;
    int W, H;
    return infer_view(net_of(hnd), in, out, sigma_range, (hipStream_t)stream, &W, &H);
}

// Where one call's fallback tile list and count live: the network's own
// workspace (grown with hipMalloc as images grow), or caller-provided
// device memory of `cap` tiles plus the count.
struct FallbackWs {
    int32_t *list = nullptr;
    int32_t *count = nullptr;
    int cap = 0;
};

static int sur_filter(Net *net, const bf_image_view *in, const bf_image_view *out, bf_lut lut,
                      int radius, int mode, float sigma_range, int tol_lsb, int32_t *accum,
                      int32_t *tiles_accum, const FallbackWs *given, hipStream_t s)
{
    // exact job for the check and the fallback (64 x 64 tiles = config 0)
    DevJob job;
    int rc = make_job(in, out, lut_from_handle(lut), radius, mode, 0, halo_class(radius), 0, &job,
                      nullptr);
    if (rc) return rc;
    const int W = std::min(in->width, out->width), H = std::min(in->height, out->height);
    const int tx = (W + 63) / 64, ty = (H + 63) / 64, nt = tx * ty;
    FallbackWs ws;
    if (given) {
        if (nt > given->cap) {
            set_error("the image has " + std::to_string(nt) + " 64x64 tiles; the workspace holds " +
                      std::to_string(given->cap));
            return kInvalidData;
        }
        ws = *given;
    }
    int Wi, Hi;
    if ((rc = infer_view(net, in, out, sigma_range, s, &Wi, &Hi))) return rc;
    if (!given) {
        if (nt > net->fb_tiles) {
            (void)hipFree(net->list);
            net->list = nullptr;
            net->fb_tiles = 0;
            BF_HIP(hipMalloc(&net->list, nt * 4));
            net->fb_tiles = nt;
        }
        if (!net->count) BF_HIP(hipMalloc(&net->count, 4));
        ws.list = net->list;
        ws.count = net->count;
        ws.cap = net->fb_tiles;
    }
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, ws.count, 1);
    BF_HIP(hipGetLastError());
    // The canonical template at r = 9 or 15 runs the unrolled 64x64-tile
    // kernels for both the check and the fallback (same taps, order and
    // rounding as the run-time tap list, so the same verdicts and pixels).
    const bool ring = ring64_job(lut_from_handle(lut), radius, mode, &job);
    CheckArgs ca{};
    ca.job = job;
    ca.rowg = lut_from_handle(lut)->d_rowg;
    ca.tiles_x = tx;
    ca.tiles_y = ty;
    ca.tol = tol_lsb;
    ca.list = ws.list;
    ca.count = ws.count;
    ca.accum = accum;
    ca.tiles_accum = tiles_accum;
    ca.salt = 0x5bd1e995u;
    // make_job has checked every tap against this halo (REF rows may reach
    // two further down, which the 4-row grouping of the sample row covers).
    ca.halo = halo_class(radius);
    if ((rc = launch_check(ca, ring ? radius : 0, nt, s))) return rc;
    // exact fallback over the listed tiles
    if (ring) return ring64_tiles(job, radius, mode, ws.list, ws.count, nt, s);
    return exact_tiles(job, ws.list, nt, ws.count, 0, halo_class(radius), s);
}

// ------------------------------------------------- allocation-free variant
//
// bf_sur_filter_noalloc keeps its fallback tile count and list in this
// file-scope device array instead of the network's workspace: word 0 is the
// count, words 64.. the list.  Every device the library runs on holds its
// own copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (the array's name is a
// host-side stand-in, not a device address).
constexpr int kScratchTiles = 65536;      // 64x64 tiles: up to 16384 x 16384 pixels
constexpr int kScratchWords = 64 + kScratchTiles;
static __device__ __align__(256) int32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static int32_t *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per
// device, with that device current), or nullptr with the error set.
static int32_t *scratch_on_current_device()
{
    const int dev = current_device();
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("no current device, or device id beyond the scratch table");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            hip_fail(e, "hipGetSymbolAddress(scratch)");
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<int32_t *>(ptr);
    }
    return g_scratch_ptr[dev];
}

int bf_sur_filter(bf_sur hnd, const bf_image_view *in, const bf_image_view *out, bf_lut lut,
                  int32_t radius, int32_t mode, float sigma_range, int32_t tol_lsb,
                  int32_t *tiles_total, int32_t *tiles_fallback, void *stream)
{
    if (!hnd || !lut) return kNullPointer;
    Net *net = net_of(hnd);
    hipStream_t s = (hipStream_t)stream;
    int rc = sur_filter(net, in, out, lut, radius, mode, sigma_range, tol_lsb, nullptr, nullptr,
                        nullptr, s);
    if (rc) return rc;
    const int W = std::min(in->width, out->width), H = std::min(in->height, out->height);
    if (tiles_total) *tiles_total = ((W + 63) / 64) * ((H + 63) / 64);
    if (tiles_fallback) {
        int32_t c = 0;
        BF_HIP(hipMemcpyAsync(&c, net->count, 4, hipMemcpyDeviceToHost, s));
        BF_HIP(hipStreamSynchronize(s));
        *tiles_fallback = c;
    }
    return kOk;
}

int bf_sur_filter_async(bf_sur hnd, const bf_image_view *in, const bf_image_view *out, bf_lut lut,
                        int32_t radius, int32_t mode, float sigma_range, int32_t tol_lsb,
                        int32_t *fallback_accum, int32_t *tiles_accum, void *stream)
{
    if (!hnd || !lut) return kNullPointer;
    return sur_filter(net_of(hnd), in, out, lut, radius, mode, sigma_range, tol_lsb,
                      fallback_accum, tiles_accum, nullptr, (hipStream_t)stream);
}

int bf_sur_filter_noalloc(bf_sur hnd, const bf_image_view *in, const bf_image_view *out,
                          bf_lut lut, int32_t radius, int32_t mode, float sigma_range,
                          int32_t tol_lsb, int32_t *fallback_accum, int32_t *tiles_accum,
                          void *stream)
{
    if (!hnd || !lut) return kNullPointer;
    int32_t *base = scratch_on_current_device();
    if (!base) return kHipError;
    FallbackWs ws;
    ws.count = base;
    ws.list = base + 64;
    ws.cap = kScratchTiles;
    return sur_filter(net_of(hnd), in, out, lut, radius, mode, sigma_range, tol_lsb,
                      fallback_accum, tiles_accum, &ws, (hipStream_t)stream);
}

}  // extern "C"
