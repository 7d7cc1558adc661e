// bf_ring_batch.hip -- batches of jobs on the canonical (GAUSS-factory)
// template, TemplateFactory.cu:1092-1168: the persistent work-queue launch
// of bf_exact_kernel, with each tile filtered by the unrolled kernel of its
// job's radius and mode (tap offsets and table rows compiled in) instead of
// the run-time tap list.  Results are the same bits; the batch only saves
// the tap-list decoding.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bf_internal.h"
#include "bf_tile.h"

namespace bfimpl {

namespace {

// One unrolled radius: instantiated only where the staged halo covers it.
template <int R, int MODE, int TW, int TGY, int HALO>
__device__ __forceinline__ void ring_if_fits(const DevJob &job, int X0, int Y0, int lx, int gy)
{
    if constexpr (R <= HALO) tile_ring<R, MODE, TW, TGY, HALO>(job, X0, Y0, lx, gy);
}

template <int TW, int TGY, int HALO>
__global__ void __launch_bounds__(TW * TGY) bf_ring_batch_kernel(KArgs args)
{
    using G = Geo<TW, TGY, HALO>;
    const int tid = threadIdx.x;
    const int lx = tid % TW, gy = tid / TW;
    const BF_CONST DevJob *jobs = as_const(args.jobs);
    int cur_key = -1;
    for (;;) {
        __syncthreads();                           // everyone done with the slot
        if (tid == 0) lds_store_u32(0, (uint32_t)atomicAdd(args.counter, 1));
        __syncthreads();
        const int tile = __builtin_amdgcn_readfirstlane((int)lds_u32(0));
        if (tile >= args.ntiles) break;
        DevJob job;
        {
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
        switch (job.radius * 2 + (job.mode == kModeRef ? 0 : 1)) {
        case 6: ring_if_fits<3, 0, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 7: ring_if_fits<3, 1, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 10: ring_if_fits<5, 0, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 11: ring_if_fits<5, 1, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 14: ring_if_fits<7, 0, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 15: ring_if_fits<7, 1, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 18: ring_if_fits<9, 0, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 19: ring_if_fits<9, 1, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 30: ring_if_fits<15, 0, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        case 31: ring_if_fits<15, 1, TW, TGY, HALO>(job, X0, Y0, lx, gy); break;
        default: break;                            // excluded by the host (ring_job)
        }
    }
}

template <int TW, int TGY, int HALO>
int launch_one(const KArgs &a, int max_lut_rows, hipStream_t s)
{
    using G = Geo<TW, TGY, HALO>;
    const size_t lds = kTileOff + (size_t)G::TILE_BYTES + (size_t)max_lut_rows * 1024;
    constexpr auto kern = bf_ring_batch_kernel<TW, TGY, HALO>;
    BF_HIP(allow_full_lds<kern>());
    const int res = resident_groups(reinterpret_cast<const void *>(kern), G::NT, lds);
    if (res < 0) return hip_fail(hipGetLastError(), "occupancy query");
    if (res < 1) {
        set_error("ring batch kernel does not fit on a CU (LDS " + std::to_string(lds) + " bytes)");
        return kInvalidData;
    }
    if (a.ntiles < 1) return kOk;
    hipLaunchKernelGGL(kern, dim3(res), dim3(G::NT), lds, s, a);
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
    default: set_error("no unrolled batch kernel for this halo"); return kInvalidData;
    }
}

}  // namespace

bool ring_batch_supported(int config, int halo)
{
    return (config == 0 || config == 1) && (halo == 4 || halo == 8 || halo == 12 || halo == 16);
}

int launch_ring_batch(int config, int halo, const DevJob *jobs_dev, int njobs, int ntiles,
                      int max_lut_rows, int *counter, hipStream_t s)
{
    if (!jobs_dev || !counter || njobs < 1) {
        set_error("ring batch needs a device job table and a queue counter");
        return kInvalidData;
    }
    KArgs a{};
    a.jobs = jobs_dev;
    a.njobs = njobs;
    a.ntiles = ntiles;
    a.counter = counter;
    switch (config) {
    case 0: return launch_halo<64, 16>(halo, a, max_lut_rows, s);
    case 1: return launch_halo<64, 8>(halo, a, max_lut_rows, s);
    default: set_error("no unrolled batch kernel for this tile config"); return kInvalidData;
    }
}

}  // namespace bfimpl
