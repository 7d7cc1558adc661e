// bf_abi.cpp -- C ABI of the exact and batched filter (include/opencuda_bf.h).
//
// Host-side bookkeeping: validation with the reference's status codes,
// device tables, tile planning, launches.  Kernels are in bf_exact.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bf_internal.h"
#include "opencuda_bf.h"

namespace bfimpl {

static thread_local std::string g_last_error;

// Table ids: unique among live tables and below 2^19, so the 19-bit id part
// of a job's LDS key (make_job) never names two live tables.  Destroyed ids
// are reused.
constexpr int kMaxLutId = 0x7FFFF;
struct LutIds {
    std::mutex mu;
    std::vector<int> free_ids;
    int next = 1;
};

static LutIds &lut_ids()
{
    static LutIds *ids = new LutIds();
    return *ids;
}

static int acquire_lut_id()
{
    LutIds &ids = lut_ids();
    std::lock_guard<std::mutex> lock(ids.mu);
    if (!ids.free_ids.empty()) {
        const int id = ids.free_ids.back();
        ids.free_ids.pop_back();
        return id;
    }
    return ids.next <= kMaxLutId ? ids.next++ : -1;
}

static void release_lut_id(int id)
{
    LutIds &ids = lut_ids();
    std::lock_guard<std::mutex> lock(ids.mu);
    ids.free_ids.push_back(id);
}

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return kHipError;
}

int current_device()
{
    int d = -1;
    if (hipGetDevice(&d) != hipSuccess) return -1;
    return d;
}

// Per-(lut, ntaps) device tap arrays live in one buffer: entry k of the
// variant with n real taps sits at variant_offset(n) + k.
struct LutVariants {
    std::unordered_map<int, int> offset_of;   // ntaps -> first entry
};

struct LutFull : Lut {
    LutVariants var;
};

Lut *lut_from_handle(bf_lut h) { return reinterpret_cast<Lut *>(h); }

// Autotuner results: measured ms per Mpix of each tile configuration for
// (mode, radius); an entry is "tuned" when at least one cost is >= 0.
struct TuneTable {
    std::mutex mu;
    bool have[2][BF_MAX_RADIUS + 1] = {};
    float cost[2][BF_MAX_RADIUS + 1][16] = {};
};

static TuneTable &tune_table()
{
    static TuneTable *t = new TuneTable();
    return *t;
}

// Tuned configuration for (radius, mode), or -1.
static int tuned_config(int radius, int mode)
{
    if (radius < 1 || radius > BF_MAX_RADIUS || (mode != 0 && mode != 1)) return -1;
    TuneTable &t = tune_table();
    std::lock_guard<std::mutex> lock(t.mu);
    if (!t.have[mode][radius]) return -1;
    int best = -1;
    for (int c = 0; c < tile_config_count(); c++)
        if (t.cost[mode][radius][c] >= 0 && (best < 0 || t.cost[mode][radius][c] < t.cost[mode][radius][best]))
            best = c;
    return best;
}

static int ntaps_for(int radius, int mode)
{
    const int g = (2 * radius + 1) * (2 * radius + 1);
    return mode == BF_MODE_REF ? g + 1 : g;
}

static int padded(int n) { return (n + kTapPad - 1) / kTapPad * kTapPad; }

int make_job(const bf_image_view *in, const bf_image_view *out, const Lut *lut_base, int radius,
             int mode, int config, int halo, int tile_begin, DevJob *job, int64_t *pixels)
{
    if (!in || !out || !lut_base || !job) {
        set_error("null argument");
        return kNullPointer;
    }
    const LutFull *lut = static_cast<const LutFull *>(lut_base);
    if (mode != BF_MODE_REF && mode != BF_MODE_STD) {
        set_error("mode must be BF_MODE_REF or BF_MODE_STD");
        return kInvalidData;
    }
    if (radius <= 0 || radius > BF_MAX_RADIUS) {
        set_error("radius must be in 1..20");
        return kInvalidData;
    }
    if (!in->data || !out->data) {
        set_error("image view has no data");
        return kNullPointer;
    }
    if (in->width <= 0 || in->height <= 0 || out->width <= 0 || out->height <= 0 ||
        in->pitch < in->width || out->pitch < out->width) {
        set_error("empty image view or pitch smaller than width");
        return kInvalidData;
    }
    const int nt = ntaps_for(radius, mode);
    if (nt > lut->count) {
        set_error("template holds " + std::to_string(lut->count) + " points, radius " +
                  std::to_string(radius) + " needs " + std::to_string(nt) +
                  (mode == BF_MODE_REF ? " (REF reads (2r+1)^2+1 taps)" : ""));
        return kInvalidData;
    }
    auto it = lut->var.offset_of.find(nt);
    if (it == lut->var.offset_of.end()) {
        set_error("no tap variant for this radius");
        return kInvalidData;
    }
    if (halo < halo_class(radius)) {
        set_error("halo smaller than the radius needs");
        return kInvalidData;
    }
    // Every tap must stay inside the staged halo.
    for (int c = 0; c < nt; c++) {
        const int dy = lut->offsets[2 * c], dx = lut->offsets[2 * c + 1];
        if (dx < -halo || dx > halo || dy < -halo || dy > halo + (mode == BF_MODE_REF ? 2 : 0) ||
            (mode == BF_MODE_STD && (dy < -radius || dy > radius || dx < -radius || dx > radius))) {
            set_error("template tap outside the radius/halo the kernel stages");
            return kInvalidData;
        }
    }
    const TileConfig &tc = tile_config(config);
    const int nrows = lut->rows_upto[nt];
    const bool sep = lds_bytes(config, halo, nrows + 1) > max_lds_bytes();
    if (lds_bytes(config, halo, sep ? 1 : nrows + 1) > max_lds_bytes()) {
        set_error("tile config " + std::to_string(config) + " does not fit in LDS");
        return kInvalidData;
    }
    const int W = std::min(in->width, out->width);
    const int H = std::min(in->height, out->height);
    std::memset(job, 0, sizeof(*job));
    job->in = in->data;
    job->out = out->data;
    job->in_pitch = in->pitch;
    job->out_pitch = out->pitch;
    job->in_w = in->width;
    job->in_h = in->height;
    job->W = W;
    job->H = H;
    job->taps = lut->d_taps + it->second;
    job->gtap = lut->d_gtap + it->second;
    job->rowg = lut->d_rowg;
    job->range = lut->d_range;
    job->ntaps_padded = padded(nt);
    job->nrows = nrows;
    job->sep = sep ? 1 : 0;
    job->lut_rows = sep ? 1 : nrows + 1;
    job->mode = mode;
    // LDS table identity: the same tables with another row count (REF has
    // the extra tap's row, and the zero row moves) or mode is another table.
    // (A fused table that fits in LDS has < 160 rows; a separate table is E only.)
    job->lut_key = (lut->id << 12) | ((sep ? 0 : nrows) << 1) | (sep ? 1 : 0);
    job->tiles_x = (W + tc.tw - 1) / tc.tw;
    job->tile_begin = tile_begin;
    job->radius = radius;
    if (pixels) *pixels = (int64_t)W * H;
    return kOk;
}

int fused_rows(const Lut *lut, int radius, int mode)
{
    if (radius < 1 || radius > BF_MAX_RADIUS) return -1;
    const int nt = ntaps_for(radius, mode);
    return nt <= lut->count ? lut->rows_upto[nt] : -1;
}

static int tiles_of(const DevJob &j, int config)
{
    const TileConfig &tc = tile_config(config);
    return j.tiles_x * ((j.H + 4 * tc.tgy - 1) / (4 * tc.tgy));
}

// Switch a job to the canonical per-q table of the unrolled kernels when
// the template is the canonical ring, it holds the radius's taps and the
// table fits next to a `config` tile with `halo`.
bool ring_job(const Lut *L, int radius, int mode, int config, int halo, DevJob *job)
{
    const int rr = ring_rows(radius, mode);
    if (!L->canonical || rr <= 0 || ntaps_for(radius, mode) > L->count || job->sep ||
        halo < halo_class(radius) || lds_bytes(config, halo, rr + 1) > max_lds_bytes())
        return false;
    job->rowg = L->d_rowq;
    job->nrows = rr;
    job->lut_rows = rr + 1;
    // LDS key as make_job forms it: with one weight per q the per-q table
    // equals the per-weight one whenever their row counts agree
    job->lut_key = (L->id << 12) | (rr << 1);
    return true;
}

bool ring64_job(const Lut *L, int radius, int mode, DevJob *job)
{
    return ring64_available(radius) && ring_job(L, radius, mode, 0, halo_class(radius), job);
}

bool views_overlap(const bf_image_view *a, const bf_image_view *b)
{
    const uint8_t *a0 = a->data, *a1 = a->data + (int64_t)(a->height - 1) * a->pitch + a->width;
    const uint8_t *b0 = b->data, *b1 = b->data + (int64_t)(b->height - 1) * b->pitch + b->width;
    return a0 < b1 && b0 < a1;
}

}  // namespace bfimpl

using namespace bfimpl;

extern "C" {

const char *bf_last_error(void) { return g_last_error.c_str(); }

int bf_version(void) { return 10000; }

int bf_lut_create(const int32_t *offsets, const float *weights, int32_t count, const float *range,
                  bf_lut *out)
{
    if (!offsets || !weights || !range || !out) {
        set_error("null argument");
        return kNullPointer;
    }
    if (count <= 0 || count > 65535) {
        set_error("template point count must be 1..65535");
        return kInvalidData;
    }
    for (int c = 0; c < count; c++) {
        if (offsets[2 * c] < -127 || offsets[2 * c] > 127 || offsets[2 * c + 1] < -127 ||
            offsets[2 * c + 1] > 127) {
            set_error("template offset out of int8 range");
            return kInvalidData;
        }
    }
    auto *L = new (std::nothrow) LutFull();
    if (!L) return kOutOfMem;
    L->id = acquire_lut_id();
    if (L->id < 0) {
        delete L;
        set_error("too many live filter tables (2^19 - 1)");
        return kOutOfMem;
    }
    L->count = count;
    L->offsets.assign(offsets, offsets + 2 * (size_t)count);
    L->weights.assign(weights, weights + count);
    L->range.assign(range, range + 256);
    // Rows of the fused table: distinct weight bit patterns, in order of
    // first appearance, so the first n taps use rows [0, rows_upto[n]).
    std::unordered_map<uint32_t, int> row_of_bits;
    std::vector<float> rowg;
    L->row_of_tap.resize(count);
    L->rows_upto.assign(count + 1, 0);
    for (int c = 0; c < count; c++) {
        uint32_t bits;
        std::memcpy(&bits, &weights[c], 4);
        auto it = row_of_bits.find(bits);
        int row;
        if (it == row_of_bits.end()) {
            row = (int)rowg.size();
            row_of_bits.emplace(bits, row);
            rowg.push_back(weights[c]);
        } else {
            row = it->second;
        }
        L->row_of_tap[c] = row;
        L->rows_upto[c + 1] = (int)rowg.size();
    }
    L->nrows_total = (int)rowg.size();
    // Canonical ring template (the GAUSS factory's order, one weight per
    // q = dy^2 + dx^2)?  Then the unrolled kernels apply; they index the
    // fused table by q in order of first appearance.
    std::vector<float> rowq;
    {
        bool canon = true;
        std::unordered_map<int, int> qrow;
        std::vector<std::pair<int, int>> ring;
        ring.emplace_back(0, 0);
        for (int i = 1; (int)ring.size() < count; i++)
            for (int j = -i + 1; j <= i; j++) {
                ring.emplace_back(i, j);
                ring.emplace_back(-i, j);
                ring.emplace_back(j, i);
                ring.emplace_back(j, -i);
            }
        for (int c = 0; c < count && canon; c++) {
            const int dy = offsets[2 * c], dx = offsets[2 * c + 1];
            if (dy != ring[c].first || dx != ring[c].second) {
                canon = false;
                break;
            }
            const int q = dy * dy + dx * dx;
            auto it = qrow.find(q);
            if (it == qrow.end()) {
                qrow.emplace(q, (int)rowq.size());
                rowq.push_back(weights[c]);
            } else if (std::memcmp(&rowq[it->second], &weights[c], 4) != 0) {
                canon = false;
            }
        }
        L->canonical = canon;
        if (!canon) rowq.clear();
    }
    // Tap variants for every radius and mode that fits in the template.
    std::vector<uint32_t> packed;
    std::vector<float> gtap;
    for (int mode = 0; mode < 2; mode++) {
        for (int r = 1; r <= BF_MAX_RADIUS; r++) {
            const int nt = ntaps_for(r, mode);
            if (nt > count || L->var.offset_of.count(nt)) continue;
            L->var.offset_of[nt] = (int)packed.size();
            const int nrows = L->rows_upto[nt];
            for (int c = 0; c < padded(nt); c++) {
                if (c < nt) {
                    packed.push_back(pack_tap(offsets[2 * c], offsets[2 * c + 1], L->row_of_tap[c]));
                    gtap.push_back(weights[c]);
                } else {
                    packed.push_back(pack_tap(0, 0, nrows));   // null tap: zero row / g = 0
                    gtap.push_back(0.0f);
                }
            }
        }
    }
    if (packed.empty()) {
        packed.push_back(0);
        gtap.push_back(0.0f);
    }
    L->device = current_device();
    hipError_t e = hipMalloc(&L->d_taps, packed.size() * 4);
    if (e == hipSuccess) e = hipMalloc(&L->d_gtap, gtap.size() * 4);
    if (e == hipSuccess) e = hipMalloc(&L->d_rowg, std::max<size_t>(1, rowg.size()) * 4);
    if (e == hipSuccess) e = hipMalloc(&L->d_range, 256 * 4);
    if (e == hipSuccess && !rowq.empty()) e = hipMalloc(&L->d_rowq, rowq.size() * 4);
    if (e == hipSuccess && !rowq.empty())
        e = hipMemcpy(L->d_rowq, rowq.data(), rowq.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_taps, packed.data(), packed.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_gtap, gtap.data(), gtap.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && !rowg.empty())
        e = hipMemcpy(L->d_rowg, rowg.data(), rowg.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->d_range, range, 256 * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(L->d_taps);
        hipFree(L->d_gtap);
        hipFree(L->d_rowg);
        hipFree(L->d_range);
        hipFree(L->d_rowq);
        release_lut_id(L->id);
        delete L;
        return hip_fail(e, "bf_lut_create");
    }
    *out = reinterpret_cast<bf_lut>(static_cast<Lut *>(L));
    return kOk;
}

int bf_lut_destroy(bf_lut h)
{
    if (!h) return kNullPointer;
    auto *L = static_cast<LutFull *>(lut_from_handle(h));
    hipFree(L->d_taps);
    hipFree(L->d_gtap);
    hipFree(L->d_rowq);
    hipFree(L->d_rowg);
    hipFree(L->d_range);
    release_lut_id(L->id);
    delete L;
    return kOk;
}

int bf_lut_rows(bf_lut h, int32_t ntaps)
{
    if (!h) return kNullPointer;
    Lut *L = lut_from_handle(h);
    if (ntaps < 0 || ntaps > L->count) return kInvalidData;
    return L->rows_upto[ntaps];
}

int bf_tune_count(void) { return tile_config_count(); }

int bf_tune_describe(int32_t config, int32_t *tile_w, int32_t *tile_h, int32_t *threads,
                     int32_t *cols_per_lane)
{
    if (config < 0 || config >= tile_config_count()) return kInvalidData;
    const TileConfig &c = tile_config(config);
    if (tile_w) *tile_w = c.tw;
    if (tile_h) *tile_h = 4 * c.tgy;
    if (threads) *threads = c.threads;
    if (cols_per_lane) *cols_per_lane = c.cpl;
    return kOk;
}

// Pick a tile configuration for (radius, fused-table rows): `want` if the
// fused table fits with it, else any configuration where it fits, else
// `want` in separate-table mode.
static int fit_config(int want, int radius, int nrows)
{
    const int halo = halo_class(radius);
    if (lds_bytes(want, halo, nrows + 1) <= max_lds_bytes()) return want;
    for (int c = 0; c < tile_config_count(); c++)
        if (lds_bytes(c, halo, nrows + 1) <= max_lds_bytes()) return c;
    return want;
}

int bf_exact_config(const bf_image_view *in, const bf_image_view *out, bf_lut lut, int32_t radius,
                    int32_t mode, int32_t config, void *stream)
{
    if (config < 0 || config >= tile_config_count()) {
        set_error("bad tile config");
        return kInvalidData;
    }
    if (!in || !out || !lut) {
        set_error("null argument");
        return kNullPointer;
    }
    if (in->data && out->data && in->width > 0 && out->width > 0 && in->height > 0 &&
        out->height > 0 && views_overlap(in, out)) {
        set_error("input and output views overlap");
        return kInvalidData;
    }
    DevJob job;
    int rc = make_job(in, out, lut_from_handle(lut), radius, mode, config, halo_class(radius), 0,
                      &job, nullptr);
    if (rc) return rc;
    return launch_exact(config, halo_class(radius), job, nullptr, 1, tiles_of(job, config),
                        job.lut_rows, nullptr, (hipStream_t)stream);
}

int bf_exact(const bf_image_view *in, const bf_image_view *out, bf_lut lut, int32_t radius,
             int32_t mode, void *stream)
{
    if (!lut) {
        set_error("null lut");
        return kNullPointer;
    }
    Lut *L = lut_from_handle(lut);
    int rcfg = -1;
    if (L->canonical && ring_available(radius, mode, &rcfg) && ring_rows(radius, mode) > 0 &&
        ntaps_for(radius, mode) <= L->count) {
        if (!in || !out) {
            set_error("null argument");
            return kNullPointer;
        }
        if (in->data && out->data && in->width > 0 && out->width > 0 && in->height > 0 &&
            out->height > 0 && views_overlap(in, out)) {
            set_error("input and output views overlap");
            return kInvalidData;
        }
        DevJob job;
        int rc = make_job(in, out, L, radius, mode, rcfg, halo_class(radius), 0, &job, nullptr);
        if (rc) return rc;
        if (ring_job(L, radius, mode, rcfg, halo_class(radius), &job))
            return launch_ring(job, radius, mode, tiles_of(job, rcfg), (hipStream_t)stream);
    }
    const int rows = fused_rows(lut_from_handle(lut), radius, mode);
    const int tuned = tuned_config(radius, mode);
    const int want = tuned >= 0 ? tuned : default_config(radius, mode);
    const int config = rows < 0 ? default_config(1, mode) : fit_config(want, radius, rows);
    return bf_exact_config(in, out, lut, radius, mode, config, stream);
}

int bf_tune_radius(int32_t radius, int32_t mode, int32_t width, int32_t height, int32_t reps,
                   int32_t *best_config, float *ms_per_mpix)
{
    if (radius < 1 || radius > BF_MAX_RADIUS || (mode != 0 && mode != 1) || width < 1 ||
        height < 1 || reps < 1) {
        set_error("bad tuning parameters");
        return kInvalidData;
    }
    const int nc = tile_config_count();
    std::vector<unsigned char> host((size_t)width * height);
    bf_synth_image(host.data(), width, height, std::max(1, width * height / 20000), 8,
                   std::max(9, std::min(width, height) / 4), 12.0, 0x7e57u + radius);
    std::vector<int32_t> off(41 * 41 * 2);
    std::vector<float> g(41 * 41), E(256);
    bf_gauss_template(std::max(0.5f, radius / 1.5f), 41, off.data(), g.data());
    bf_euclid_template(30.0f, 256, E.data());
    bf_lut lut = nullptr;
    int rc = bf_lut_create(off.data(), g.data(), 41 * 41, E.data(), &lut);
    if (rc) return rc;
    unsigned char *din = nullptr, *dout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<float> cost(nc, -1.0f);
    hipError_t e = hipMalloc(&din, host.size());
    if (e == hipSuccess) e = hipMalloc(&dout, host.size());
    if (e == hipSuccess) e = hipMemcpy(din, host.data(), host.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) rc = hip_fail(e, "bf_tune_radius");
    bf_image_view vin{ din, width, height, width }, vout{ dout, width, height, width };
    // Each configuration is timed as a one-job batch: the launch a batch
    // with that configuration makes (the unrolled kernels where they apply,
    // configurations 0 and 1 at r = 3, 5, 7, 9, 15; the tap-list kernel
    // otherwise), which is what the table's costs steer.
    const bf_job job{ vin, vout, lut, radius, mode };
    for (int c = 0; c < nc && rc == kOk; c++) {
        bf_batch b = nullptr;
        if (bf_batch_create(&job, 1, c, &b) != kOk) continue;
        int r2 = bf_batch_run(b, nullptr);
        if (r2 == kOk) {
            hipEventRecord(e0, nullptr);
            for (int k = 0; k < reps && r2 == kOk; k++) r2 = bf_batch_run(b, nullptr);
            hipEventRecord(e1, nullptr);
        }
        if (hipEventSynchronize(e1) != hipSuccess) {
            bf_batch_destroy(b);
            rc = kHipError;
            break;
        }
        bf_batch_destroy(b);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        if (r2 == kOk) cost[c] = ms / reps / ((float)width * height / 1e6f);
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(din);
    hipFree(dout);
    bf_lut_destroy(lut);
    if (rc) return rc;
    int rc2 = bf_tune_set(radius, mode, cost.data());
    if (rc2) return rc2;
    if (best_config) *best_config = tuned_config(radius, mode);
    if (ms_per_mpix) std::copy(cost.begin(), cost.end(), ms_per_mpix);
    return kOk;
}

int bf_tune_set(int32_t radius, int32_t mode, const float *ms_per_mpix)
{
    if (!ms_per_mpix) return kNullPointer;
    if (radius < 1 || radius > BF_MAX_RADIUS || (mode != 0 && mode != 1)) return kInvalidData;
    TuneTable &t = tune_table();
    std::lock_guard<std::mutex> lock(t.mu);
    bool any = false;
    for (int c = 0; c < tile_config_count(); c++) {
        t.cost[mode][radius][c] = ms_per_mpix[c];
        any = any || ms_per_mpix[c] >= 0;
    }
    t.have[mode][radius] = any;
    return kOk;
}

int bf_tune_get(int32_t radius, int32_t mode, float *ms_per_mpix)
{
    if (radius < 1 || radius > BF_MAX_RADIUS || (mode != 0 && mode != 1)) return kInvalidData;
    const int best = tuned_config(radius, mode);
    if (best >= 0 && ms_per_mpix) {
        TuneTable &t = tune_table();
        std::lock_guard<std::mutex> lock(t.mu);
        for (int c = 0; c < tile_config_count(); c++) ms_per_mpix[c] = t.cost[mode][radius][c];
    }
    return best;
}

int bf_tune_reset(void)
{
    TuneTable &t = tune_table();
    std::lock_guard<std::mutex> lock(t.mu);
    for (auto &m : t.have)
        for (auto &r : m) r = false;
    return kOk;
}

int bf_exact_inplace(const bf_image_view *img, bf_lut lut, int32_t radius, int32_t mode,
                     int32_t repeat, unsigned char *work, void *stream)
{
    if (!img || !lut) {
        set_error("null argument");
        return kNullPointer;
    }
    if (repeat <= 0) {
        set_error("repeat must be >= 1");
        return kInvalidData;
    }
    if (!img->data || img->width <= 0 || img->height <= 0) {
        set_error("empty image view");
        return kInvalidData;
    }
    hipStream_t s = (hipStream_t)stream;
    unsigned char *tmp = work;
    const size_t bytes = (size_t)img->width * img->height;
    if (!tmp) BF_HIP(hipMallocAsync(reinterpret_cast<void **>(&tmp), bytes, s));
    bf_image_view scratch{ tmp, img->width, img->height, img->width };
    int rc = kOk;
    // Each pass: snapshot the image into scratch, filter scratch -> image.
    for (int i = 0; i < repeat && rc == kOk; i++) {
        hipError_t e = hipMemcpy2DAsync(tmp, img->width, img->data, img->pitch, img->width,
                                        img->height, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipMemcpy2DAsync");
            break;
        }
        rc = bf_exact(&scratch, img, lut, radius, mode, stream);
    }
    if (!work) {
        hipError_t e = hipFreeAsync(tmp, s);
        if (rc == kOk && e != hipSuccess) rc = hip_fail(e, "hipFreeAsync");
    }
    return rc;
}

}  // extern "C"

// --------------------------------------------------------------- batches

namespace bfimpl {

struct Batch {
    std::vector<DevJob> jobs;      // sorted, tile_begin assigned
    DevJob *d_jobs = nullptr;
    int32_t *d_counter = nullptr;
    int ntiles = 0;
    int config = 0;
    int halo = 4;
    int max_rows = 0;
    int64_t pixels = 0;
    bool ring = false;             // every job on the unrolled kernels (bf_ring_batch.hip)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = -1;
};

}  // namespace bfimpl

extern "C" {

int bf_batch_create(const bf_job *jobs, int32_t njobs, int32_t config, bf_batch *out)
{
    if (!jobs || !out) {
        set_error("null argument");
        return kNullPointer;
    }
    if (njobs <= 0) {
        set_error("empty batch");
        return kInvalidData;
    }
    int rmax = 0;
    for (int i = 0; i < njobs; i++) rmax = std::max(rmax, (int)jobs[i].radius);
    if (rmax < 1 || rmax > BF_MAX_RADIUS) {
        set_error("radius must be in 1..20");
        return kInvalidData;
    }
    const int halo = halo_class(rmax);
    // Validate and size every job with a provisional config: the one with
    // the least estimated time when every (radius, mode) is tuned, else the
    // built-in default for the largest radius.
    int cfg = config >= 0 ? config : default_config(rmax, jobs[0].mode);
    if (config < 0) {
        TuneTable &t = tune_table();
        std::lock_guard<std::mutex> lock(t.mu);
        bool all = true;
        for (int i = 0; i < njobs && all; i++) {
            const int r = jobs[i].radius, m = jobs[i].mode;
            all = r >= 1 && r <= BF_MAX_RADIUS && (m == 0 || m == 1) && t.have[m][r];
        }
        if (all) {
            double best = -1.0;
            for (int c = 0; c < tile_config_count(); c++) {
                double est = 0.0;
                bool ok = true;
                for (int i = 0; i < njobs && ok; i++) {
                    const float k = t.cost[jobs[i].mode][jobs[i].radius][c];
                    if (k < 0) ok = false;
                    const double px = (double)std::min(jobs[i].in.width, jobs[i].out.width) *
                                      std::min(jobs[i].in.height, jobs[i].out.height);
                    est += px * k;
                }
                if (ok && (best < 0 || est < best)) {
                    best = est;
                    cfg = c;
                }
            }
        }
    }
    if (cfg >= tile_config_count()) {
        set_error("bad tile config");
        return kInvalidData;
    }
    std::vector<DevJob> dj(njobs);
    std::vector<int64_t> cost(njobs);
    int max_rows = 0, max_lut_rows = 0;
    int64_t pixels = 0;
    for (int i = 0; i < njobs; i++) {
        int64_t px = 0;
        if (!jobs[i].lut) {
            set_error("job " + std::to_string(i) + " has no lut");
            return kNullPointer;
        }
        if (views_overlap(&jobs[i].in, &jobs[i].out)) {
            set_error("job " + std::to_string(i) + ": input and output views overlap");
            return kInvalidData;
        }
        int rc = make_job(&jobs[i].in, &jobs[i].out, lut_from_handle(jobs[i].lut), jobs[i].radius,
                          jobs[i].mode, cfg, halo, 0, &dj[i], &px);
        if (rc) {
            set_error("job " + std::to_string(i) + ": " + g_last_error);
            return rc;
        }
        max_rows = std::max(max_rows, dj[i].nrows);
        cost[i] = (int64_t)dj[i].ntaps_padded;
        pixels += px;
    }
    if (config < 0) {
        const int c = fit_config(cfg, rmax, max_rows);
        if (c != cfg) {   // re-plan the jobs with the configuration that fits
            cfg = c;
            for (int i = 0; i < njobs; i++) {
                int rc = make_job(&jobs[i].in, &jobs[i].out, lut_from_handle(jobs[i].lut),
                                  jobs[i].radius, jobs[i].mode, cfg, halo, 0, &dj[i], nullptr);
                if (rc) return rc;
            }
        }
    }
    // Every job on the canonical template at a radius with an unrolled
    // kernel: switch the whole batch to the per-q tables of those kernels.
    bool ring = ring_batch_supported(cfg, halo);
    if (ring) {
        std::vector<DevJob> rj = dj;
        for (int i = 0; i < njobs && ring; i++)
            ring = ring_job(lut_from_handle(jobs[i].lut), jobs[i].radius, jobs[i].mode, cfg, halo,
                            &rj[i]);
        if (ring) dj.swap(rj);
    }
    for (int i = 0; i < njobs; i++) max_lut_rows = std::max(max_lut_rows, dj[i].lut_rows);
    if (lds_bytes(cfg, halo, max_lut_rows) > max_lds_bytes()) {
        set_error("tile config does not fit the weight tables in LDS");
        return kInvalidData;
    }
    // Longest (per-tile cost) first, grouped by table so a workgroup walking
    // the queue rebuilds its LDS table rarely.
    std::vector<int> order(njobs);
    for (int i = 0; i < njobs; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (cost[a] != cost[b]) return cost[a] > cost[b];
        return dj[a].lut_key < dj[b].lut_key;
    });
    auto *B = new (std::nothrow) Batch();
    if (!B) return kOutOfMem;
    B->config = cfg;
    B->halo = halo;
    B->ring = ring;
    B->max_rows = max_lut_rows;
    B->pixels = pixels;
    B->device = current_device();
    const TileConfig &tc = tile_config(cfg);
    int next = 0;
    for (int k = 0; k < njobs; k++) {
        DevJob j = dj[order[k]];
        j.tiles_x = (j.W + tc.tw - 1) / tc.tw;
        j.tile_begin = next;
        next += tiles_of(j, cfg);
        B->jobs.push_back(j);
    }
    B->ntiles = next;
    hipError_t e = hipMalloc(&B->d_jobs, sizeof(DevJob) * B->jobs.size());
    if (e == hipSuccess) e = hipMalloc(&B->d_counter, sizeof(int32_t));
    if (e == hipSuccess)
        e = hipMemcpy(B->d_jobs, B->jobs.data(), sizeof(DevJob) * B->jobs.size(),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(B->d_jobs);
        hipFree(B->d_counter);
        delete B;
        return hip_fail(e, "bf_batch_create");
    }
    *out = reinterpret_cast<bf_batch>(B);
    return kOk;
}

int bf_batch_config(bf_batch h)
{
    if (!h) return kNullPointer;
    return reinterpret_cast<Batch *>(h)->config;
}

int bf_batch_unrolled(bf_batch h)
{
    if (!h) return kNullPointer;
    return reinterpret_cast<Batch *>(h)->ring ? 1 : 0;
}

int bf_batch_destroy(bf_batch h)
{
    if (!h) return kNullPointer;
    auto *B = reinterpret_cast<Batch *>(h);
    if (B->exec) hipGraphExecDestroy(B->exec);
    if (B->graph) hipGraphDestroy(B->graph);
    hipFree(B->d_jobs);
    hipFree(B->d_counter);
    delete B;
    return kOk;
}

int bf_batch_run(bf_batch h, void *stream)
{
    if (!h) return kNullPointer;
    auto *B = reinterpret_cast<Batch *>(h);
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_reset_counter(B->d_counter, s);
    if (rc) return rc;
    if (B->ring)
        return launch_ring_batch(B->config, B->halo, B->d_jobs, (int)B->jobs.size(), B->ntiles,
                                 B->max_rows, B->d_counter, s);
    return launch_exact(B->config, B->halo, B->jobs[0], B->d_jobs, (int)B->jobs.size(), B->ntiles,
                        B->max_rows, B->d_counter, s);
}

int bf_batch_run_graph(bf_batch h, void *stream)
{
    if (!h) return kNullPointer;
    auto *B = reinterpret_cast<Batch *>(h);
    hipStream_t s = (hipStream_t)stream;
    if (!B->exec) {
        // Capture on a private stream so the caller's stream state is untouched.
        hipStream_t cap;
        BF_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int rc = kOk;
        if (e == hipSuccess) rc = bf_batch_run(h, cap);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        hipStreamDestroy(cap);
        if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
        B->graph = g;
        BF_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    }
    BF_HIP(hipGraphLaunch(B->exec, s));
    return kOk;
}

int64_t bf_batch_pixels(bf_batch h)
{
    if (!h) return kNullPointer;
    return reinterpret_cast<Batch *>(h)->pixels;
}

}  // extern "C"
