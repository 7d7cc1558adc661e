// synth.cpp -- deterministic synthetic test images (bf_synth_image).
//
// No reference counterpart: the reference ships no test images.  Content:
//   background 200*x/(W-1);
//   for k in 0..nshapes-1: rectangle k, then disc k (overwriting), sizes
//     uniform in [smin, smax], positions uniform over the image, grey
//     uniform in 0..255;
//   Gaussian noise (Box-Muller in double, counter-based per pixel pair so
//     rows can be generated in parallel), rounded half-to-even, clipped.
// Random numbers: splitmix64.  Shape draws use a sequential stream seeded
// with `seed`; the noise uses sm64(key + n) with key = sm64(seed ^ C).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "opencuda_bf.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr double kTwoPi = 6.283185307179586476925286766559;

inline uint64_t sm64(uint64_t x)
{
    uint64_t z = x + kGolden;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t r = sm64(s);
        s += kGolden;
        return r;
    }
    int uniform(int lo, int hi)   // inclusive
    {
        return lo + (int)(next() % (uint64_t)(hi - lo + 1));
    }
};

}  // namespace

extern "C" int bf_synth_image(unsigned char *dst, int32_t W, int32_t H, int32_t nshapes,
                              int32_t smin, int32_t smax, double noise_sigma, uint64_t seed)
{
    if (!dst) return -2;
    if (W <= 0 || H <= 0 || nshapes < 0 || smin < 1 || smax < smin || noise_sigma < 0) return -1;
    std::vector<float> base((size_t)W * H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            base[(size_t)y * W + x] = W > 1 ? (float)(200.0 * x / (W - 1)) : 0.0f;
    Rng rng{ seed };
    for (int k = 0; k < nshapes; k++) {
        // rectangle k
        {
            const int w = rng.uniform(smin, smax), h = rng.uniform(smin, smax);
            const int cx = rng.uniform(0, W - 1), cy = rng.uniform(0, H - 1);
            const float grey = (float)rng.uniform(0, 255);
            const int x0 = std::max(0, cx - w / 2), x1 = std::min(W, cx - w / 2 + w);
            const int y0 = std::max(0, cy - h / 2), y1 = std::min(H, cy - h / 2 + h);
            for (int y = y0; y < y1; y++)
                for (int x = x0; x < x1; x++) base[(size_t)y * W + x] = grey;
        }
        // disc k
        {
            const int d = rng.uniform(smin, smax);
            const int cx = rng.uniform(0, W - 1), cy = rng.uniform(0, H - 1);
            const float grey = (float)rng.uniform(0, 255);
            const double rad = d / 2.0, r2 = rad * rad;
            const int ri = (int)std::ceil(rad);
            const int y0 = std::max(0, cy - ri), y1 = std::min(H - 1, cy + ri);
            for (int y = y0; y <= y1; y++) {
                const double dy = y - cy;
                const double rem = r2 - dy * dy;
                if (rem < 0) continue;
                const int hw = (int)std::floor(std::sqrt(rem));
                const int x0 = std::max(0, cx - hw), x1 = std::min(W - 1, cx + hw);
                for (int x = x0; x <= x1; x++) base[(size_t)y * W + x] = grey;
            }
        }
    }
    const uint64_t key = sm64(seed ^ 0xD1B54A32D192ED03ull);
    const size_t n = (size_t)W * H;
    auto work = [&](size_t pair_lo, size_t pair_hi) {
        for (size_t p = pair_lo; p < pair_hi; p++) {
            const uint64_t a = sm64(key + 2 * p), b = sm64(key + 2 * p + 1);
            const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;   // (0, 1]
            const double u2 = (double)(b >> 11) * 0x1.0p-53;          // [0, 1)
            const double rr = std::sqrt(-2.0 * std::log(u1));
            const double z[2] = { rr * std::cos(kTwoPi * u2), rr * std::sin(kTwoPi * u2) };
            for (int k = 0; k < 2; k++) {
                const size_t i = 2 * p + k;
                if (i >= n) break;
                double v = std::nearbyint((double)base[i] + noise_sigma * z[k]);
                v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
                dst[i] = (unsigned char)v;
            }
        }
    };
    const size_t pairs = (n + 1) / 2;
    unsigned nth = std::thread::hardware_concurrency();
    if (nth == 0) nth = 1;
    if (nth > 16) nth = 16;
    if (pairs < 65536) nth = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++) {
        const size_t lo = pairs * t / nth, hi = pairs * (t + 1) / nth;
        th.emplace_back(work, lo, hi);
    }
    for (auto &t : th) t.join();
    return 0;
}
