// template_factory.cpp -- the five template vendors (BOX, CIRCLE, ARC,
// GAUSS, EUCLIDE) and the pooled TemplateFactory (okano_3_0/TemplateFactory.cu).
//
// The weight formulas, the ring order of the spatial template, the pool
// hash, the 1e-6 sigma equality and the replacement policy are the
// reference's, because they decide which float bits the filter multiplies:
// a sigma within 1e-6 of a pooled one reuses that template's weights.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "TemplateFactory.h"
#include "opencuda_bf.h"

namespace {

// ---------------------------------------------------------------- vendors

// okano_3_0/TemplateFactory.cu:1057-1076
bool gauss_legal(TFSize size, const void *priv)
{
    if (priv == NULL || size.z != 1) return false;
    if (size.x < 3 || size.y != size.x || size.x % 2 == 0) return false;
    return std::fabs(*static_cast<const float *>(priv)) >= 1.0e-8f;
}

// okano_3_0/TemplateFactory.cu:1079-1089
int gauss_hash(TFSize size, const void *priv)
{
    if (priv == NULL) return -1;
    return ((int)(size.x / 2 - 1) ^ (int)(std::fabs(*static_cast<const float *>(priv)) * 10.0f)) %
           TF_VOL_SHAPE;
}

// Fill `offsets` (2 ints per point) and `weights` for an edge of `size`:
// (0,0) with weight 1, then rings i = 1..size/2, each as j = -i+1..i of
// (i,j), (-i,j), (j,i), (j,-i) sharing exp((float)(-(i*i+j*j)) / s22)
// (okano_3_0/TemplateFactory.cu:1092-1168).
int gauss_fill(float sigma, int size, int32_t *offsets, float *weights)
{
    const float sigma22 = 2 * sigma * sigma;
    int radius = size / 2;
    if (radius < 1) radius = 1;
    int n = 0;
    offsets[0] = 0;
    offsets[1] = 0;
    weights[n++] = 1.0f;
    for (int i = 1; i <= radius; i++) {
        for (int j = -i + 1; j <= i; j++) {
            const float w = std::exp((float)(-(i * i + j * j)) / sigma22);
            const int pts[4][2] = { { i, j }, { -i, j }, { j, i }, { j, -i } };
            for (const auto &p : pts) {
                offsets[2 * n] = p[0];
                offsets[2 * n + 1] = p[1];
                weights[n++] = w;
            }
        }
    }
    return n;
}

// okano_3_0/TemplateFactory.cu:1293-1310
bool euclid_legal(TFSize size, const void *priv)
{
    if (priv == NULL || size.z != 1) return false;
    if (size.x < 1 || (size.y != size.x && size.y != 1)) return false;
    return std::fabs(*static_cast<const float *>(priv)) >= 1.0e-8f;
}

// okano_3_0/TemplateFactory.cu:1313-1320
int euclid_hash(TFSize size, const void *priv)
{
    if (priv == NULL) return -1;
    return ((int)size.x ^ (int)(std::fabs(*static_cast<const float *>(priv)) * 10.0f)) %
           TF_VOL_SHAPE;
}

// okano_3_0/TemplateFactory.cu:1323-1373: point i = (i, 0), weight
// exp((float)(-(i*i)) / s22).
void euclid_fill(float sigma, int size, int32_t *offsets, float *weights)
{
    const float sigma22 = 2 * sigma * sigma;
    for (int i = 0; i < size; i++) {
        if (offsets) {
            offsets[2 * i] = i;
            offsets[2 * i + 1] = 0;
        }
        weights[i] = std::exp((float)(-(i * i)) / sigma22);
    }
}

// Bit interleaving of the reference's shape hashes
// (okano_3_0/TemplateFactory.cu:157-205): bit i of num moves to bit
// (extent + 1) i, keeping the low 32 / (extent + 1) bits.
unsigned stretch_digit(unsigned num, int extent)
{
    const int maxbit = 8 * (int)sizeof(unsigned) / (extent + 1);
    unsigned res = 0u;
    for (int i = 0; i < maxbit; i++) res |= (num & (1u << i)) << (extent * i);
    return res;
}

unsigned combine_digit(TFSize num, int count)
{
    if (count <= 1) return num.x;
    unsigned res = 0u;
    if (count >= 3) {
        count = 3;
        res = stretch_digit(num.z, count) << 2;
    }
    res |= stretch_digit(num.x, count);
    res |= stretch_digit(num.y, count) << 1;
    return res;
}

// ---- BOX (okano_3_0/TemplateFactory.cu:340-402)
bool box_legal(TFSize size, const void *priv)
{
    if (size.z != 1 || priv != NULL) return false;
    return size.x >= 1 && size.y >= 1;
}

int box_hash(TFSize size, const void *) { return (int)(combine_digit(size, 2) % TF_VOL_SHAPE); }

bool size_xy_equal(TFSize a, const void *, TFSize b, const void *) { return a.x == b.x && a.y == b.y; }

// x*y points are allocated; the reference's loops run c over
// [-(x-1)/2, x/2) and r over [-(y-1)/2, y/2) (half-open: (x-1)(y-1) points,
// row-major, as (c, r) pairs) and leave the other entries unwritten.  Here
// those entries are the centre with weight 0.  Every written point has
// weight 1 / (x*y).
int box_fill(int sx, int sy, int32_t *xy, float *w)
{
    const int count = sx * sy;
    const float att = 1.0f / count;
    int n = 0;
    for (int r = -((sy - 1) / 2); r < sy / 2; r++)
        for (int c = -((sx - 1) / 2); c < sx / 2; c++) {
            xy[2 * n] = c;
            xy[2 * n + 1] = r;
            w[n++] = att;
        }
    for (; n < count; n++) {
        xy[2 * n] = 0;
        xy[2 * n + 1] = 0;
        w[n] = 0.0f;
    }
    return count;
}

// ---- CIRCLE (okano_3_0/TemplateFactory.cu:501-666)
bool circle_legal(TFSize size, const void *priv)
{
    if (size.z != 1 || priv != NULL) return false;
    if (size.x < 3 || size.x != size.y) return false;
    return size.x % 2 != 0;
}

int circle_hash(TFSize size, const void *) { return (int)(((size.x - 1) / 2) % TF_VOL_SHAPE); }

bool size_x_equal(TFSize a, const void *, TFSize b, const void *) { return a.x == b.x; }

// Filled digital disc by the midpoint rule on one octant: (0, +-r), row 0,
// then per column step either the four boundary points (+-x, +-y) or, when
// y steps down, rows +-y over [-x, x]; rows +-x are filled over [-y, y]
// while x < y.  Points as (x, y) pairs in that order, weights 1 / count.
void circle_points(int size, std::vector<int32_t> &p)
{
    int radius = (size - 1) / 2;
    if (radius < 1) radius = 1;
    const int r2 = radius * radius;
    p.clear();
    auto put = [&](int a, int b) {
        p.push_back(a);
        p.push_back(b);
    };
    put(0, radius);
    put(0, -radius);
    for (int ix = -radius; ix <= radius; ix++) put(ix, 0);
    int x = 0, y = radius;
    while (x < y) {
        x++;
        int d1 = x * x + y * y - r2, d2 = x * x + (y - 1) * (y - 1) - r2;
        d1 = d1 < 0 ? -d1 : d1;
        d2 = d2 < 0 ? -d2 : d2;
        if (d1 < d2) {
            put(x, y);
            put(-x, y);
            put(x, -y);
            put(-x, -y);
        } else {
            y--;
            if (x > y) break;
            for (int ix = -x; ix <= x; ix++) {
                put(ix, y);
                put(ix, -y);
            }
        }
        if (x < y)
            for (int iy = -y; iy <= y; iy++) {
                put(iy, x);
                put(iy, -x);
            }
    }
}

// ---- ARC (okano_3_0/TemplateFactory.cu:771-936)
bool arc_legal(TFSize size, const void *priv)
{
    if (size.z != 1 || priv != NULL) return false;
    if (size.x < 1 || size.y <= size.x) return false;
    return size.x % 2 != 0 && size.y % 2 != 0;
}

int arc_hash(TFSize size, const void *)
{
    TFSize h(size.x >> 1, size.y >> 1, size.z);
    return (int)(combine_digit(h, 2) % TF_VOL_SHAPE);
}

// The ring between the inner circle of diameter size_x and the outer one of
// diameter size_y: axis points for y in [sr, lr), then per column x the
// points between the inner circle (or the diagonal once it is passed) and
// the outer circle, with their 8 symmetric images (4 on the diagonal).
void arc_points(int size_x, int size_y, std::vector<int32_t> &p)
{
    int sr = (size_x - 1) / 2;
    sr = sr < 1 ? 1 : sr;
    const int sr2 = sr * sr;
    int lr = (size_y - 1) / 2;
    lr = lr <= sr ? sr + 1 : lr;
    const int lr2 = lr * lr;
    p.clear();
    auto put = [&](int a, int b) {
        p.push_back(a);
        p.push_back(b);
    };
    for (int y = sr; y < lr; y++) {
        put(0, y);
        put(0, -y);
        put(y, 0);
        put(-y, 0);
    }
    int x = 0, sy = sr, ly = lr;
    while (x < ly) {
        x++;
        if (x - 1 < sy) {
            const int sd1 = std::abs(x * x + sy * sy - sr2);
            const int sd2 = std::abs(x * x + (sy - 1) * (sy - 1) - sr2);
            sy = sd1 <= sd2 ? sy : sy - 1;
        }
        if (x >= sy) sy = x;
        const int ld1 = std::abs(x * x + ly * ly - lr2);
        const int ld2 = std::abs(x * x + (ly - 1) * (ly - 1) - lr2);
        ly = ld1 <= ld2 ? ly : ly - 1;
        if (x > ly) break;
        for (int y = sy; y < ly; y++) {
            put(x, y);
            put(-x, y);
            put(x, -y);
            put(-x, -y);
            if (x == y) continue;
            put(y, x);
            put(-y, x);
            put(y, -x);
            put(-y, -x);
        }
    }
}

Template *points_template(const std::vector<int32_t> &p);

Template *box_create(TFSize size, const void *);
Template *circle_create(TFSize size, const void *);
Template *arc_create(TFSize size, const void *);

Template *make_host_template(size_t count)
{
    Template *t = NULL;
    if (TemplateBasicOp::newTemplate(&t) != NO_ERROR) return NULL;
    if (TemplateBasicOp::makeAtHost(t, count) != NO_ERROR) {
        TemplateBasicOp::deleteTemplate(t);
        return NULL;
    }
    return t;
}

Template *gauss_create(TFSize size, const void *priv)
{
    const float sigma = *static_cast<const float *>(priv);
    if (std::fabs(sigma) < 1.0e-8f) return NULL;
    int radius = (int)size.x / 2;
    if (radius < 1) radius = 1;
    const int edge = 2 * radius + 1;
    Template *t = make_host_template((size_t)edge * edge);
    if (!t) return NULL;
    gauss_fill(sigma, edge, t->tplData, ATTACHED_DATA(t));
    return t;
}

Template *euclid_create(TFSize size, const void *priv)
{
    const float sigma = *static_cast<const float *>(priv);
    if (size.x < 1 || std::fabs(sigma) < 1.0e-8f) return NULL;
    Template *t = make_host_template(size.x);
    if (!t) return NULL;
    euclid_fill(sigma, (int)size.x, t->tplData, ATTACHED_DATA(t));
    return t;
}

Template *points_template(const std::vector<int32_t> &p)
{
    const size_t n = p.size() / 2;
    if (n == 0) return NULL;
    Template *t = make_host_template(n);
    if (!t) return NULL;
    std::memcpy(t->tplData, p.data(), p.size() * sizeof(int32_t));
    const float att = 1.0f / (float)n;
    float *w = ATTACHED_DATA(t);
    for (size_t i = 0; i < n; i++) w[i] = att;
    return t;
}

Template *box_create(TFSize size, const void *)
{
    Template *t = make_host_template((size_t)size.x * size.y);
    if (!t) return NULL;
    box_fill((int)size.x, (int)size.y, t->tplData, ATTACHED_DATA(t));
    return t;
}

Template *circle_create(TFSize size, const void *)
{
    std::vector<int32_t> p;
    circle_points((int)size.x, p);
    return points_template(p);
}

Template *arc_create(TFSize size, const void *)
{
    std::vector<int32_t> p;
    arc_points((int)size.x, (int)size.y, p);
    return points_template(p);
}

// okano_3_0/TemplateFactory.cu:1171-1188 (gauss) and :1376-1393 (euclide)
bool sigma_equal(TFSize a, const void *pa, TFSize b, const void *pb)
{
    if (pa == NULL || pb == NULL) return false;
    if (std::fabs(*static_cast<const float *>(pa) - *static_cast<const float *>(pb)) >= 1.0e-6f)
        return false;
    return a.x == b.x;
}

// ------------------------------------------------------------------- pool

constexpr int kSlots = TF_VOL_SHAPE * TF_SET_SHAPE;

struct Slot {
    Template *tpl = NULL;
    TFSize size{ 0, 0, 0 };
    float sigma = 0.0f;   // private parameter copy
    bool has_sigma = false;
    int count = 0;        // outstanding getTemplate references
};

struct Pool {
    std::mutex mu;
    Slot slot[TF_CNT_SHAPE][kSlots];
};

Pool &pool()
{
    static Pool *p = new Pool();   // never destroyed: templates may outlive statics
    return *p;
}

// Swap an entry with its predecessor in the set (okano_3_0/TemplateFactory.cu
// :1449-1502) so that frequently used templates are kicked last.
void boost(int shape, int idx)
{
    if (idx % TF_SET_SHAPE == 0) return;
    Slot tmp = pool().slot[shape][idx];
    pool().slot[shape][idx] = pool().slot[shape][idx - 1];
    pool().slot[shape][idx - 1] = tmp;
}

struct Vendor {
    bool (*legal)(TFSize, const void *);
    int (*hash)(TFSize, const void *);
    Template *(*create)(TFSize, const void *);
    bool (*equal)(TFSize, const void *, TFSize, const void *);
    bool has_sigma;
};

// indexed by TF_SHAPE_* (okano_3_0/TemplateFactory.cu:209-216, :331-338,
// :492-499, :762-769, the GAUSS and EUCLIDE vendors further down)
const Vendor kVendors[TF_CNT_SHAPE] = {
    { box_legal, box_hash, box_create, size_xy_equal, false },
    { circle_legal, circle_hash, circle_create, size_x_equal, false },
    { arc_legal, arc_hash, arc_create, size_xy_equal, false },
    { gauss_legal, gauss_hash, gauss_create, sigma_equal, true },
    { euclid_legal, euclid_hash, euclid_create, sigma_equal, true },
};

}  // namespace

// okano_3_0/TemplateFactory.cu:1505-1650
int TemplateFactory::getTemplate(Template **tpl, int shape, TFSize size, void *privated)
{
    if (shape < 0 || shape >= TF_CNT_SHAPE) return INVALID_DATA;
    if (tpl == NULL) return NULL_POINTER;
    const Vendor &v = kVendors[shape];
    if (!v.legal(size, privated)) return INVALID_DATA;
    const int hash = v.hash(size, privated);
    if (hash < 0 || hash > TF_VOL_SHAPE) return INVALID_DATA;
    std::lock_guard<std::mutex> lock(pool().mu);
    Slot *slots = pool().slot[shape];
    const int start = hash * TF_SET_SHAPE, end = start + TF_SET_SHAPE;
    int pos;
    for (pos = start; pos < end; pos++) {
        if (slots[pos].tpl == NULL) break;
        if (v.equal(slots[pos].size, slots[pos].has_sigma ? &slots[pos].sigma : NULL, size,
                    privated)) {
            *tpl = slots[pos].tpl;
            slots[pos].count++;
            boost(shape, pos);
            return NO_ERROR;
        }
    }
    *tpl = v.create(size, privated);
    if (*tpl == NULL) return OUT_OF_MEM;
    if (pos >= end) {
        // Set full: replace the last unused entry in the back half.
        bool replaced = false;
        pos = end - 1;
        const int stop = (start + end) / 2;
        do {
            if (slots[pos].count == 0) {
                replaced = true;
                TemplateBasicOp::deleteTemplate(slots[pos].tpl);
                break;
            }
        } while (pos-- >= stop);
        if (!replaced) pos = -1;
    }
    if (pos >= start && pos < end) {
        slots[pos].tpl = *tpl;
        slots[pos].size = size;
        slots[pos].has_sigma = v.has_sigma;
        slots[pos].sigma = v.has_sigma ? *static_cast<const float *>(privated) : 0.0f;
        slots[pos].count = 1;
    }
    return NO_ERROR;
}

// okano_3_0/TemplateFactory.cu:1653-1675
int TemplateFactory::putTemplate(Template *tpl)
{
    if (tpl == NULL) return NULL_POINTER;
    {
        std::lock_guard<std::mutex> lock(pool().mu);
        for (int shape = 0; shape < TF_CNT_SHAPE; shape++) {
            for (int i = 0; i < kSlots; i++) {
                Slot &s = pool().slot[shape][i];
                if (s.tpl == tpl) {
                    if (s.count > 0) s.count--;
                    return NO_ERROR;
                }
            }
        }
    }
    TemplateBasicOp::deleteTemplate(tpl);
    return NO_ERROR;
}

extern "C" int bf_gauss_template(float sigma, int32_t size, int32_t *offsets, float *weights)
{
    if (!offsets || !weights) return NULL_POINTER;
    TFSize sz((unsigned)size, (unsigned)size);
    if (size < 3 || !gauss_legal(sz, &sigma)) return INVALID_DATA;
    return gauss_fill(sigma, size, offsets, weights);
}

extern "C" int bf_euclid_template(float sigma, int32_t size, float *weights)
{
    if (!weights) return NULL_POINTER;
    TFSize sz((unsigned)size, 1);
    if (size < 1 || !euclid_legal(sz, &sigma)) return INVALID_DATA;
    euclid_fill(sigma, size, NULL, weights);
    return size;
}

// Host-side shape templates for FFI callers and tests: return the point
// count, or an ErrorCode.h status; `points` (x, y pairs) and `weights` may
// be NULL to query the count first.
extern "C" int bf_box_template(int32_t size_x, int32_t size_y, int32_t *points, float *weights)
{
    TFSize sz((unsigned)size_x, (unsigned)size_y);
    if (size_x < 1 || size_y < 1 || !box_legal(sz, NULL)) return INVALID_DATA;
    if ((int64_t)size_x * size_y > (1 << 26)) return INVALID_DATA;
    if (!points || !weights) return size_x * size_y;
    return box_fill(size_x, size_y, points, weights);
}

static int copy_points(const std::vector<int32_t> &p, int32_t *points, float *weights)
{
    const int n = (int)(p.size() / 2);
    if (!points || !weights) return n;
    std::memcpy(points, p.data(), p.size() * sizeof(int32_t));
    const float att = 1.0f / (float)n;
    for (int i = 0; i < n; i++) weights[i] = att;
    return n;
}

extern "C" int bf_circle_template(int32_t size, int32_t *points, float *weights)
{
    TFSize sz((unsigned)size, (unsigned)size);
    if (size < 3 || size > 8191 || !circle_legal(sz, NULL)) return INVALID_DATA;
    std::vector<int32_t> p;
    circle_points(size, p);
    return copy_points(p, points, weights);
}

extern "C" int bf_arc_template(int32_t size_x, int32_t size_y, int32_t *points, float *weights)
{
    TFSize sz((unsigned)size_x, (unsigned)size_y);
    if (size_x < 1 || size_y > 8191 || !arc_legal(sz, NULL)) return INVALID_DATA;
    std::vector<int32_t> p;
    arc_points(size_x, size_y, p);
    return copy_points(p, points, weights);
}
